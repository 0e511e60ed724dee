"""The fused engine: bit-identical to the element-by-element path on the
device (same kernels' arithmetic, deterministic reductions), independent of
graph capture and reproducible run to run; and within tolerance of the
oracle over many turns."""
import numpy as np
import pytest
import torch

from oracle import refpath as R
from oracle import workloads as W
from tests.machines import (COORDS, coord_err, linear_machine, nonlinear_machine,
                            to_numpy, to_product)

pytestmark = pytest.mark.gpu


def _print_case(name):
    print("\n[case] %s" % name, flush=True)


def _fused(machine, **kw):
    from pyheadtail_amd.engine import FusedOneTurnMap
    return FusedOneTurnMap(machine, **kw)


def test_engine_equals_element_path():
    _print_case("engine vs elements, N=1e5, 500 slices, 12 turns")
    ob = W.nonlinear_bunch(100000, seed=11)
    a, b = to_product(ob), to_product(ob)
    m_el = nonlinear_machine()
    for _ in range(12):
        for el in m_el:
            el.track(a)
    eng = _fused(nonlinear_machine(), graph_turns=0)
    eng.track_turns(b, 12)
    ga, gb = to_numpy(a), to_numpy(b)
    for k in COORDS:
        np.testing.assert_array_equal(ga[k], gb[k], err_msg=k)
    assert a.gamma == b.gamma and a.p0 == b.p0


def test_engine_graph_and_repeatability():
    _print_case("engine graph vs direct, repeat")
    ob = W.nonlinear_bunch(50000, seed=12)
    outs = []
    for graph_turns in (0, 4, 4):
        pb = to_product(ob)
        eng = _fused(nonlinear_machine(), graph_turns=graph_turns)
        eng.track_turns(pb, 10)
        outs.append(to_numpy(pb))
    for k in COORDS:
        np.testing.assert_array_equal(outs[0][k], outs[1][k], err_msg=k)
        np.testing.assert_array_equal(outs[1][k], outs[2][k], err_msg=k)


def test_engine_split_calls_equal_one_call():
    _print_case("engine 4+6 turns == 10 turns")
    ob = W.nonlinear_bunch(30000, seed=13)
    a, b = to_product(ob), to_product(ob)
    e1 = _fused(nonlinear_machine(), graph_turns=0)
    e1.track_turns(a, 10)
    e2 = _fused(nonlinear_machine(), graph_turns=0)
    e2.track_turns(b, 4)
    e2.track_turns(b, 6)
    ga, gb = to_numpy(a), to_numpy(b)
    for k in COORDS:
        np.testing.assert_array_equal(ga[k], gb[k], err_msg=k)


def test_engine_vs_oracle_nonlinear():
    _print_case("engine vs oracle, nonlinear+wake, N=2e4, 30 turns")
    n_turns = 30
    ob = W.nonlinear_bunch(20000, seed=14)
    pb = to_product(ob)
    eng = _fused(nonlinear_machine(), record_stats=True, graph_turns=8)
    eng.track_turns(pb, n_turns)
    m = W.nonlinear_map_oracle()
    hist = [R.bunch_statistics(ob)]
    for _ in range(n_turns):
        for el in m:
            el.track(ob)
        hist.append(R.bunch_statistics(ob))
    err = coord_err(to_numpy(pb), ob)
    print("[err] coordinates after %d turns %.3e" % (n_turns, err), flush=True)
    assert err <= 1e-12, err
    z_bins, counts, mx, my = eng.slice_data
    np.testing.assert_array_equal(counts.cpu().numpy(), m[2].slices[0].n_macroparticles_per_slice)
    st = eng.last_stats
    for key in ('mean_x', 'mean_y', 'mean_z', 'sigma_z', 'epsn_x', 'epsn_y', 'epsn_z'):
        ref = np.array([h[key] for h in hist])
        scale = np.max(np.abs(ref)) if not key.startswith('mean') else \
            np.max([h['sigma_' + key[5:]] for h in hist])
        err = np.max(np.abs(st[key] - ref)) / scale
        print("[err] %s history %.3e" % (key, err), flush=True)
        assert err <= 1e-12, (key, err)


def test_linear_workload_slicing_every_turn():
    _print_case("linear workload N=1e4, 100 turns, slicing every turn")
    from pyheadtail_amd.particles.slicing import UniformBinSlicer
    ob = W.linear_bunch(10000, seed=12345)
    pb = to_product(ob)
    mach = linear_machine()
    omach = W.linear_map_oracle()
    sl = UniformBinSlicer(50, n_sigma_z=3)
    for turn in range(100):
        for el in mach:
            el.track(pb)
        for el in omach:
            el.track(ob)
        ss = pb.get_slices(sl, statistics=['mean_x', 'mean_y'])
        sd = R.slice_uniform(ob, 50, n_sigma_z=3)
        np.testing.assert_array_equal(ss.slice_index_of_particle.cpu().numpy(),
                                      sd.slice_index_of_particle)
        np.testing.assert_array_equal(ss.n_macroparticles_per_slice.cpu().numpy(),
                                      sd.n_macroparticles_per_slice)
        for name, u in (('mean_x', ob.x), ('mean_y', ob.y)):
            ref = R.mean_per_slice(sd, u)
            assert np.max(np.abs(getattr(ss, name).cpu().numpy() - ref)) <= 1e-12 * np.std(u)
    assert coord_err(to_numpy(pb), ob) <= 1e-12


def test_engine_multiturn_wake_equals_elements():
    _print_case("engine vs elements, 8-turn resonator memory (dot product)")
    ob = W.nonlinear_bunch(40000, seed=15)
    a, b = to_product(ob), to_product(ob)
    m_el = nonlinear_machine(n_slices=100, n_turns_wake=8)
    for _ in range(12):
        for el in m_el:
            el.track(a)
    eng = _fused(nonlinear_machine(n_slices=100, n_turns_wake=8), graph_turns=4)
    eng.track_turns(b, 12)
    ga, gb = to_numpy(a), to_numpy(b)
    for k in COORDS:
        np.testing.assert_array_equal(ga[k], gb[k], err_msg=k)

