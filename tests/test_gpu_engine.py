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


def _standalone_grid():
    """The pass grid whose fold order is that of the standalone slicing, so
    that the engine's cuts equal the element path's bit for bit."""
    from pyheadtail_amd import _phx
    return _phx.STANDALONE_GRID


def test_engine_equals_element_path():
    _print_case("engine vs elements, N=1e5, 500 slices, 12 turns")
    ob = W.nonlinear_bunch(100000, seed=11)
    a, b = to_product(ob), to_product(ob)
    m_el = nonlinear_machine()
    for _ in range(12):
        for el in m_el:
            el.track(a)
    eng = _fused(nonlinear_machine(), graph_turns=0, pass_grid=_standalone_grid())
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


def test_engine_coordinates_changed_between_calls():
    """The shifts of the pass's sums come from particle 0 as it is when a
    call starts: the coordinates changed between two calls of one engine
    give what the element path gives from the changed coordinates, also
    when the pass grid is the automatic one (many blocks start while block 0
    has already rewritten particle 0)."""
    _print_case("engine: coordinates changed between calls, 6+6 turns, N=2e5")
    ob = W.nonlinear_bunch(200000, seed=14)
    a, b = to_product(ob), to_product(ob)
    m_el = nonlinear_machine()
    eng = _fused(nonlinear_machine(), graph_turns=3, pass_grid=_standalone_grid())
    for el in m_el * 6:
        el.track(a)
    eng.track_turns(b, 6)
    for pb in (a, b):
        pb.x[0] += 1e-3
        pb.z.mul_(1.01)
    for _ in range(6):
        for el in m_el:
            el.track(a)
    eng.track_turns(b, 6)
    ga, gb = to_numpy(a), to_numpy(b)
    for k in COORDS:
        np.testing.assert_array_equal(ga[k], gb[k], err_msg=k)
    # automatic grid, twice from the same start: identical
    outs = []
    for _ in range(2):
        pb = to_product(ob)
        e = _fused(nonlinear_machine(), graph_turns=4)
        e.track_turns(pb, 5)
        pb.z.mul_(0.99)
        e.track_turns(pb, 7)
        outs.append(to_numpy(pb))
    for k in COORDS:
        np.testing.assert_array_equal(outs[0][k], outs[1][k], err_msg=k)


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
    eng = _fused(nonlinear_machine(n_slices=100, n_turns_wake=8), graph_turns=4,
                 pass_grid=_standalone_grid())
    eng.track_turns(b, 12)
    ga, gb = to_numpy(a), to_numpy(b)
    for k in COORDS:
        np.testing.assert_array_equal(ga[k], gb[k], err_msg=k)


def test_engine_automatic_grid_close_to_standalone_grid():
    """The automatic pass grid changes only the summation order of the
    bunch sums behind the cuts: coordinates agree to rounding."""
    _print_case("engine automatic grid vs standalone grid, 10 turns")
    ob = W.nonlinear_bunch(100000, seed=16)
    outs = []
    for grid in (0, _standalone_grid(), 300):
        pb = to_product(ob)
        eng = _fused(nonlinear_machine(), graph_turns=0, pass_grid=grid)
        eng.track_turns(pb, 10)
        outs.append(to_numpy(pb))
    for o in outs[1:]:
        for k in COORDS:
            ref = outs[0][k]
            assert np.max(np.abs(o[k] - ref)) <= 1e-12 * np.sqrt(np.mean(ref * ref)), k


def _shape_case_machines():
    """One map per compiled element order, with round-chamber (dipole x, y)
    and parallel-plates (four kinds) wakes."""
    from pyheadtail_amd.trackers.longitudinal_tracking import LinearMap, RFSystems
    from pyheadtail_amd.particles.slicing import UniformBinSlicer
    from pyheadtail_amd.impedances.wakes import (WakeField, CircularResonator,
                                                 ParallelHorizontalPlatesResonator)
    from scipy.constants import e, m_p

    def rf(nh, phase=0.):
        h = [W.H_RF, 4 * W.H_RF][:nh]
        v = [W.V_RF, 0.1 * W.V_RF][:nh]
        return RFSystems(W.CIRCUMFERENCE, h, v, [phase, np.pi][:nh], [W.ALPHA_C], W.GAMMA,
                         p_increment=0., charge=e, mass=m_p)

    def tseg(dets='c1_amp'):
        from tests.machines import smooth_transverse
        if dets == 'none':
            return smooth_transverse()[0]
        if dets == 'c1':
            return smooth_transverse(5.)[0]
        return smooth_transverse(5., 1e-10 * W.P0, -5e-11 * W.P0)[0]

    def lin():
        return LinearMap(np.atleast_1d(W.ALPHA_C), W.CIRCUMFERENCE, 0.017)

    def wake(plates):
        src = ParallelHorizontalPlatesResonator(1e6, 1e9, 1) if plates else \
            CircularResonator(1e6, 1e9, 1)
        return WakeField(UniformBinSlicer(64, n_sigma_z=3), src)

    return {
        'RF1_T': lambda: [rf(1), tseg(), wake(False)],
        'RF2_T': lambda: [rf(2), tseg(), wake(True)],
        'T_RF1': lambda: [tseg(), rf(1), wake(True)],
        'T_RF2': lambda: [tseg(), rf(2), wake(False)],
        'LIN_T': lambda: [lin(), tseg(), wake(False)],
        'T_LIN': lambda: [tseg(), lin(), wake(True)],
        'T': lambda: [tseg(), wake(False)],
        'RF1_T_nowake': lambda: [rf(1), tseg()],
        'LIN_T_nodet': lambda: [lin(), tseg('none'), wake(True)],
        'T_RF1_c1': lambda: [tseg('c1'), rf(1), wake(False)],
        'RF1_T_c1_nowake': lambda: [rf(1), tseg('c1')],
        'RF1_T_phase': lambda: [rf(1, 0.3), tseg(), wake(False)],
    }


@pytest.mark.parametrize("case", ['RF1_T', 'LIN_T', 'T_RF1_c1'])
def test_smooth_segments_equal_general_segments(case, monkeypatch):
    """The smooth-segment form (M00 = cos, M01 = J01 sin, ...) against the
    general one (I cos + J sin) on the same smooth machine: bit-identical."""
    _print_case("smooth vs general segment form: %s" % case)
    from pyheadtail_amd import _phx
    from pyheadtail_amd.engine import FusedOneTurnMap
    make = _shape_case_machines()[case]
    ob = W.nonlinear_bunch(20000, seed=22)
    outs, dets = [], []
    for smooth in (True, False):
        if not smooth:
            monkeypatch.setattr(FusedOneTurnMap, '_smooth', staticmethod(lambda t: False))
        pb = to_product(ob)
        eng = _fused(make(), graph_turns=0, pass_grid=_standalone_grid())
        eng.track_turns(pb, 6)
        outs.append(to_numpy(pb))
        dets.append(eng._shape[2])
    assert dets[0] & _phx.DETS_SMOOTH and not dets[1] & _phx.DETS_SMOOTH
    for k in outs[0]:
        np.testing.assert_array_equal(outs[0][k], outs[1][k], err_msg=k)


@pytest.mark.parametrize("case", ['RF1_T', 'RF1_T_c1_nowake'])
def test_stationary_rf_equals_general_rf(case, monkeypatch):
    """The stationary RF form (no beta_ratio, phase offsets, momentum
    increment or dispersion terms) against the general one on the same
    stationary machine: bit-identical."""
    _print_case("stationary vs general RF form: %s" % case)
    from pyheadtail_amd import _phx
    from pyheadtail_amd.engine import FusedOneTurnMap
    make = _shape_case_machines()[case]
    ob = W.nonlinear_bunch(20000, seed=23)
    outs, shapes = [], []
    for stat in (True, False):
        if not stat:
            monkeypatch.setattr(FusedOneTurnMap, '_stationary', staticmethod(lambda p: False))
        pb = to_product(ob)
        eng = _fused(make(), graph_turns=0, pass_grid=_standalone_grid())
        eng.track_turns(pb, 6)
        outs.append(to_numpy(pb))
        shapes.append(eng._shape[0])
    assert shapes[0] == _phx.SHAPE_RF1_T | _phx.SHAPE_STATIONARY
    assert shapes[1] == _phx.SHAPE_RF1_T
    for k in outs[0]:
        np.testing.assert_array_equal(outs[0][k], outs[1][k], err_msg=k)


@pytest.mark.parametrize("case", sorted(_shape_case_machines()))
@pytest.mark.parametrize("rec", [False, True])
def test_specialised_pass_equals_generic(case, rec):
    """The passes compiled for an element order and kick set run the same
    arithmetic as the generic interpreter: bit-identical coordinates and
    statistics."""
    _print_case("specialised vs generic pass: %s, record_stats=%s" % (case, rec))
    from pyheadtail_amd import _phx
    make = _shape_case_machines()[case]
    ob = W.nonlinear_bunch(20000, seed=21)
    outs, stats, shapes = [], [], []
    for spec in (True, False):
        pb = to_product(ob)
        eng = _fused(make(), graph_turns=0, specialise=spec, record_stats=rec,
                     pass_grid=_standalone_grid())
        eng.track_turns(pb, 6)
        outs.append(to_numpy(pb))
        shapes.append(eng._shape)
        if rec:
            stats.append(eng.last_stats['raw'])
    assert shapes[1] == (_phx.SHAPE_GENERIC, _phx.KICKS_GENERIC, _phx.DETS_GENERIC)
    base = (case.replace('_nowake', '').replace('_nodet', '').replace('_c1', '')
            .replace('_phase', ''))
    want_shape = getattr(_phx, 'SHAPE_' + base)
    if base == 'RF1_T' and 'phase' not in case:
        want_shape |= _phx.SHAPE_STATIONARY       # phi = 0, no acceleration, D = 0
    assert shapes[0][0] == want_shape
    want_dets = (_phx.DETS_NONE if 'nodet' in case else
                 _phx.DETS_C1 if '_c1' in case else _phx.DETS_C1_AMP)
    if want_dets != _phx.DETS_NONE:
        want_dets |= _phx.DETS_SMOOTH        # smooth optics: I = 1, J00 = J11 = 0
    assert shapes[0][2] == want_dets
    if case in ('RF2_T', 'T_RF1', 'T_LIN', 'LIN_T_nodet'):
        assert shapes[0][1] == _phx.KICKS_PLATES
    elif 'nowake' not in case:
        assert shapes[0][1] == _phx.KICKS_DXDY
    for k in COORDS:
        np.testing.assert_array_equal(outs[0][k], outs[1][k], err_msg=k)
    if rec:
        np.testing.assert_array_equal(stats[0], stats[1])


@pytest.mark.parametrize("mixed", [False, True])
def test_engine_resistive_wall_memory_equals_elements(mixed):
    """Multi-turn resistive wall (its own convolution kernel), alone and next
    to a multi-turn resonator (the general kernel): the engine equals the
    element path bit for bit and launches only the kernels the kicks need."""
    _print_case("engine vs elements, 8-turn resistive wall, mixed=%s" % mixed)
    from pyheadtail_amd import _phx
    from pyheadtail_amd.particles.slicing import UniformBinSlicer
    from pyheadtail_amd.impedances.wakes import (WakeField, CircularResonator,
                                                 ParallelHorizontalPlatesResistiveWall)

    def make():
        srcs = [ParallelHorizontalPlatesResistiveWall(0.02, W.CIRCUMFERENCE, 5.9e7, 1e-12,
                                                      n_turns_wake=8)]
        if mixed:
            srcs.append(CircularResonator(1e6, 1e9, 1, n_turns_wake=4))
        return nonlinear_machine(wake=False) + [WakeField(UniformBinSlicer(100, n_sigma_z=3),
                                                          *srcs)]
    ob = W.nonlinear_bunch(40000, seed=17)
    a, b = to_product(ob), to_product(ob)
    m_el = make()
    for _ in range(12):
        for el in m_el:
            el.track(a)
    eng = _fused(make(), graph_turns=4, pass_grid=_standalone_grid())
    eng.track_turns(b, 12)
    want = _phx.CONV_RW_DOT | (_phx.CONV_GENERAL if mixed else 0)
    assert eng._conv_kinds == want
    ga, gb = to_numpy(a), to_numpy(b)
    for k in COORDS:
        np.testing.assert_array_equal(ga[k], gb[k], err_msg=k)


def test_engine_reset_between_calls_keeps_last_kick():
    """reset() between calls: the kick of the last turn before the reset is
    applied, then the history restarts; the same as the element path with a
    new WakeField after the first call."""
    _print_case("engine: reset() between calls, 5 + 4 turns, vs elements")
    ob = W.nonlinear_bunch(50000, seed=41)
    a, b = to_product(ob), to_product(ob)
    m1 = nonlinear_machine(n_turns_wake=3)
    for _ in range(5):
        for el in m1:
            el.track(a)
    m2 = nonlinear_machine(n_turns_wake=3)  # a fresh wake history
    for _ in range(4):
        for el in m2:
            el.track(a)
    eng = _fused(nonlinear_machine(n_turns_wake=3), graph_turns=0, pass_grid=_standalone_grid())
    eng.track_turns(b, 5)
    eng.reset()
    eng.track_turns(b, 4)
    ga, gb = to_numpy(a), to_numpy(b)
    for k in COORDS:
        np.testing.assert_array_equal(ga[k], gb[k], err_msg=k)


def test_engine_two_bunches_alternating():
    """Two bunches tracked in turn by one engine (one shared wake history,
    as the reference's WakeField shares its deque): each bunch's last kick
    goes to that bunch, bit for bit as the element path."""
    _print_case("engine: two bunches alternating on one engine, 3 rounds of 2 turns")
    oa, ob_ = W.nonlinear_bunch(30000, seed=42), W.nonlinear_bunch(30000, seed=43)
    ea, eb = to_product(oa), to_product(ob_)
    fa, fb = to_product(oa), to_product(ob_)
    m_el = nonlinear_machine(n_turns_wake=2)
    eng = _fused(nonlinear_machine(n_turns_wake=2), graph_turns=0, pass_grid=_standalone_grid())
    for _ in range(3):
        for bunch_el, bunch_f in ((ea, fa), (eb, fb)):
            for _ in range(2):
                for el in m_el:
                    el.track(bunch_el)
            eng.track_turns(bunch_f, 2)
    for x, y in ((ea, fa), (eb, fb)):
        gx, gy = to_numpy(x), to_numpy(y)
        for k in COORDS:
            np.testing.assert_array_equal(gx[k], gy[k], err_msg=k)


@pytest.mark.parametrize("source", ["resonator", "resistive_wall"])
def test_engine_1000_slices_equals_elements(source):
    """Above 512 slices the merge takes 8 slices per block: at 1000 slices
    the engine equals the element path bit for bit (counts, means and
    coordinates), with a one-turn resonator and an 8-turn wall."""
    _print_case("engine vs elements, 1000 slices, %s" % source)
    from pyheadtail_amd.particles.slicing import UniformBinSlicer
    from pyheadtail_amd.impedances.wakes import (WakeField, CircularResonator,
                                                 ParallelHorizontalPlatesResistiveWall)

    def make():
        src = (CircularResonator(1e6, 1e9, 1) if source == "resonator" else
               ParallelHorizontalPlatesResistiveWall(0.02, W.CIRCUMFERENCE, 5.9e7, 1e-12,
                                                     n_turns_wake=8))
        return nonlinear_machine(wake=False) + [WakeField(UniformBinSlicer(1000, n_sigma_z=3),
                                                          src)]
    ob = W.nonlinear_bunch(200000, seed=44)
    a, b = to_product(ob), to_product(ob)
    m_el = make()
    for _ in range(10):
        for el in m_el:
            el.track(a)
    eng = _fused(make(), graph_turns=5, pass_grid=_standalone_grid())
    eng.track_turns(b, 10)
    ss = m_el[-1].slice_set_deque[0]
    z_bins, counts, mx, my = eng.slice_data
    np.testing.assert_array_equal(counts.cpu().numpy(), ss.n_macroparticles_per_slice.cpu().numpy())
    np.testing.assert_array_equal(mx.cpu().numpy(), ss.mean_x.cpu().numpy())
    np.testing.assert_array_equal(my.cpu().numpy(), ss.mean_y.cpu().numpy())
    ga, gb = to_numpy(a), to_numpy(b)
    for k in COORDS:
        np.testing.assert_array_equal(ga[k], gb[k], err_msg=k)
