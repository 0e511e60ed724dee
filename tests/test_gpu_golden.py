"""The HIP path against the golden vectors the reference itself wrote
(tests/golden/make_golden.py): the element-by-element path and the fused
engine, started from the reference's own initial coordinates.

Bars: LinearMap and the transverse segment without detuning are the same
IEEE operations in the same order, so the linear workload's coordinates are
bit-identical; slice counts are bit-identical; the cuts agree to 1e-14 of
the slicing window (the device sums z in a fixed tree order, NumPy pairwise,
and mean z is a sum with cancellation); per-slice means agree to 1e-12 of
the coordinate's rms.  Where sin/cos enter (RF, detuning) the device uses
its own libm, so coordinates agree to TOL = 1e-12 of their rms per turn,
and the bunch statistics to 1e-12 of their scale."""
import os

import numpy as np
import pytest
import torch
from scipy.constants import e, m_p

from oracle import workloads as W

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), 'golden')
COORDS = ('x', 'xp', 'y', 'yp', 'z', 'dp')


def _case(name):
    print("\n[case] %s" % name, flush=True)


def gold(name):
    return np.load(os.path.join(GOLD, name + '.npz'))


def product_bunch(arr, gamma, n_particles=1.5e11):
    from pyheadtail_amd.particles.particles import Particles
    n = arr.shape[1]
    return Particles(n, n_particles / n, e, m_p, W.CIRCUMFERENCE, gamma,
                     {k: arr[i].copy() for i, k in enumerate(COORDS)})


def snap(pb):
    return np.stack([getattr(pb, k).cpu().numpy() for k in COORDS])


TOL = 1e-12


def rel(a, ref):
    """max |a - ref| per coordinate over that coordinate's rms (printed)."""
    rms = np.sqrt(np.mean(ref * ref, axis=1))
    rms[rms == 0] = 1.
    err = float(np.max(np.max(np.abs(a - ref), axis=1) / rms))
    print("[err] coordinates %.3e" % err, flush=True)
    return err


def machine(longitudinal_mode, **kw):
    from pyheadtail_amd.machines.synchrotron import Synchrotron
    args = dict(optics_mode='smooth', charge=e, mass=m_p, p0=W.P0,
                circumference=W.CIRCUMFERENCE, n_segments=1, alpha_x=0., beta_x=W.BETA_X,
                D_x=0., alpha_y=0., beta_y=W.BETA_Y, D_y=0., accQ_x=W.Q_X, accQ_y=W.Q_Y,
                alpha_mom_compaction=W.ALPHA_C, longitudinal_mode=longitudinal_mode)
    args.update(kw)
    return Synchrotron(**args)


def nonlinear_machine():
    p0 = W.P0
    return machine('non-linear', Qp_x=5., Qp_y=5., app_x=1e-10 * p0, app_y=1e-10 * p0,
                   app_xy=-5e-11 * p0, h_RF=W.H_RF, V_RF=W.V_RF, dphi_RF=0., p_increment=0.)


def order(elements):
    return [type(el).__name__ for el in elements]


@pytest.mark.parametrize("fused", [False, True])
def test_linear_workload_vs_reference(fused):
    _case("linear workload vs reference golden, fused=%s" % fused)
    from pyheadtail_amd.particles.slicing import UniformBinSlicer
    g = gold('linear')
    mach = machine('linear', Q_s=0.017)
    assert order(mach.one_turn_map) == list(g['order'])
    pb = product_bunch(g['initial'], g['gamma'][0])
    keep = list(g['coord_turns'])
    sl = UniformBinSlicer(50, n_sigma_z=3)
    eng = None
    if fused:
        from pyheadtail_amd.engine import FusedOneTurnMap
        eng = FusedOneTurnMap(list(mach.one_turn_map), graph_turns=0)
    for turn in range(1, 11):
        if fused:
            eng.track_turns(pb, 1)
        else:
            for el in mach.one_turn_map:
                el.track(pb)
        ss = pb.get_slices(sl, statistics=['mean_x', 'mean_y'])
        zb = ss.z_bins.cpu().numpy()
        cut = np.array([zb[0], zb[-1], zb[1] - zb[0]])
        ref = g['cuts'][turn - 1]
        # mean z is a sum of terms much larger than the result: the device's
        # fixed tree and NumPy's pairwise order round differently (a few ulp)
        assert np.max(np.abs(cut - ref)) <= 1e-14 * (ref[1] - ref[0])
        np.testing.assert_array_equal(ss.n_macroparticles_per_slice.cpu().numpy(),
                                      g['counts'][turn - 1])
        got = snap(pb)
        for i, name in ((0, 'mean_x'), (1, 'mean_y')):
            m = getattr(ss, name).cpu().numpy()
            scale = np.std(got[2 * i])
            assert np.max(np.abs(m - g['means'][turn - 1][i])) <= 1e-12 * scale, name
        if turn in keep:
            np.testing.assert_array_equal(got, g['coords'][keep.index(turn)])


@pytest.mark.parametrize("fused", [False, True])
def test_nonlinear_workload_vs_reference(fused):
    _case("nonlinear + resonator wake vs reference golden, fused=%s" % fused)
    from pyheadtail_amd.particles.slicing import UniformBinSlicer
    from pyheadtail_amd.impedances.wakes import WakeField, CircularResonator
    g = gold('nonlinear')
    mach = nonlinear_machine()
    wake = WakeField(UniformBinSlicer(50, n_sigma_z=3), CircularResonator(1e6, 1e9, 1))
    one_turn = list(mach.one_turn_map) + [wake]
    assert order(one_turn) == list(g['order'])
    pb = product_bunch(g['initial'], g['gamma'][0])
    keep = list(g['coord_turns'])
    if fused:
        from pyheadtail_amd.engine import FusedOneTurnMap
        eng = FusedOneTurnMap(one_turn, graph_turns=0)
    for turn in range(1, 6):
        if fused:
            eng.track_turns(pb, 1)
            counts = eng.slice_data[1].cpu().numpy()
        else:
            for el in one_turn:
                el.track(pb)
            counts = wake.slice_set_deque[0].n_macroparticles_per_slice.cpu().numpy()
        np.testing.assert_array_equal(counts, g['counts'][turn - 1])
        got = snap(pb)
        if turn in keep:
            assert rel(got, g['coords'][keep.index(turn)]) <= TOL
        st = [pb.mean_x(), pb.mean_y(), pb.mean_z(), pb.sigma_z(), pb.epsn_x(), pb.epsn_y(),
              pb.epsn_z()]
        ref = g['stats'][turn - 1]
        scale = np.array([np.std(got[0]), np.std(got[2]), ref[3], ref[3], ref[4], ref[5],
                          ref[6]])
        err = np.max(np.abs(np.array(st) - ref) / scale)
        print("[err] bunch statistics %.3e" % err, flush=True)
        assert err <= TOL, err


@pytest.mark.parametrize("fused", [False, True])
def test_multiturn_wake_vs_reference(fused):
    _case("4-turn resonator + resistive wall (dot product) vs reference, fused=%s" % fused)
    from pyheadtail_amd.particles.slicing import UniformBinSlicer
    from pyheadtail_amd.impedances.wakes import (WakeField, CircularResonator,
                                                 CircularResistiveWall)
    g = gold('multiturn')
    mach = machine('linear', Q_s=0.017)
    wake = WakeField(UniformBinSlicer(20, n_sigma_z=3),
                     CircularResonator(1e6, 1e9, 1, n_turns_wake=4),
                     CircularResistiveWall(0.02, W.CIRCUMFERENCE, 5.9e7, 1e-12,
                                           n_turns_wake=4))
    one_turn = list(mach.one_turn_map) + [wake]
    assert order(one_turn) == list(g['order'])
    pb = product_bunch(g['initial'], g['gamma'][0])
    keep = list(g['coord_turns'])
    if fused:
        from pyheadtail_amd.engine import FusedOneTurnMap
        eng = FusedOneTurnMap(one_turn, graph_turns=0)
    for turn in range(1, 9):
        if fused:
            eng.track_turns(pb, 1)
        else:
            for el in one_turn:
                el.track(pb)
        if turn in keep:
            assert rel(snap(pb), g['coords'][keep.index(turn)]) <= TOL


def test_rf_machine_bunch_generation_vs_reference():
    _case("Synchrotron.generate_6D_Gaussian_bunch vs reference golden")
    g = gold('nonlinear')
    mach = nonlinear_machine()
    np.random.seed(12345)
    pb = mach.generate_6D_Gaussian_bunch(1000, 1.5e11, W.EPSN, W.EPSN, W.SIGMA_Z)
    np.testing.assert_array_equal(snap(pb), g['initial'])
    assert pb.gamma == g['gamma'][0]


@pytest.mark.parametrize("fused", [False, True])
def test_extra_wake_kinds_vs_reference(fused):
    _case("quadrupolar / constant-z kicks, Q = 0.3 and 0.5, parallel-plates RW, fixed cuts "
          "and n_sigma_z, fused=%s" % fused)
    from pyheadtail_amd.particles.slicing import UniformBinSlicer
    from pyheadtail_amd.impedances.wakes import (WakeField, Resonator,
                                                 ParallelHorizontalPlatesResistiveWall)
    g = gold('extra_wakes')

    def sources(T):
        res = [Resonator(1e6, 1.2e9, 0.5, 0.7, 1., 0.3, -0.3, 1, n_turns_wake=T),
               ParallelHorizontalPlatesResistiveWall(0.02, W.CIRCUMFERENCE, 5.9e7, 1e-12,
                                                     n_turns_wake=T)]
        if T == 1:
            res.insert(0, Resonator(2e6, 7e8, 0.3, 1., 0.8, -0.4, 0.4, 1, n_turns_wake=T))
        return res
    mach = machine('linear', Q_s=0.017)
    wakes = [WakeField(UniformBinSlicer(30, z_cuts=(-0.6, 0.6)), *sources(1)),
             WakeField(UniformBinSlicer(30, n_sigma_z=3), *sources(2))]
    assert [type(k).__name__ for w in wakes for k in w.wake_kicks] == list(g['kicks'])
    pb = product_bunch(g['initial'], g['gamma'][0])
    if fused:
        # the engine takes one WakeField at the end of the map: the first one
        # runs as an element, the rest of the turn fused
        from pyheadtail_amd.engine import FusedOneTurnMap
        eng = FusedOneTurnMap([wakes[1]], graph_turns=0)
    for turn in range(3):
        for el in list(mach.one_turn_map) + wakes[:1 if fused else 2]:
            el.track(pb)
        if fused:
            eng.track_turns(pb, 1)
        assert rel(snap(pb), g['coords'][turn]) <= TOL, turn


@pytest.mark.parametrize("fused", [False, True])
def test_dispersion_chromaticity_acceleration_vs_reference(fused):
    _case("two segments with alpha and dispersion, third-order chromaticity, two-harmonic "
          "RF with acceleration, fused=%s" % fused)
    from pyheadtail_amd.trackers.transverse_tracking import TransverseMap
    from pyheadtail_amd.trackers.detuners import Chromaticity, AmplitudeDetuning
    from pyheadtail_amd.trackers.longitudinal_tracking import RFSystems
    g = gold('variants')
    pb = product_bunch(g['initial'], g['gamma'][0])
    C = W.CIRCUMFERENCE
    p0 = pb.p0
    tm = TransverseMap(s=np.array([0., 0.5 * C, C]), alpha_x=np.array([0.4, -0.3, 0.4]),
                       beta_x=np.array([50., 35., 50.]), D_x=np.array([1.2, 0.7, 1.2]),
                       alpha_y=np.array([-0.2, 0.5, -0.2]), beta_y=np.array([60., 45., 60.]),
                       D_y=np.array([0.3, 0.1, 0.3]), accQ_x=W.Q_X, accQ_y=W.Q_Y,
                       detuners=[Chromaticity([5., 120., -3000.], [-2., 50., 800.]),
                                 AmplitudeDetuning(2e-10 * p0, 1e-10 * p0, -5e-11 * p0)])
    rf = RFSystems(C, [4620, 9240], [4.5e6, 1e6], [0., np.pi], [W.ALPHA_C], g['gamma'][0],
                   p_increment=2e-21, D_x=0.7, D_y=0.1, charge=e, mass=m_p)
    one_turn = [tm[0], rf, tm[1]]
    assert order(one_turn) == list(g['order'])
    if fused:
        from pyheadtail_amd.engine import FusedOneTurnMap
        eng = FusedOneTurnMap(one_turn, graph_turns=0)
    for turn in range(5):
        if fused:
            eng.track_turns(pb, 1)
        else:
            for el in one_turn:
                el.track(pb)
        assert rel(snap(pb), g['coords'][turn]) <= TOL, turn
        assert (pb.gamma, pb.p0) == tuple(g['gamma_p0'][turn])
