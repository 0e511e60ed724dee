"""The oracle against the golden vectors written by the reference itself
(tests/golden/make_golden.py runs PyHEADTAIL's CPU path in the build
container).  These pin the oracle: the statistics functions, the seeded bunch
generators (draw order, matching, separatrix cut), the slicing, and whole
turns of the three workload shapes, including the multi-turn dot-product wake
with moving bins."""
import os

import numpy as np
import pytest
from scipy.constants import c, e, m_p

from oracle import refpath as R
from oracle import workloads as W

GOLD = os.path.join(os.path.dirname(__file__), 'golden')
COORDS = ('x', 'xp', 'y', 'yp', 'z', 'dp')


def gold(name):
    return np.load(os.path.join(GOLD, name + '.npz'))


def bunch_from(arr, gamma, n_particles=1.5e11):
    n = arr.shape[1]
    return R.Bunch(n, n_particles / n, e, m_p, W.CIRCUMFERENCE, gamma,
                   {k: arr[i] for i, k in enumerate(COORDS)})


def snapshot(b):
    return np.stack([getattr(b, k) for k in COORDS])


def rel_err(a, ref):
    scale = np.sqrt(np.mean(ref * ref, axis=-1, keepdims=True))
    scale[scale == 0] = 1.
    return float(np.max(np.abs(a - ref) / scale))


def test_statistics_functions_bit_exact():
    g = gold('stats')
    for n in (1000, 2, 1):
        a, b, d = g['a%d' % n], g['b%d' % n], g['d%d' % n]
        # NaN (a negative determinant from rounding) must be NaN on both sides
        np.testing.assert_array_equal([R.cov(a, b), R.cov(a, a)], g['cov%d' % n])
        np.testing.assert_array_equal([R.std(a)], g['std%d' % n])
        np.testing.assert_array_equal([R.emittance(a, b, None), R.emittance(a, b, d)],
                                      g['emit%d' % n])


def test_bunch_and_slice_statistics():
    g = gold('stats')
    b = bunch_from(g['beam'], g['gamma'][0], 1e11)
    st = R.bunch_statistics(b)
    keys = ('mean_x', 'mean_y', 'mean_z', 'sigma_x', 'sigma_y', 'sigma_z', 'sigma_dp',
            'epsn_x', 'epsn_y', 'epsn_z')
    for k, ref in zip(keys, g['beam_stats']):
        assert st[k] == ref, k
    sd = R.slice_uniform(b, 10, n_sigma_z=2)
    np.testing.assert_array_equal(sd.z_bins, g['ss_z_bins'])
    np.testing.assert_array_equal(sd.slice_index_of_particle, g['ss_idx'])
    np.testing.assert_array_equal(sd.n_macroparticles_per_slice, g['ss_counts'])
    np.testing.assert_allclose(R.mean_per_slice(sd, b.x), g['ss_mean_x'], rtol=1e-13,
                               atol=1e-13 * np.std(b.x))
    np.testing.assert_allclose(R.mean_per_slice(sd, b.y), g['ss_mean_y'], rtol=1e-13,
                               atol=1e-13 * np.std(b.y))
    np.testing.assert_allclose(R.std_per_slice(sd, b.x), g['ss_sigma_x'], rtol=1e-12)
    np.testing.assert_allclose(R.std_per_slice(sd, b.dp), g['ss_sigma_dp'], rtol=1e-12)
    bg = b.betagamma
    np.testing.assert_allclose(R.emittance_per_slice(sd, b.x, b.xp, b.dp) * bg,
                               g['ss_epsn_x'], rtol=1e-9)
    np.testing.assert_allclose(R.emittance_per_slice(sd, b.x, b.xp) * bg,
                               g['ss_eff_epsn_x'], rtol=1e-9)
    epsz = R.emittance_per_slice(sd, b.z, b.dp) * b.p0 / e * 4 * np.pi
    np.testing.assert_allclose(epsz, g['ss_epsn_z'], rtol=1e-9)


def test_generators_reproduce_reference_draws():
    """Draw order, matching and (non-linear RF) the separatrix cut."""
    np.testing.assert_array_equal(snapshot(W.linear_bunch(1000, seed=12345)),
                                  gold('linear')['initial'])
    np.testing.assert_array_equal(snapshot(W.nonlinear_bunch(1000, seed=12345)),
                                  gold('nonlinear')['initial'])


def test_linear_workload_turns():
    g = gold('linear')
    assert list(g['order']) == ['LinearMap', 'TransverseSegmentMap']
    b = bunch_from(g['initial'], g['gamma'][0])
    m = W.linear_map_oracle()
    keep = list(g['coord_turns'])
    for turn in range(1, 11):
        for el in m:
            el.track(b)
        sd = R.slice_uniform(b, 50, n_sigma_z=3)
        np.testing.assert_array_equal(sd.n_macroparticles_per_slice, g['counts'][turn - 1])
        np.testing.assert_array_equal([sd.z_bins[0], sd.z_bins[-1],
                                       sd.z_bins[1] - sd.z_bins[0]], g['cuts'][turn - 1])
        np.testing.assert_array_equal(R.mean_per_slice(sd, b.x), g['means'][turn - 1][0])
        np.testing.assert_array_equal(R.mean_per_slice(sd, b.y), g['means'][turn - 1][1])
        if turn in keep:
            np.testing.assert_array_equal(snapshot(b), g['coords'][keep.index(turn)])


def test_nonlinear_workload_turns():
    g = gold('nonlinear')
    assert list(g['order']) == ['RFSystems', 'TransverseSegmentMap', 'WakeField']
    b = bunch_from(g['initial'], g['gamma'][0])
    m = W.nonlinear_map_oracle(n_slices=50)
    keep = list(g['coord_turns'])
    for turn in range(1, 6):
        for el in m:
            el.track(b)
        sd = m[2].slices[0]
        np.testing.assert_array_equal(sd.n_macroparticles_per_slice, g['counts'][turn - 1])
        np.testing.assert_array_equal([sd.z_bins[0], sd.z_bins[-1],
                                       sd.z_bins[1] - sd.z_bins[0]], g['cuts'][turn - 1])
        st = R.bunch_statistics(b)
        got = [st[k] for k in ('mean_x', 'mean_y', 'mean_z', 'sigma_z', 'epsn_x', 'epsn_y',
                               'epsn_z')]
        np.testing.assert_array_equal(got, g['stats'][turn - 1])
        if turn in keep:
            np.testing.assert_array_equal(snapshot(b), g['coords'][keep.index(turn)])


def multiturn_oracle(T=4, n_slices=20):
    lin, seg = W.linear_map_oracle()
    wr = R.resonator_transverse(1., 1e6, 1e9, 1.)
    ww = R.resistive_wall_transverse(1., 0.02, W.CIRCUMFERENCE, 5.9e7, 1e-12)
    kicks = [R.Kick('DipoleX', wr, T), R.Kick('DipoleY', wr, T),
             R.Kick('DipoleX', ww, T), R.Kick('DipoleY', ww, T)]
    return [lin, seg, R.WakeFieldOracle(n_slices, kicks, n_sigma_z=3)]


def test_multiturn_dot_product_wake():
    g = gold('multiturn')
    assert list(g['order']) == ['LinearMap', 'TransverseSegmentMap', 'WakeField']
    b = bunch_from(g['initial'], g['gamma'][0])
    m = multiturn_oracle()
    keep = list(g['coord_turns'])
    for turn in range(1, 9):
        for el in m:
            el.track(b)
        sd = m[2].slices[0]
        np.testing.assert_array_equal([sd.z_bins[0], sd.z_bins[-1],
                                       sd.z_bins[1] - sd.z_bins[0]], g['cuts'][turn - 1])
        if turn in keep:
            np.testing.assert_array_equal(snapshot(b), g['coords'][keep.index(turn)])


def test_rf_machine_elements_and_bucket():
    """One turn of the RF + resonator machine element by element, the bucket's
    Q_s and the separatrix acceptance (margin 0.05) on a grid."""
    g = gold('elements')
    assert W.q_s_single_harmonic(W.GAMMA, W.ALPHA_C, W.H_RF, W.V_RF) == g['Q_s'][0]
    assert W.nonlinear_sigma_dp() == g['sigma_dp'][0]
    acc = W.stationary_acceptance(W.GAMMA, W.ALPHA_C, W.H_RF, W.V_RF, W.CIRCUMFERENCE, 0.05)
    np.testing.assert_array_equal(acc(g['grid_z'], g['grid_dp']), g['accepted'].astype(bool))
    b = bunch_from(g['initial'], g['gamma'][0])
    m = W.nonlinear_map_oracle(n_slices=50)
    for name, el in zip(('after_rf', 'after_transverse', 'after_wake'), m):
        el.track(b)
        np.testing.assert_array_equal(snapshot(b), g[name], err_msg=name)
    assert (b.p0, b.gamma) == tuple(g['p0_after'])


def extra_wakes_oracle():
    """The two WakeFields of the extra_wakes fixture (see make_golden.py)."""
    pp = (np.pi**2 / 24., np.pi**2 / 12., -np.pi**2 / 24., np.pi**2 / 24.)

    def resonator(T, z_cuts, Rs, f, Q, yx1, yy1, yx2, yy2):
        return [R.Kick('DipoleX', R.resonator_transverse(yx1, Rs, f, Q), T, z_cuts=z_cuts),
                R.Kick('QuadX', R.resonator_transverse(yx2, Rs, f, Q), T, z_cuts=z_cuts),
                R.Kick('DipoleY', R.resonator_transverse(yy1, Rs, f, Q), T, z_cuts=z_cuts),
                R.Kick('QuadY', R.resonator_transverse(yy2, Rs, f, Q), T, z_cuts=z_cuts),
                R.Kick('ConstantZ', R.resonator_longitudinal(Rs, f, Q), T, z_cuts=z_cuts)]

    def sources(T, z_cuts):
        kicks = []
        if T == 1:
            kicks += resonator(T, z_cuts, 2e6, 7e8, 0.3, 1., 0.8, -0.4, 0.4)
        kicks += resonator(T, z_cuts, 1e6, 1.2e9, 0.5, 0.7, 1., 0.3, -0.3)
        for kind, y in (('DipoleX', pp[0]), ('QuadX', pp[2]), ('DipoleY', pp[1]),
                        ('QuadY', pp[3])):
            wake = R.resistive_wall_transverse(y, 0.02, W.CIRCUMFERENCE, 5.9e7, 1e-12)
            kicks.append(R.Kick(kind, wake, T, z_cuts=z_cuts))
        return kicks
    return [R.WakeFieldOracle(30, sources(1, (-0.6, 0.6)), z_cuts=(-0.6, 0.6)),
            R.WakeFieldOracle(30, sources(2, None), n_sigma_z=3)]


def test_extra_wake_kinds():
    """Quadrupolar and constant-z kicks, resonators with Q = 0.3 and 0.5 and
    switch_Z, parallel-plates resistive wall, fixed cuts and n_sigma_z."""
    g = gold('extra_wakes')
    b = bunch_from(g['initial'], g['gamma'][0])
    lin, seg = W.linear_map_oracle()
    wakes = extra_wakes_oracle()
    assert [k.kind for w in wakes for k in w.kicks] == [
        n.replace('WakeKick', '').replace('Quadrupole', 'Quad') for n in g['kicks']]
    for turn in range(3):
        for el in [lin, seg] + wakes:
            el.track(b)
        np.testing.assert_array_equal(snapshot(b), g['coords'][turn])


def test_dispersion_chromaticity_acceleration_variants():
    """Two segments with alpha != 0 and dispersion (the last segment closes
    on the optics of the first boundary), third-order chromaticity and
    amplitude detuning scaled by the segment's phase advance, a
    two-harmonic RFSystems with phase lock, dispersion and p_increment."""
    g = gold('variants')
    assert list(g['order']) == ['TransverseSegmentMap', 'RFSystems', 'TransverseSegmentMap']
    b = bunch_from(g['initial'], g['gamma'][0])
    C = W.CIRCUMFERENCE
    s = np.array([0., 0.5 * C, C])
    ax, bx, Dx = [0.4, -0.3, 0.4], [50., 35., 50.], [1.2, 0.7, 1.2]
    ay, by, Dy = [-0.2, 0.5, -0.2], [60., 45., 60.], [0.3, 0.1, 0.3]
    Qp_x, Qp_y = [5., 120., -3000.], [-2., 50., 800.]
    p0 = b.p0
    app = (2e-10 * p0, 1e-10 * p0, -5e-11 * p0)
    seg_len = np.diff(s) / s[-1]
    dQ_x, dQ_y = W.Q_X * seg_len, W.Q_Y * seg_len
    dmu_x, dmu_y = dQ_x / W.Q_X, dQ_y / W.Q_Y
    segs = []
    for k in range(2):
        s0, s1 = k, (k + 1) % 2
        dets = [R.Chroma([q * dmu_x[k] for q in Qp_x], [q * dmu_y[k] for q in Qp_y]),
                R.Amplitude(app[0] * dmu_x[k], app[1] * dmu_y[k], app[2] * dmu_x[k],
                            app[2] * dmu_y[k], ax[s0], bx[s0], ay[s0], by[s0])]
        segs.append(R.TransverseSegment(ax[s0], bx[s0], Dx[s0], ax[s1], bx[s1], Dx[s1],
                                        ay[s0], by[s0], Dy[s0], ay[s1], by[s1], Dy[s1],
                                        dQ_x[k], dQ_y[k], dets))
    rf = R.RFSystemsOracle(C, [4620, 9240], [4.5e6, 1e6], [0., np.pi], [W.ALPHA_C],
                           g['gamma'][0], e, p_increment=2e-21, D_x=0.7, D_y=0.1)
    for turn in range(5):
        for el in (segs[0], rf, segs[1]):
            el.track(b)
        np.testing.assert_array_equal(snapshot(b), g['coords'][turn])
        assert (b.gamma, b.p0) == tuple(g['gamma_p0'][turn])
