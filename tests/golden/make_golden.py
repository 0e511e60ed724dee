"""Writes the golden vectors tests/golden/*.npz by running the reference
PyHEADTAIL CPU path in this container (never on the GPU box, which has no
copy of the reference).  The fixtures hold inputs and outputs only.

Usage, from the repository root:

    make -C oracle reference             # builds the reference's Cython
                                         # statistics module into oracle/_ref
    python tests/golden/make_golden.py [REFERENCE_ROOT [FIXTURE ...]]

REFERENCE_ROOT defaults to /root/reference (the directory holding the
PyHEADTAIL package).  The script imports the reference package from there
and its compiled statistics module (cobra_functions/stats.pyx) from
oracle/_ref, with bytecode writing off (the reference tree is read-only).

Fixtures (all float64 unless noted; seeds fixed; 'gamma' is the bunch's):
  stats.npz      cov / std / emittance of cobra_functions.stats on seeded
                 arrays (n = 1000, 2, 1), bunch-level statistics (with the
                 statistical Twiss parameters and dispersions) and per-slice
                 statistics of a SliceSet (UniformBinSlicer(10, n_sigma_z=2)),
                 UniformBinSlicer's slicing from z_sample_points and z_cuts
  linear.npz     linear-map ring: ParticleGenerator bunch (N = 1000,
                 sigma_z = 0.23, sigma_dp = 1e-3 as stated), Synchrotron
                 one-turn map with LinearMap (Q_s = 0.017), 10 turns; per
                 turn UniformBinSlicer(50, n_sigma_z=3) counts / mean_x /
                 mean_y / (tail, head, width), coordinates after turns
                 1, 2 and 10
  nonlinear.npz  RF + resonator ring: Synchrotron.generate_6D_Gaussian_bunch
                 (N = 1000, separatrix cut) of the non-linear RF machine
                 with Q' = 5 and amplitude detuning, plus
                 WakeField(UniformBinSlicer(50, n_sigma_z=3),
                 CircularResonator(1e6, 1e9, 1)), 5 turns: bunch statistics
                 and the wake slicing per turn, coordinates after turns 1, 5
  elements.npz   one turn of the RF + resonator ring element by element (after
                 RFSystems, after the transverse segment, after the
                 WakeField), RFBucket Q_s, beta_z, sigma_dp and the
                 separatrix acceptance (margin 0.05) on an 81 x 81 grid
  extra_wakes.npz  the kick types and regimes the workloads leave out
                 (quadrupolar and constant-z kicks; resonators with Q = 0.5
                 and switch_Z, and Q = 0.3 with one turn of memory;
                 parallel-plates resistive wall) on a 500-particle linear
                 machine, one WakeField over fixed cuts (convolution) and one
                 with n_sigma_z and two turns of memory (dot product);
                 coordinates after each of 3 turns
  variants.npz   branches the workloads leave out: two transverse segments
                 with alpha != 0 and dispersion, third-order chromaticity,
                 two-harmonic RFSystems (phase lock, dispersion, p_increment
                 with transverse shrinking); 500 particles, coordinates and
                 (gamma, p0) after each of 5 turns
  growth.npz     3000 turns of the broadband-resonator instability ring of
                 tests/simulation_tests/test_bb_resonator.py (two RF
                 harmonics, 200 slices over +-6 sigma_z) from a seeded,
                 separatrix-cut Gaussian bunch of 1e5 macroparticles:
                 checksums and the first 64 particles of the initial bunch,
                 bunch statistics after every turn, fitted growth rates
  config2.npz    1000 turns of the full plain one-turn map (RF h = 4620,
                 V = 4.5 MV, Q' = 5, amplitude detuning, 500 slices, broadband
                 resonator 1 MOhm/m at 1 GHz) from its separatrix-cut Gaussian
                 bunch of 1e5 macroparticles (seed 12345): checksums of the
                 initial bunch, bunch statistics after every turn, growth
                 rates fitted over turns 500-999
  machines.npz   Synchrotron element order and Q_s for linear (given Q_s
                 or from the RF bucket) and non-linear longitudinal modes,
                 1 / 4 / 5 segments, RF_at 'middle' / 'end_of_transverse'
  multiturn.npz  multi-turn wake memory at small size: linear longitudinal map,
                 WakeField(UniformBinSlicer(20, n_sigma_z=3),
                 CircularResonator(1e6, 1e9, 1, n_turns_wake=4),
                 CircularResistiveWall(0.02, C, 5.9e7, 1e-12, n_turns_wake=4)),
                 8 turns (dot-product path, moving bins): cuts per turn,
                 coordinates after turns 1, 4, 8
"""
import glob
import importlib.util
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))


def load_reference(ref_root):
    sys.dont_write_bytecode = True
    so = glob.glob(os.path.join(ROOT, 'oracle', '_ref', 'stats*.so'))
    if not so:
        raise SystemExit("build the statistics module first: make -C oracle reference")
    spec = importlib.util.spec_from_file_location('PyHEADTAIL.cobra_functions.stats', so[0])
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    sys.modules['PyHEADTAIL.cobra_functions.stats'] = mod
    sys.path.insert(0, ref_root)
    import PyHEADTAIL  # noqa: F401
    return mod


COORDS = ('x', 'xp', 'y', 'yp', 'z', 'dp')
# turns (1-based) whose full coordinates are kept
KEEP_LINEAR = (1, 2, 10)
KEEP_NONLINEAR = (1, 5)
KEEP_MULTITURN = (1, 4, 8)
GAMMA, C, BETA, EPSN = 27.7, 6911.5, 54.6, 2.5e-6
QX, QY, ALPHA_C = 20.13, 20.18, 1.9e-3


def snapshot(bunch):
    return np.stack([np.array(getattr(bunch, k), dtype=np.float64) for k in COORDS])


def machine(longitudinal_mode, **kw):
    from scipy.constants import c, e, m_p
    from PyHEADTAIL.machines.synchrotron import Synchrotron
    p0 = np.sqrt(GAMMA**2 - 1) * m_p * c
    args = dict(optics_mode='smooth', charge=e, mass=m_p, p0=p0, circumference=C,
                n_segments=1, alpha_x=0., beta_x=BETA, D_x=0., alpha_y=0., beta_y=BETA,
                D_y=0., accQ_x=QX, accQ_y=QY, alpha_mom_compaction=ALPHA_C,
                longitudinal_mode=longitudinal_mode)
    args.update(kw)
    return Synchrotron(**args), p0


def gen_stats(cp):
    from scipy.constants import c, e, m_p
    from PyHEADTAIL.particles.generators import ParticleGenerator, gaussian2D
    from PyHEADTAIL.particles.slicing import UniformBinSlicer
    rng = np.random.RandomState(1)
    out = {}
    for n in (1000, 2, 1):
        a, b, d = rng.normal(1., 2., n), rng.normal(-3., 0.5, n), rng.normal(0., 1e-3, n)
        out['a%d' % n], out['b%d' % n], out['d%d' % n] = a, b, d
        out['cov%d' % n] = np.array([cp.cov(a, b), cp.cov(a, a)])
        out['std%d' % n] = np.array([cp.std(a)])
        out['emit%d' % n] = np.array([cp.emittance(a, b, None), cp.emittance(a, b, d)])
    np.random.seed(7)
    beam = ParticleGenerator(macroparticlenumber=400, intensity=1e11, charge=e, mass=m_p,
                             circumference=C, gamma=GAMMA,
                             distribution_x=gaussian2D(1e-8), alpha_x=0., beta_x=BETA,
                             D_x=0., distribution_y=gaussian2D(1e-8), alpha_y=0.,
                             beta_y=BETA, D_y=0.,
                             distribution_z=gaussian2D(1e-3)).generate()
    out['beam'] = snapshot(beam)
    out['gamma'] = np.array([beam.gamma])
    out['beam_stats'] = np.array([beam.mean_x(), beam.mean_y(), beam.mean_z(),
                                  beam.sigma_x(), beam.sigma_y(), beam.sigma_z(),
                                  beam.sigma_dp(), beam.epsn_x(), beam.epsn_y(),
                                  beam.epsn_z()])
    # UniformBinSlicer(z_sample_points=, z_cuts=): n_slices and cuts
    zsp = []
    for n_pts, dz, a, cuts in ((10, 0.1, -0.45, None), (10, 0.1, -0.45, (-1.0, 1.0)),
                               (20, 0.05, -0.5, (-0.2, 0.3)), (7, 0.3, 0.0, (-0.7, 0.05)),
                               (15, 0.02, 0.1, (0.3, 2.0)), (5, 1.0, -2.0, (-2.5, 2.5))):
        pts = a + dz * np.arange(n_pts)
        sl = UniformBinSlicer(0, z_sample_points=pts, z_cuts=cuts)
        zsp.append([n_pts, dz, a, np.nan if cuts is None else cuts[0],
                    np.nan if cuts is None else cuts[1], sl.n_slices, sl.z_cuts[0],
                    sl.z_cuts[1]])
    out['z_sample_points'] = np.array(zsp)
    out['beam_twiss'] = np.array([beam.alpha_Twiss_x(), beam.beta_Twiss_x(),
                                  beam.gamma_Twiss_x(), beam.alpha_Twiss_y(),
                                  beam.beta_Twiss_y(), beam.gamma_Twiss_y(),
                                  beam.dispersion_x(), beam.dispersion_y()])
    ss = beam.get_slices(UniformBinSlicer(10, n_sigma_z=2),
                         statistics=['mean_x', 'mean_y', 'sigma_x', 'sigma_dp',
                                     'epsn_x', 'epsn_z', 'eff_epsn_x'])
    out['ss_z_bins'] = np.array(ss.z_bins)
    out['ss_idx'] = np.array(ss.slice_index_of_particle, dtype=np.int32)
    out['ss_counts'] = np.array(ss.n_macroparticles_per_slice, dtype=np.int32)
    for k in ('mean_x', 'mean_y', 'sigma_x', 'sigma_dp', 'epsn_x', 'epsn_z', 'eff_epsn_x'):
        out['ss_' + k] = np.array(getattr(ss, k))
    return out


def slice_record(ss):
    zb = np.array(ss.z_bins)
    return (np.array(ss.n_macroparticles_per_slice, dtype=np.int32),
            np.array([zb[0], zb[-1], zb[1] - zb[0]]))


def gen_linear():
    from scipy.constants import e, m_p
    from PyHEADTAIL.particles.generators import (ParticleGenerator, gaussian2D,
                                                 gaussian2D_asymmetrical)
    from PyHEADTAIL.particles.slicing import UniformBinSlicer
    mach, p0 = machine('linear', Q_s=0.017)
    np.random.seed(12345)
    eps = EPSN / np.sqrt(GAMMA**2 - 1)
    beam = ParticleGenerator(macroparticlenumber=1000, intensity=1.5e11, charge=e, mass=m_p,
                             circumference=C, gamma=GAMMA,
                             distribution_x=gaussian2D(eps), alpha_x=0., beta_x=BETA, D_x=0.,
                             distribution_y=gaussian2D(eps), alpha_y=0., beta_y=BETA, D_y=0.,
                             distribution_z=gaussian2D_asymmetrical(sigma_u=0.23,
                                                                    sigma_up=1e-3)).generate()
    out = {'initial': snapshot(beam),
           'order': np.array([type(el).__name__ for el in mach.one_turn_map]),
           'gamma': np.array([beam.gamma])}
    slicer = UniformBinSlicer(50, n_sigma_z=3)
    coords, counts, means, cuts = [], [], [], []
    for turn in range(1, 11):
        for el in mach.one_turn_map:
            el.track(beam)
        ss = beam.get_slices(slicer, statistics=['mean_x', 'mean_y'])
        cnt, cut = slice_record(ss)
        if turn in KEEP_LINEAR:
            coords.append(snapshot(beam))
        counts.append(cnt)
        cuts.append(cut)
        means.append(np.stack([np.array(ss.mean_x), np.array(ss.mean_y)]))
    out.update(coords=np.stack(coords), coord_turns=np.array(KEEP_LINEAR),
               counts=np.stack(counts), means=np.stack(means), cuts=np.stack(cuts))
    return out


def gen_nonlinear():
    from PyHEADTAIL.impedances.wakes import WakeField, CircularResonator
    from PyHEADTAIL.particles.slicing import UniformBinSlicer
    from scipy.constants import c, m_p
    p0 = np.sqrt(GAMMA**2 - 1) * m_p * c
    mach, _ = machine('non-linear', Qp_x=5., Qp_y=5., app_x=1e-10 * p0, app_y=1e-10 * p0,
                      app_xy=-5e-11 * p0, h_RF=4620, V_RF=4.5e6, dphi_RF=0., p_increment=0.)
    np.random.seed(12345)
    beam = mach.generate_6D_Gaussian_bunch(1000, 1.5e11, EPSN, EPSN, 0.23)
    wake = WakeField(UniformBinSlicer(50, n_sigma_z=3), CircularResonator(1e6, 1e9, 1))
    one_turn = list(mach.one_turn_map) + [wake]
    out = {'initial': snapshot(beam),
           'order': np.array([type(el).__name__ for el in one_turn]),
           'gamma': np.array([beam.gamma])}
    coords, stats, counts, cuts = [], [], [], []
    for turn in range(1, 6):
        for el in one_turn:
            el.track(beam)
        cnt, cut = slice_record(wake.slice_set_deque[0])
        if turn in KEEP_NONLINEAR:
            coords.append(snapshot(beam))
        counts.append(cnt)
        cuts.append(cut)
        stats.append([beam.mean_x(), beam.mean_y(), beam.mean_z(), beam.sigma_z(),
                      beam.epsn_x(), beam.epsn_y(), beam.epsn_z()])
    out.update(coords=np.stack(coords), coord_turns=np.array(KEEP_NONLINEAR),
               stats=np.array(stats), counts=np.stack(counts), cuts=np.stack(cuts))
    return out


def gen_elements():
    """Element by element through one turn of the RF + resonator ring, and the
    quantities that fix its bunch: RFBucket Q_s, sigma_dp, the separatrix
    acceptance on a grid."""
    from scipy.constants import c, m_p
    from PyHEADTAIL.impedances.wakes import WakeField, CircularResonator
    from PyHEADTAIL.particles.slicing import UniformBinSlicer
    p0 = np.sqrt(GAMMA**2 - 1) * m_p * c
    mach, _ = machine('non-linear', Qp_x=5., Qp_y=5., app_x=1e-10 * p0, app_y=1e-10 * p0,
                      app_xy=-5e-11 * p0, h_RF=4620, V_RF=4.5e6, dphi_RF=0., p_increment=0.)
    bucket = mach.longitudinal_map.get_bucket(gamma=mach.gamma, mass=mach.mass,
                                              charge=mach.charge)
    eta = mach.longitudinal_map.alpha_array[0] - mach.gamma ** -2
    beta_z = np.abs(eta) * mach.circumference / 2.0 / np.pi / bucket.Q_s
    zz, dd = np.meshgrid(np.linspace(-0.8, 0.8, 81), np.linspace(-4e-3, 4e-3, 81))
    acc = bucket.make_is_accepted(margin=0.05)(zz.ravel(), dd.ravel())
    out = {'Q_s': np.array([bucket.Q_s]), 'beta_z': np.array([beta_z]),
           'sigma_dp': np.array([0.23 / beta_z]), 'grid_z': zz.ravel(),
           'grid_dp': dd.ravel(), 'accepted': np.array(acc, dtype=np.int8)}
    np.random.seed(12345)
    beam = mach.generate_6D_Gaussian_bunch(1000, 1.5e11, EPSN, EPSN, 0.23)
    wake = WakeField(UniformBinSlicer(50, n_sigma_z=3), CircularResonator(1e6, 1e9, 1))
    out['initial'] = snapshot(beam)
    out['gamma'] = np.array([beam.gamma])
    for name, el in (('after_rf', mach.one_turn_map[0]), ('after_transverse',
                                                          mach.one_turn_map[1]),
                     ('after_wake', wake)):
        el.track(beam)
        out[name] = snapshot(beam)
    out['p0_after'] = np.array([beam.p0, beam.gamma])
    return out


def gen_multiturn():
    from PyHEADTAIL.impedances.wakes import (WakeField, CircularResonator,
                                             CircularResistiveWall)
    from PyHEADTAIL.particles.generators import (ParticleGenerator, gaussian2D,
                                                 gaussian2D_asymmetrical)
    from scipy.constants import e, m_p
    from PyHEADTAIL.particles.slicing import UniformBinSlicer
    mach, p0 = machine('linear', Q_s=0.017)
    np.random.seed(2024)
    eps = EPSN / np.sqrt(GAMMA**2 - 1)
    beam = ParticleGenerator(macroparticlenumber=1000, intensity=1.5e11, charge=e, mass=m_p,
                             circumference=C, gamma=GAMMA,
                             distribution_x=gaussian2D(eps), alpha_x=0., beta_x=BETA, D_x=0.,
                             distribution_y=gaussian2D(eps), alpha_y=0., beta_y=BETA, D_y=0.,
                             distribution_z=gaussian2D_asymmetrical(sigma_u=0.23,
                                                                    sigma_up=1e-3)).generate()
    wake = WakeField(UniformBinSlicer(20, n_sigma_z=3),
                     CircularResonator(1e6, 1e9, 1, n_turns_wake=4),
                     CircularResistiveWall(0.02, C, 5.9e7, 1e-12, n_turns_wake=4))
    one_turn = list(mach.one_turn_map) + [wake]
    out = {'initial': snapshot(beam),
           'order': np.array([type(el).__name__ for el in one_turn]),
           'gamma': np.array([beam.gamma])}
    coords, cuts = [], []
    for turn in range(1, 9):
        for el in one_turn:
            el.track(beam)
        if turn in KEEP_MULTITURN:
            coords.append(snapshot(beam))
        cuts.append(slice_record(wake.slice_set_deque[0])[1])
    out.update(coords=np.stack(coords), coord_turns=np.array(KEEP_MULTITURN),
               cuts=np.stack(cuts))
    return out


def extra_wakes(sources):
    """Per kick type and resonator regime: linear machine, 30 slices over
    fixed cuts (np.convolve path) then n_sigma_z (dot product with T = 2)."""
    from PyHEADTAIL.impedances.wakes import WakeField
    from PyHEADTAIL.particles.generators import (ParticleGenerator, gaussian2D,
                                                 gaussian2D_asymmetrical)
    from PyHEADTAIL.particles.slicing import UniformBinSlicer
    from scipy.constants import e, m_p
    mach, _ = machine('linear', Q_s=0.017)
    np.random.seed(77)
    eps = EPSN / np.sqrt(GAMMA**2 - 1)
    beam = ParticleGenerator(macroparticlenumber=500, intensity=1.5e11, charge=e, mass=m_p,
                             circumference=C, gamma=GAMMA,
                             distribution_x=gaussian2D(eps), alpha_x=0., beta_x=BETA, D_x=0.,
                             distribution_y=gaussian2D(eps), alpha_y=0., beta_y=BETA, D_y=0.,
                             distribution_z=gaussian2D_asymmetrical(sigma_u=0.23,
                                                                    sigma_up=1e-3)).generate()
    out = {'initial': snapshot(beam), 'gamma': np.array([beam.gamma])}
    wakes = [WakeField(UniformBinSlicer(30, z_cuts=(-0.6, 0.6)), *sources(1)),
             WakeField(UniformBinSlicer(30, n_sigma_z=3), *sources(2))]
    coords = []
    for turn in range(3):
        for el in list(mach.one_turn_map) + wakes:
            el.track(beam)
        coords.append(snapshot(beam))
    out['coords'] = np.stack(coords)
    out['kicks'] = np.array([type(k).__name__ for w in wakes for k in w.wake_kicks])
    return out


def gen_extra_wakes():
    """The kick types and wake regimes the workloads leave out: quadrupolar
    and constant-z kicks, resonator Q < 1/2, Q = 1/2, longitudinal resonator,
    parallel-plates resistive wall."""
    from PyHEADTAIL.impedances.wakes import (Resonator, ParallelHorizontalPlatesResistiveWall)

    # Q < 1/2 only with one turn of memory: for past turns the reference's
    # exp(alpha dt) * sinh(omegabar dt) is 0 * inf = NaN
    def sources(T):
        res = [Resonator(1e6, 1.2e9, 0.5, 0.7, 1., 0.3, -0.3, 1, n_turns_wake=T),
               ParallelHorizontalPlatesResistiveWall(0.02, C, 5.9e7, 1e-12, n_turns_wake=T)]
        if T == 1:
            res.insert(0, Resonator(2e6, 7e8, 0.3, 1., 0.8, -0.4, 0.4, 1, n_turns_wake=T))
        return res
    return extra_wakes(sources)


def gen_variants():
    """Branches the workloads leave out: two transverse segments with
    alpha != 0 and dispersion, third-order chromaticity, a two-harmonic
    RFSystems with phase lock, dispersion and acceleration (p_increment, with
    the transverse shrinking), 5 turns."""
    from scipy.constants import e, m_p
    from PyHEADTAIL.particles.generators import (ParticleGenerator, gaussian2D,
                                                 gaussian2D_asymmetrical)
    from PyHEADTAIL.trackers.transverse_tracking import TransverseMap
    from PyHEADTAIL.trackers.detuners import Chromaticity, AmplitudeDetuning
    from PyHEADTAIL.trackers.longitudinal_tracking import RFSystems
    np.random.seed(99)
    eps = EPSN / np.sqrt(GAMMA**2 - 1)
    beam = ParticleGenerator(macroparticlenumber=500, intensity=1.5e11, charge=e, mass=m_p,
                             circumference=C, gamma=GAMMA,
                             distribution_x=gaussian2D(eps), alpha_x=0.4, beta_x=50., D_x=1.2,
                             distribution_y=gaussian2D(eps), alpha_y=-0.2, beta_y=60., D_y=0.3,
                             distribution_z=gaussian2D_asymmetrical(sigma_u=0.2,
                                                                    sigma_up=8e-4)).generate()
    p0 = beam.p0
    s = np.array([0., 0.5 * C, C])
    tm = TransverseMap(s=s, alpha_x=np.array([0.4, -0.3, 0.4]),
                       beta_x=np.array([50., 35., 50.]), D_x=np.array([1.2, 0.7, 1.2]),
                       alpha_y=np.array([-0.2, 0.5, -0.2]), beta_y=np.array([60., 45., 60.]),
                       D_y=np.array([0.3, 0.1, 0.3]), accQ_x=QX, accQ_y=QY,
                       detuners=[Chromaticity([5., 120., -3000.], [-2., 50., 800.]),
                                 AmplitudeDetuning(2e-10 * p0, 1e-10 * p0, -5e-11 * p0)])
    rf = RFSystems(C, [4620, 9240], [4.5e6, 1e6], [0., np.pi], [ALPHA_C], GAMMA,
                   p_increment=2e-21, D_x=0.7, D_y=0.1, charge=e, mass=m_p)
    one_turn = [tm[0], rf, tm[1]]
    out = {'initial': snapshot(beam), 'gamma': np.array([beam.gamma]),
           'order': np.array([type(el).__name__ for el in one_turn])}
    coords, gam = [], []
    for turn in range(5):
        for el in one_turn:
            el.track(beam)
        coords.append(snapshot(beam))
        gam.append([beam.gamma, beam.p0])
    out.update(coords=np.stack(coords), gamma_p0=np.array(gam))
    return out


# the broadband-resonator instability of tests/simulation_tests/
# test_bb_resonator.py:13-67 (expected growth 59e-4 / 25e-4 per turn there,
# with a matched bunch)
GROWTH = dict(n=100000, turns=3000, fit_from=1500, seed=42, intensity=2.5e11, epsn=2e-6,
              sigma_z=0.23, circumference=6911.5038378975451, p0_eV=26e9,
              beta_x=54.644808743169399, beta_y=54.509415262636274, alpha=0.0030864197530864196,
              R_shunt=10.2e6, frequency=1.3e9, Q=1., n_slices=200, n_sigma_z=6.)


def growth_rate(signal, fit_from):
    """Slope of log|analytic signal| over turns fit_from.. (the fit of
    test_bb_resonator.py:97-104)."""
    from scipy.signal import hilbert
    from scipy.stats import linregress
    turns = np.arange(len(signal))
    return linregress(turns[fit_from:], np.log(np.abs(hilbert(signal))[fit_from:])).slope


def gen_growth():
    """3000 turns of the broadband-resonator ring (two RF harmonics, 200
    slices over +-6 sigma_z) from a separatrix-cut Gaussian bunch of 1e5
    macroparticles: bunch statistics after every turn and the fitted
    horizontal and vertical growth rates."""
    from scipy.constants import c, e, m_p
    from PyHEADTAIL.machines.synchrotron import Synchrotron
    from PyHEADTAIL.particles.slicing import UniformBinSlicer
    from PyHEADTAIL.impedances.wakes import WakeField, CircularResonator
    G = GROWTH
    mach = Synchrotron(optics_mode='smooth', circumference=G['circumference'], n_segments=1,
                       beta_x=G['beta_x'], beta_y=G['beta_y'], D_x=0., D_y=0., accQ_x=QX,
                       accQ_y=QY, Qp_x=0., Qp_y=0., alpha_mom_compaction=G['alpha'],
                       longitudinal_mode='non-linear', h_RF=[4620, 4 * 4620],
                       V_RF=[4.5e6, 0.45e6], dphi_RF=[0., np.pi], p_increment=0.,
                       p0=G['p0_eV'] * e / c, charge=e, mass=m_p)
    wake = WakeField(UniformBinSlicer(G['n_slices'], n_sigma_z=G['n_sigma_z']),
                     CircularResonator(G['R_shunt'], G['frequency'], G['Q']))
    one_turn = list(mach.one_turn_map) + [wake]
    np.random.seed(G['seed'])
    beam = mach.generate_6D_Gaussian_bunch(G['n'], G['intensity'], G['epsn'], G['epsn'],
                                           G['sigma_z'])
    init = snapshot(beam)
    out = dict(order=np.array([type(el).__name__ for el in one_turn]),
               gamma=np.array([beam.gamma]), initial_sum=init.sum(axis=1),
               initial_sumsq=(init * init).sum(axis=1), initial_head=init[:, :64],
               params=np.array(sorted(G.items()), dtype=object).astype(str))
    names = ('mean_x', 'mean_y', 'mean_z', 'sigma_z', 'epsn_x', 'epsn_y', 'epsn_z')
    hist = []
    for turn in range(G['turns']):
        for el in one_turn:
            el.track(beam)
        hist.append([getattr(beam, k)() for k in names])
    hist = np.array(hist)
    out.update(stat_names=np.array(names), history=hist, final_head=snapshot(beam)[:, :64],
               growth=np.array([growth_rate(hist[:, 0], G['fit_from']),
                                growth_rate(hist[:, 1], G['fit_from'])]))
    print("growth rates x, y: %.3e %.3e per turn" % tuple(out['growth']))
    return out


# the full plain one-turn map (RF h = 4620, V = 4.5 MV; Q' = 5 and amplitude
# detuning; 500 slices over +-3 sigma_z; CircularResonator 1 MOhm/m, 1 GHz,
# Q = 1) at 1e5 macroparticles for 1000 turns
CONFIG2 = dict(n=100000, turns=1000, fit_from=500, seed=12345, intensity=1.5e11)


def gen_config2():
    """1000 turns of the full plain one-turn map from its separatrix-cut
    Gaussian bunch of 1e5 macroparticles: bunch statistics after every turn
    and the growth rates fitted as test_bb_resonator.py:97-104 fits them."""
    from PyHEADTAIL.impedances.wakes import WakeField, CircularResonator
    from PyHEADTAIL.particles.slicing import UniformBinSlicer
    from scipy.constants import c, m_p
    G = CONFIG2
    p0 = np.sqrt(GAMMA**2 - 1) * m_p * c
    mach, _ = machine('non-linear', Qp_x=5., Qp_y=5., app_x=1e-10 * p0, app_y=1e-10 * p0,
                      app_xy=-5e-11 * p0, h_RF=4620, V_RF=4.5e6, dphi_RF=0., p_increment=0.)
    np.random.seed(G['seed'])
    beam = mach.generate_6D_Gaussian_bunch(G['n'], G['intensity'], EPSN, EPSN, 0.23)
    wake = WakeField(UniformBinSlicer(500, n_sigma_z=3), CircularResonator(1e6, 1e9, 1))
    one_turn = list(mach.one_turn_map) + [wake]
    init = snapshot(beam)
    out = dict(order=np.array([type(el).__name__ for el in one_turn]),
               gamma=np.array([beam.gamma]), initial_sum=init.sum(axis=1),
               initial_sumsq=(init * init).sum(axis=1), initial_head=init[:, :64])
    names = ('mean_x', 'mean_y', 'mean_z', 'sigma_z', 'epsn_x', 'epsn_y', 'epsn_z')
    hist = []
    for turn in range(G['turns']):
        for el in one_turn:
            el.track(beam)
        hist.append([getattr(beam, k)() for k in names])
    hist = np.array(hist)
    out.update(stat_names=np.array(names), history=hist, final_head=snapshot(beam)[:, :64],
               growth=np.array([growth_rate(hist[:, 0], G['fit_from']),
                                growth_rate(hist[:, 1], G['fit_from'])]))
    print("growth rates x, y: %.3e %.3e per turn" % tuple(out['growth']))
    return out


MACHINE_CASES = [(mode, n_seg, at) for mode in ('linear Q_s', 'non-linear', 'linear bucket')
                 for n_seg in (1, 4, 5) for at in ('middle', 'end_of_transverse')]


def gen_machines():
    """Synchrotron one-turn-map element order and Q_s for each longitudinal
    mode, segment count and RF_at placement."""
    from scipy.constants import c, e, m_p
    from PyHEADTAIL.machines.synchrotron import Synchrotron
    rf = dict(h_RF=[4620, 4 * 4620], V_RF=[4.5e6, 0.45e6], dphi_RF=[0., np.pi], p_increment=0.)
    orders, qs = [], []
    for mode, n_seg, at in MACHINE_CASES:
        kw = {'linear Q_s': dict(Q_s=0.017), 'non-linear': rf,
              'linear bucket': dict(h_RF=4620, V_RF=4.5e6, dphi_RF=0., p_increment=0.)}[mode]
        m = Synchrotron('smooth', charge=e, mass=m_p, p0=26e9 * e / c, circumference=6911.5,
                        n_segments=n_seg, beta_x=54., beta_y=54., D_x=0., D_y=0.,
                        accQ_x=20.13, accQ_y=20.18, alpha_mom_compaction=3e-3,
                        longitudinal_mode=mode.split()[0], RF_at=at, **kw)
        orders.append(','.join(type(el).__name__ for el in m.one_turn_map))
        qs.append(m.Q_s)
    return dict(cases=np.array(['%s|%d|%s' % cse for cse in MACHINE_CASES]),
                orders=np.array(orders), Q_s=np.array(qs))


GENERATORS = ('stats', 'linear', 'nonlinear', 'multiturn', 'elements', 'extra_wakes',
              'variants', 'growth', 'machines', 'config2')


def main():
    """make_golden.py [reference root] [fixture names...] (default: all)."""
    ref = sys.argv[1] if len(sys.argv) > 1 else '/root/reference'
    only = sys.argv[2:] or GENERATORS
    cp = load_reference(ref)
    fns = dict(stats=lambda: gen_stats(cp), linear=gen_linear, nonlinear=gen_nonlinear,
               multiturn=gen_multiturn, elements=gen_elements, extra_wakes=gen_extra_wakes,
               variants=gen_variants, growth=gen_growth, machines=gen_machines,
               config2=gen_config2)
    for name in only:
        out = fns[name]()
        path = os.path.join(HERE, name + '.npz')
        np.savez_compressed(path, **out)
        print("wrote %s (%d bytes)" % (path, os.path.getsize(path)))


if __name__ == '__main__':
    main()
