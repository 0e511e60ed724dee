"""The wake convolution over a synthetic multi-turn history, against the
FP64 formula of the reference's dot-product path (impedances/wake_kicks.py:
75-85, 118-139) evaluated with the oracle's wake functions.

Covers the resistive wall in the dot-product form (a kernel of its own) alone
and next to resonator kicks (the general kernel), slice counts below and
between multiples of the 64-target tile, fewer stored turns than the memory,
source turns of age zero (pairs on both sides of dt = -|dt_min|) and times on
a grid of |dt_min|, so that dt = -|dt_min| (half the wall's value) occurs
exactly."""
import numpy as np
import pytest
import torch
from scipy.constants import c, e, m_p

from oracle import refpath as R
from oracle import workloads as W

pytestmark = pytest.mark.gpu

BETA = np.sqrt(1 - W.GAMMA**-2)
T_REV = W.CIRCUMFERENCE / (BETA * c)
DT_MIN = 2.0 ** -40


class _Bunch(object):
    """What WakeKick._wake_factor reads."""
    charge, mass, gamma, beta = e, m_p, W.GAMMA, BETA
    particlenumber_per_mp = 1.5e5


def _history(S, T, n_filled, seed, grid=False, young=False):
    """A device history with n_filled entries; entry t (t turns old) has its
    own slice times, moments, age and beta.  Returns the ring and the host
    copies, newest first."""
    from pyheadtail_amd import _phx
    from pyheadtail_amd.impedances.wake_kicks import WakeHistory
    rng = np.random.default_rng(seed)
    h = WakeHistory(S, T)
    times, moms, ages, betas = [], [], [], []
    for t in range(n_filled):
        if grid:
            tt = (np.arange(S) - S // 2 + t) * DT_MIN
        else:
            width = 4.6e-12 * (1.0 + 0.05 * rng.standard_normal())
            tt = (np.arange(S) - 0.5 * S) * width + 1e-12 * rng.standard_normal()
        counts = rng.integers(0, 3000, S).astype(np.float64)
        mom = np.stack([counts, counts * rng.normal(0, 1e-3, S), counts * rng.normal(0, 1e-3, S)])
        age = 0.0 if t == 0 or (young and t == 1) else t * T_REV * (1 + 1e-6 * rng.standard_normal())
        times.append(tt)
        moms.append(mom)
        ages.append(age)
        betas.append(BETA * (1 + 1e-9 * rng.standard_normal()))
    head = n_filled - 1
    ht, hm = h.times.cpu(), h.mom.cpu()
    ha, hb = h.age.cpu(), h.beta.cpu()
    for t in range(n_filled):
        slot = (head - t) % T
        ht[slot] = torch.from_numpy(times[t])
        hm[slot] = torch.from_numpy(moms[t])
        ha[slot], hb[slot] = ages[t], betas[t]
    h.times.copy_(ht)
    h.mom.copy_(hm)
    h.age.copy_(ha)
    h.beta.copy_(hb)
    rec = _phx.WakeHist()
    rec.n_slices, rec.capacity, rec.n_filled, rec.head = S, T, n_filled, head
    rec.times, rec.mom = h.times.data_ptr(), h.mom.data_ptr()
    rec.age, rec.beta = h.age.data_ptr(), h.beta.data_ptr()
    h.record = _phx.device_record(rec)
    h.n_pushed = n_filled
    return h, times, moms, ages, betas


def _reference(kick_fn, moment, n_turns_wake, factor, times, moms, ages, betas):
    """wake_factor * sum over stored turns of np.dot(moments, W(dt))."""
    n_t = min(len(times), n_turns_wake)
    acc = 0.
    for t in range(n_t):
        dt = times[0][None, :] - (times[t][:, None] + ages[t])
        acc = acc + np.dot(moms[t][moment], kick_fn(dt, beta=betas[t]))
    return factor * acc


_MOMENT = {'n': 0, 'x': 1, 'y': 2}


def _kicks(T, mixed, dt_min):
    from pyheadtail_amd.particles.slicing import UniformBinSlicer
    from pyheadtail_amd.impedances.wakes import (ParallelHorizontalPlatesResistiveWall,
                                                 CircularResonator)
    sl = UniformBinSlicer(10, n_sigma_z=3)
    srcs = [ParallelHorizontalPlatesResistiveWall(0.02, W.CIRCUMFERENCE, 5.9e7, dt_min,
                                                  n_turns_wake=T)]
    if mixed:
        srcs.insert(0, CircularResonator(1e6, 1e9, 1, n_turns_wake=max(2, T // 2)))
    kicks = [k for s in srcs for k in s.get_wake_kicks(sl)]
    yok = {'DipoleWakeKickX': np.pi**2 / 24., 'QuadrupoleWakeKickX': -np.pi**2 / 24.,
           'DipoleWakeKickY': np.pi**2 / 12., 'QuadrupoleWakeKickY': np.pi**2 / 24.}
    oracle = []
    for s in srcs:
        for k in s.get_wake_kicks(sl):
            name = type(k).__name__
            kind = name.replace('WakeKick', '').replace('Quadrupole', 'Quad')
            mom = _MOMENT[R.KICK_KINDS[kind][0]]
            if type(s).__name__ == 'CircularResonator':
                fn = R.resonator_transverse(1., 1e6, 1e9, 1)
            else:
                fn = R.resistive_wall_transverse(yok[name], 0.02, W.CIRCUMFERENCE, 5.9e7, dt_min)
            oracle.append((fn, mom, s.n_turns_wake))
    return kicks, oracle


@pytest.mark.parametrize("S,T,n_filled,mixed,grid,young", [
    (50, 8, 8, False, False, False),
    (300, 16, 16, False, False, False),
    (300, 16, 5, True, False, False),
    (130, 6, 6, False, True, False),
    (64, 4, 4, True, False, True),
])
def test_multiturn_convolution_vs_formula(S, T, n_filled, mixed, grid, young):
    print("\n[case] convolution S=%d T=%d filled=%d mixed=%s grid=%s young=%s"
          % (S, T, n_filled, mixed, grid, young), flush=True)
    h, times, moms, ages, betas = _history(S, T, n_filled, seed=S + T, grid=grid, young=young)
    kicks, oracle = _kicks(T, mixed, DT_MIN if grid else 1e-12)
    b = _Bunch()
    recs = [k.record(b) for k in kicks]
    assert all(r.dot_product for r in recs)
    got = h.convolve(recs).cpu().numpy()
    for i, (k, (fn, mom, ntw)) in enumerate(zip(kicks, oracle)):
        ref = _reference(fn, mom, ntw, k._wake_factor(b), times, moms, ages, betas)
        scale = np.max(np.abs(ref))
        err = np.max(np.abs(got[i] - ref)) / scale
        print("[err] %s %.3e" % (type(k).__name__, err), flush=True)
        assert scale > 0 and err <= 1e-12, (type(k).__name__, err)


def test_rw_kicks_of_one_source_share_sums():
    """The four kicks of the parallel-plates wall read three moments: the two
    quadrupolar kicks (moment n) must come out in the ratio of their Yokoya
    factors exactly up to rounding of the prefix chain."""
    h, times, moms, ages, betas = _history(200, 8, 8, seed=3)
    kicks, _ = _kicks(8, False, 1e-12)
    names = [type(k).__name__ for k in kicks]
    got = h.convolve([k.record(_Bunch()) for k in kicks]).cpu().numpy()
    qx, qy = got[names.index('QuadrupoleWakeKickX')], got[names.index('QuadrupoleWakeKickY')]
    np.testing.assert_allclose(qx, -qy, rtol=1e-15, atol=0)


@pytest.mark.parametrize("prec,bound", [("fp64", 1e-13), ("fp32", 1e-6), ("mfma_fp32", 1e-6),
                                        ("mfma_bf16x3", 2e-5), ("mfma_bf16x6", 1e-6)])
def test_resistive_wall_kick_error_per_component(prec, bound):
    """Each precision of the multi-turn wall against FP64 pair by pair (the
    reference's sums), 1000 slices x 64 turns of a parallel-plates wall: max
    and RMS error per kick component, relative to the component's max and
    RMS.  FP64 takes the past turns through the separable expansion (bound
    1e-13); the FP32 and matrix-core paths sum every pair (bound 1e-6; the
    two-part bfloat16 split carries about 16 bits, bound 2e-5)."""
    from pyheadtail_amd import _phx
    from pyheadtail_amd.impedances.wakes import WakeField
    print("\n[case] %s vs FP64 pair by pair, resistive wall, S=1000 T=64" % prec, flush=True)
    h, times, moms, ages, betas = _history(1000, 64, 64, seed=64)
    kicks, oracle = _kicks(64, False, 1e-12)
    b = _Bunch()
    ref_recs = [k.record(b) for k in kicks]
    for r in ref_recs:
        r.precision = _phx.PREC_FP64_DIRECT
    recs = [k.record(b) for k in kicks]
    for r in recs:
        r.precision = WakeField.PRECISIONS[prec]
    kref = h.convolve(ref_recs).cpu().numpy()
    got = h.convolve(recs).cpu().numpy()
    # the direct FP64 path itself against the formula on the host
    fn, mom, ntw = oracle[0]
    host = _reference(fn, mom, ntw, kicks[0]._wake_factor(b), times, moms, ages, betas)
    assert np.max(np.abs(kref[0] - host)) <= 1e-12 * np.max(np.abs(host))
    for i, k in enumerate(kicks):
        d = got[i] - kref[i]
        emax = np.max(np.abs(d)) / np.max(np.abs(kref[i]))
        erms = np.sqrt(np.mean(d * d)) / np.sqrt(np.mean(kref[i] ** 2))
        print("[err] %s %s max %.3e rms %.3e" % (type(k).__name__, prec, emax, erms), flush=True)
        assert emax <= bound and erms <= bound, (type(k).__name__, emax, erms)


def test_fp32_resistive_wall_tracking():
    """500 turns of a 64-turn parallel-plates wall, FP32 convolution against
    FP64: centroid and emittance histories within 1e-6 of their scale."""
    print("\n[case] FP32 vs FP64 tracking, 1e5 MP, S=500, T=64, 500 turns", flush=True)
    from pyheadtail_amd.engine import FusedOneTurnMap
    from pyheadtail_amd.particles.slicing import UniformBinSlicer
    from pyheadtail_amd.impedances.wakes import WakeField, ParallelHorizontalPlatesResistiveWall
    from tests.machines import linear_machine, to_product
    ob = W.linear_bunch(100000, seed=5)
    hist = {}
    for prec in ('fp64', 'fp32'):
        wake = WakeField(UniformBinSlicer(500, n_sigma_z=3),
                         ParallelHorizontalPlatesResistiveWall(0.02, W.CIRCUMFERENCE, 5.9e7, 1e-12,
                                                               n_turns_wake=64),
                         precision=prec)
        pb = to_product(ob)
        pb.particlenumber_per_mp = 3e11 / pb.macroparticlenumber
        eng = FusedOneTurnMap(linear_machine() + [wake], record_stats=True, graph_turns=8)
        eng.track_turns(pb, 500)
        eng.flush(pb)
        hist[prec] = eng.last_stats
    for key in ('mean_x', 'mean_y', 'epsn_x', 'epsn_y'):
        a, b = hist['fp32'][key], hist['fp64'][key]
        scale = np.max(np.abs(b)) if key.startswith('epsn') else \
            np.max(hist['fp64']['sigma_' + key[5:]])
        err = np.max(np.abs(a - b)) / scale
        print("[err] %s history FP32 vs FP64 %.3e" % (key, err), flush=True)
        assert err <= 1e-6, (key, err)
