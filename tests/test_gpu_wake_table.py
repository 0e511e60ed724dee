"""Wake tables and the transverse damper on the device.

- Wake-table values: a one-hot source moment makes every kick one wake value
  (all other products are exact zeros), so the device interpolation is
  checked bit for bit against the oracle (scipy's interp1d, as the
  reference calls it) on the reference's LHC table and on synthetic
  low-beta and longitudinal tables.
- A WakeField with a wake table against the oracle's WakeField over turns.
- -dt outside a transverse table raises ValueError, as interp1d does.
- TransverseDamper against its formula (feedback/transverse_damper.py).
- The reference's own known answer (tests/simulation_tests/
  test_wake_table.py): LHC at 7 TeV with the 2021 wake table and a damper,
  horizontal growth rate 12.5e-4 per turn within 5 %.
"""
import os

import numpy as np
import pytest
import torch
from scipy.constants import c, e, m_p

from oracle import refpath as R
from oracle import workloads as W
from tests.machines import COORDS, coord_err, to_numpy, to_product

pytestmark = pytest.mark.gpu

LHC = os.path.join(os.path.dirname(__file__), 'golden', 'lhc_7tev_2021_wake.dat.gz')
COLS = ['time', 'dipole_x', 'dipole_y', 'quadrupole_x', 'quadrupole_y']


def _case(name):
    print("\n[case] %s" % name, flush=True)


def _one_hot_values(fn_product, x, age0):
    """Device values wake_factor * W(-x[j]) through the dot-product
    convolution over two stored turns: the one-turn-old entry has a one-hot
    moment at a source of time 0, so its sum is one wake value; the newest
    entry (the targets themselves, moment 0, age age0 chosen so that its
    pairs stay inside the table) adds exact zeros."""
    from pyheadtail_amd import _phx
    from pyheadtail_amd.impedances.wake_kicks import WakeHistory, DipoleWakeKickX
    from pyheadtail_amd.particles.slicing import UniformBinSlicer
    S = len(x)
    h = WakeHistory(S, 2)
    times = torch.zeros((2, S), dtype=torch.float64)
    mom = torch.zeros((2, 3, S), dtype=torch.float64)
    times[1] = torch.from_numpy(-np.asarray(x))   # head: slot 1
    i0 = S // 3
    mom[0, 1, i0] = 1.0                           # one turn old: slot 0, times 0
    h.times.copy_(times)
    h.mom.copy_(mom)
    h.age.copy_(torch.tensor([0.0, age0], dtype=torch.float64))
    h.beta.fill_(1.0)
    rec = _phx.WakeHist()
    rec.n_slices, rec.capacity, rec.n_filled, rec.head = S, 2, 2, 1
    rec.times, rec.mom = h.times.data_ptr(), h.mom.data_ptr()
    rec.age, rec.beta = h.age.data_ptr(), h.beta.data_ptr()
    h.record = _phx.device_record(rec)
    h.n_pushed = 2

    class B(object):
        charge, mass, gamma, beta, particlenumber_per_mp = e, m_p, W.GAMMA, 0.9, 1e5
    k = DipoleWakeKickX(fn_product, UniformBinSlicer(10, n_sigma_z=3), 2)
    r = k.record(B())
    assert r.dot_product == 1
    return h.convolve([r]).cpu().numpy()[0], k._wake_factor(B())


def _check_values(fn_product, fn_oracle, x, age0=0.0):
    """-dt = x (within the table): device against the oracle, bit for bit."""
    x = np.asarray(x, dtype=np.float64)
    got, factor = _one_hot_values(fn_product, x, age0)
    # the turn sum: 0 + (the newest entry's exact zero) + W
    want = factor * ((0.0 + 0.0) + fn_oracle(-x - (0.0 + 0.0)))
    bad = np.nonzero(got != want)[0]
    print("[err] %d of %d wake values differ" % (len(bad), len(x)), flush=True)
    assert len(bad) == 0, (bad[:5], got[bad[:5]], want[bad[:5]])


def test_table_values_bitwise_lhc():
    _case("LHC table values on the device, bit for bit")
    from pyheadtail_amd.impedances.wakes import WakeTable
    data = np.loadtxt(LHC)
    wt = WakeTable(LHC, COLS)
    t = 1e-9 * data[:, 0]
    rng = np.random.default_rng(2)
    x = np.concatenate([t[:400], 0.5 * (t[1:300] + t[:299]), rng.uniform(0, t[-1], 600),
                        rng.uniform(0, 2e-9, 300), [t[-1]]])
    for col in ('dipole_x', 'quadrupole_y'):
        _check_values(wt.function_transverse(col),
                      R.wake_table_transverse(data[:, 0], data[:, COLS.index(col)]), x)


def test_table_values_bitwise_low_beta_and_longitudinal():
    _case("low-beta and longitudinal tables on the device, bit for bit")
    from pyheadtail_amd import _phx
    from pyheadtail_amd.impedances.wakes import TableWake
    rng = np.random.default_rng(4)
    w = rng.normal(size=200)
    t_lb = np.concatenate([[-2.], np.sort(rng.uniform(-2., 40., 198)), [40.]])
    # samples within [-1.9, 18.9] ns; the newest entry's pairs have -dt =
    # x_j - x_i + 19 ns, inside [-2, 40] ns with a margin
    nodes = t_lb[(t_lb >= -1.9) & (t_lb <= 18.9)]
    x = 1e-9 * np.concatenate([nodes, rng.uniform(-1.9, 18.9, 400), [18.9]])
    _check_values(TableWake(1e-9 * t_lb, -1e15 * w, _phx.TABLE_TRANS_LOWBETA, 'd'),
                  R.wake_table_transverse(t_lb, w), x, age0=19e-9)
    for t0 in (0.0, -3.0):
        t = np.concatenate([[t0], np.sort(rng.uniform(t0 + 0.1, 50., 199))])
        x = np.concatenate([1e-9 * t, rng.uniform(-10e-9, 60e-9, 400), [0.0]])
        _check_values(TableWake(1e-9 * t, 1e12 * w,
                                _phx.TABLE_LONG_BL if t0 == 0 else _phx.TABLE_LONG, 'l'),
                      R.wake_table_longitudinal(t, w), x)


@pytest.mark.parametrize("n_turns_wake", [1, 3])
def test_wakefield_with_table_vs_oracle(n_turns_wake):
    _case("WakeField(WakeTable) vs oracle, n_turns_wake=%d" % n_turns_wake)
    from pyheadtail_amd.impedances.wakes import WakeField, WakeTable
    from pyheadtail_amd.particles.slicing import UniformBinSlicer
    from tests.machines import linear_machine
    data = np.loadtxt(LHC)
    ob = W.linear_bunch(20000, seed=31)
    pb = to_product(ob)
    zc = (-0.7, 0.7)
    wf = WakeField(UniformBinSlicer(80, z_cuts=zc),
                   WakeTable(LHC, COLS, n_turns_wake=n_turns_wake))
    names = {'dipole_x': 'DipoleX', 'dipole_y': 'DipoleY', 'quadrupole_x': 'QuadX',
             'quadrupole_y': 'QuadY'}
    okicks = [R.Kick(names[col], R.wake_table_transverse(data[:, 0], data[:, i]), n_turns_wake,
                     uniform_bin=True, z_cuts=zc) for i, col in enumerate(COLS) if i > 0]
    owf = R.WakeFieldOracle(80, okicks, z_cuts=zc)
    mach, omach = linear_machine(), W.linear_map_oracle()
    for turn in range(5):
        for el in mach:
            el.track(pb)
        for el in omach:
            el.track(ob)
        wf.track(pb)
        owf.track(ob)
        err = coord_err(to_numpy(pb), ob)
        print("[err] turn %d coordinates %.3e" % (turn, err), flush=True)
        assert err <= 1e-12, (turn, err)


def test_table_range_error():
    _case("transverse table shorter than the bunch: ValueError")
    import tempfile
    from pyheadtail_amd.impedances.wakes import WakeField, WakeTable
    from pyheadtail_amd.particles.slicing import UniformBinSlicer
    data = np.loadtxt(LHC)[:60]          # up to ~1e-3 ns: far shorter than the bunch
    with tempfile.NamedTemporaryFile('w', suffix='.dat', delete=False) as f:
        np.savetxt(f, data, fmt='%.17g')
    try:
        wf = WakeField(UniformBinSlicer(50, n_sigma_z=3), WakeTable(f.name, COLS))
        pb = to_product(W.linear_bunch(5000, seed=3))
        with pytest.raises(ValueError):
            wf.track(pb)
    finally:
        os.unlink(f.name)


def test_transverse_damper_formula():
    _case("TransverseDamper against its formula")
    from pyheadtail_amd.feedback.transverse_damper import TransverseDamper
    ob = W.linear_bunch(50000, seed=8)
    for args, kw in (((50., 80.), {}), ((40., 0.), {}), ((0., 25.), {}),
                     ((30., 60.), dict(phase=60., local_beta_function=45.))):
        pb = to_product(ob)
        d = TransverseDamper(*args, **kw)
        d.track(pb)
        got = to_numpy(pb)
        want = {k: getattr(ob, k).copy() for k in COORDS}
        ph = d.phase_in_2pi
        mx, my = np.mean(ob.x), np.mean(ob.y)
        mxp, myp = np.mean(ob.xp), np.mean(ob.yp)
        if args[0]:
            want['xp'] = want['xp'] - d.gain_x * np.sin(ph) * mxp
            if kw.get('local_beta_function'):
                want['xp'] = want['xp'] - d.gain_x * np.cos(ph) * mx / kw['local_beta_function']
        if args[1]:
            want['yp'] = want['yp'] - d.gain_y * np.sin(ph) * myp
            if kw.get('local_beta_function'):
                want['yp'] = want['yp'] - d.gain_y * np.cos(ph) * my / kw['local_beta_function']
        for k in COORDS:
            scale = np.sqrt(np.mean(want[k] ** 2))
            assert np.max(np.abs(got[k] - want[k])) <= 1e-14 * scale, (args, k)


def test_wake_table_growth_rate_known_answer():
    """tests/simulation_tests/test_wake_table.py: 12.5e-4 per turn, 5 %, up
    to five attempts (here with seeds 1..5)."""
    from scipy.signal import hilbert
    from scipy.stats import linregress
    from pyheadtail_amd.feedback.transverse_damper import TransverseDamper
    from pyheadtail_amd.impedances.wakes import WakeTable, WakeField
    from pyheadtail_amd.machines.synchrotron import Synchrotron
    from pyheadtail_amd.particles.slicing import UniformBinSlicer
    n_turns, n_mp = 5000, int(1e4)
    intensity, epsn = 2e11, 1.8e-6
    sigma_z = 1.2e-9 / 4.0 * c
    gamma = 7e12 / (m_p * c**2 / e)
    betar = np.sqrt(1 - 1 / gamma**2)
    p0 = m_p * betar * gamma * c
    alpha_mom = 3.483575072011584e-04
    eta = alpha_mom - 1.0 / gamma**2
    Q_s = np.sqrt(e * 12e6 * eta * 35640 / (2 * np.pi * betar * c * p0))
    det_xx, det_xy = 1.4e5 * 15. / 550.0, -1.0e5 * 15. / 550.0
    machine = Synchrotron(optics_mode='smooth', circumference=26658.883199999, n_segments=1,
                          alpha_x=0.0, beta_x=68.9, D_x=0.0, alpha_y=0.0, beta_y=70.34,
                          D_y=0.0, accQ_x=64.31, accQ_y=59.32, Qp_x=-20.0, Qp_y=0.0,
                          app_x=det_xx * p0, app_y=det_xx * p0, app_xy=det_xy * p0,
                          alpha_mom_compaction=alpha_mom, longitudinal_mode='linear',
                          Q_s=Q_s, dphi_RF=0.0, p_increment=0.0, p0=p0, charge=e, mass=m_p,
                          RF_at='end_of_transverse')
    limit_z = 3 * sigma_z
    wake_field = WakeField(UniformBinSlicer(200, z_cuts=(-limit_z, limit_z)),
                           WakeTable(LHC, COLS))
    machine.one_turn_map.append(wake_field)
    machine.one_turn_map.append(TransverseDamper(dampingrate_x=7000, dampingrate_y=7000))
    rates = []
    for attempt in range(1, 6):
        _case("LHC wake table growth rate, attempt %d" % attempt)
        np.random.seed(attempt)
        bunch = machine.generate_6D_Gaussian_bunch_matched(n_mp, intensity=intensity,
                                                           epsn_x=epsn, epsn_y=epsn,
                                                           sigma_z=sigma_z)
        x = np.zeros(n_turns)
        for turn in range(n_turns):
            machine.track(bunch)
            x[turn] = bunch.mean_x()
            if turn % 1000 == 999:
                print("  turn %d" % (turn + 1), flush=True)
        turns = np.arange(n_turns)
        ampl = np.abs(hilbert(x))
        b = linregress(turns[1000:n_turns - 1000], np.log(ampl[1000:n_turns - 1000]))[0]
        rates.append(b)
        print("[err] growth rate %.3e (expected 12.5e-4)" % b, flush=True)
        if np.isclose(b, 12.5e-4, rtol=5e-2):
            return
    raise AssertionError("growth rates %s not within 5%% of 12.5e-4" % rates)
