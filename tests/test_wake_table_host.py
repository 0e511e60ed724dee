"""WakeTable host logic (impedances/wakes.py:167-376) and its interpolation:
the product's host evaluation against the oracle, which calls scipy's
interp1d exactly as the reference does, on the reference's own LHC table
(tests/golden/lhc_7tev_2021_wake.dat.gz: the 7 TeV wake file of the
reference's test_wake_table.py, gzipped) and on synthetic low-beta and
longitudinal tables."""
import os

import numpy as np
import pytest

from oracle import refpath as R

GOLDEN = os.path.join(os.path.dirname(__file__), 'golden')
LHC = os.path.join(GOLDEN, 'lhc_7tev_2021_wake.dat.gz')
COLS = ['time', 'dipole_x', 'dipole_y', 'quadrupole_x', 'quadrupole_y']


def _table():
    from pyheadtail_amd.impedances.wakes import WakeTable
    return WakeTable(LHC, COLS)


def _points(t_ns):
    """-dt samples: every node, midpoints, random points, the range ends."""
    t = 1e-9 * t_ns
    rng = np.random.default_rng(5)
    mid = 0.5 * (t[1:] + t[:-1])
    rnd = rng.uniform(t[0], t[-1], 2000)
    return np.concatenate([t, mid, rnd, [t[0], t[-1]]])


def test_kick_order_and_errors(tmp_path):
    from pyheadtail_amd.impedances.wakes import WakeTable
    from pyheadtail_amd.particles.slicing import UniformBinSlicer
    kicks = _table().get_wake_kicks(UniformBinSlicer(10, n_sigma_z=3))
    assert [type(k).__name__ for k in kicks] == ['DipoleWakeKickX', 'DipoleWakeKickY',
                                                  'QuadrupoleWakeKickX', 'QuadrupoleWakeKickY']
    with pytest.raises(ValueError):
        WakeTable(LHC, COLS[:-1])
    with pytest.raises(ValueError):
        WakeTable(LHC, ['t'] + COLS[1:])
    # every component, the reference's order (quadrupole_xy included)
    data = np.loadtxt(LHC)
    full = np.column_stack([data[:, 0]] + [data[:, 1]] * 10 + [data[:, 2]])
    f = tmp_path / 'all.dat'
    np.savetxt(f, full, fmt='%.17g')
    names = ['time', 'constant_x', 'constant_y', 'dipole_x', 'dipole_y', 'dipole_xy',
             'dipole_yx', 'quadrupole_x', 'quadrupole_y', 'quadrupole_xy', 'quadrupole_yx',
             'longitudinal']
    kinds = [type(k).__name__ for k in
             WakeTable(str(f), names).get_wake_kicks(UniformBinSlicer(10, n_sigma_z=3))]
    assert kinds == ['ConstantWakeKickX', 'ConstantWakeKickY', 'ConstantWakeKickZ',
                     'DipoleWakeKickX', 'DipoleWakeKickY', 'DipoleWakeKickXY', 'DipoleWakeKickYX',
                     'QuadrupoleWakeKickX', 'QuadrupoleWakeKickY', 'QuadrupoleWakeKickXY',
                     'QuadrupoleWakeKickYX']
    # a transverse table starting above 0 is rejected
    bad = full.copy()
    bad[:, 0] += 1.
    fb = tmp_path / 'bad.dat'
    np.savetxt(fb, bad[:, :2], fmt='%.17g')
    with pytest.raises(ValueError):
        WakeTable(str(fb), ['time', 'dipole_x']).function_transverse('dipole_x')


def test_ultrarelativistic_table_equals_interp1d():
    wt = _table()
    data = np.loadtxt(LHC)
    for col in ('dipole_x', 'quadrupole_y'):
        mine = wt.function_transverse(col)
        ref = R.wake_table_transverse(data[:, 0], data[:, COLS.index(col)])
        x = _points(data[:, 0])
        dt = np.concatenate([-x, [1e-9, 0.0, -0.0]])
        np.testing.assert_array_equal(mine(dt), ref(dt))
        with pytest.raises(ValueError):
            mine(np.array([-1.01e-9 * data[-1, 0]]))
        with pytest.raises(ValueError):
            ref(np.array([-1.01e-9 * data[-1, 0]]))


def test_low_beta_and_longitudinal_tables_equal_interp1d():
    from pyheadtail_amd.impedances.wakes import TableWake
    from pyheadtail_amd import _phx
    rng = np.random.default_rng(9)
    t_lb = np.sort(rng.uniform(-2., 40., 300))
    w = rng.normal(size=300)
    mine = TableWake(1e-9 * t_lb, -1e15 * w, _phx.TABLE_TRANS_LOWBETA, 'dipole_x')
    ref = R.wake_table_transverse(t_lb, w)
    dt = -_points(t_lb)
    np.testing.assert_array_equal(mine(dt), ref(dt))
    for t0 in (0.0, -3.0):
        t = np.concatenate([[t0], np.sort(rng.uniform(t0 + 0.1, 50., 299))])
        mine = TableWake(1e-9 * t, 1e12 * w, _phx.TABLE_LONG_BL if t0 == 0 else _phx.TABLE_LONG,
                         'longitudinal')
        ref = R.wake_table_longitudinal(t, w)
        dt = np.concatenate([-_points(t), [5e-8, 1e-7, 0.0, -0.0, -6e-8]])
        np.testing.assert_array_equal(mine(dt), ref(dt))
