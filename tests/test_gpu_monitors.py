"""BunchMonitor and SliceMonitor (monitors/monitors.py:39-330) with their
histories on the device: what they store at each step equals what the
bunch's own statistics methods and the SliceSet return at that step, bit
for bit, across buffer wrap-arounds and partial buffer writes; the file has
the reference's groups, dataset names and shapes."""
import os

import numpy as np
import pytest

from oracle import workloads as W
from tests.machines import linear_machine, to_product

pytestmark = pytest.mark.gpu


def _load(path):
    with np.load(path) as f:
        return {k: f[k] for k in f.files}


def test_bunch_monitor_history(tmp_path):
    print("\n[case] BunchMonitor, 23 steps, buffer 8, written every 5", flush=True)
    from pyheadtail_amd.monitors.monitors import BunchMonitor, BUNCH_STATS
    pb = to_product(W.linear_bunch(20000, seed=4))
    mach = linear_machine()
    fn = str(tmp_path / 'bunch')
    n_steps = 23
    mon = BunchMonitor(fn, n_steps, parameters_dict={'Q_x': 20.13}, write_buffer_every=5,
                       buffer_size=8)
    want = {s: np.empty(n_steps) for s in BUNCH_STATS}
    for step in range(n_steps):
        for el in mach:
            el.track(pb)
        mon.dump(pb)
        for s in BUNCH_STATS:
            v = getattr(pb, s)
            want[s][step] = v() if callable(v) else v
    got = _load(mon._store.path)
    assert float(got['attrs/Q_x']) == 20.13
    for s in BUNCH_STATS:
        assert got['Bunch/' + s].shape == (n_steps,)
        np.testing.assert_array_equal(got['Bunch/' + s], want[s], err_msg=s)


def test_bunch_monitor_custom_stats(tmp_path):
    from pyheadtail_amd.monitors.monitors import BunchMonitor
    pb = to_product(W.linear_bunch(5000, seed=5))
    mon = BunchMonitor(str(tmp_path / 'b2'), 3, stats_to_store=['mean_x', 'intensity', 'epsn_y'])
    want = []
    for _ in range(3):
        mon.dump(pb)
        want.append((pb.mean_x(), pb.intensity, pb.epsn_y()))
    got = _load(mon._store.path)
    assert sorted(k for k in got if k.startswith('Bunch/')) == \
        ['Bunch/epsn_y', 'Bunch/intensity', 'Bunch/mean_x']
    np.testing.assert_array_equal(got['Bunch/mean_x'], [w[0] for w in want])
    np.testing.assert_array_equal(got['Bunch/intensity'], [w[1] for w in want])
    np.testing.assert_array_equal(got['Bunch/epsn_y'], [w[2] for w in want])


def test_slice_monitor_history(tmp_path):
    print("\n[case] SliceMonitor, 7 steps, buffer 4, written every 3", flush=True)
    from pyheadtail_amd.monitors.monitors import SliceMonitor, SLICE_STATS, BUNCH_STATS
    from pyheadtail_amd.particles.slicing import UniformBinSlicer
    pb = to_product(W.linear_bunch(20000, seed=6))
    mach = linear_machine()
    slicer = UniformBinSlicer(40, n_sigma_z=3)
    n_steps = 7
    mon = SliceMonitor(str(tmp_path / 'slices'), n_steps, slicer, write_buffer_every=3,
                       buffer_size=4)
    want_s = {s: np.empty((40, n_steps)) for s in SLICE_STATS}
    want_b = {s: np.empty(n_steps) for s in BUNCH_STATS}
    for step in range(n_steps):
        for el in mach:
            el.track(pb)
        mon.dump(pb)
        ss = pb.get_slices(slicer, statistics=True)
        for s in SLICE_STATS:
            want_s[s][:, step] = getattr(ss, s).cpu().numpy()
        for s in BUNCH_STATS:
            v = getattr(pb, s)
            want_b[s][step] = v() if callable(v) else v
    got = _load(mon._store.path)
    for s in SLICE_STATS:
        assert got['Slices/' + s].shape == (40, n_steps)
        np.testing.assert_array_equal(got['Slices/' + s], want_s[s], err_msg=s)
    for s in BUNCH_STATS:
        np.testing.assert_array_equal(got['Bunch/' + s], want_b[s], err_msg=s)


def test_monitors_vs_oracle_12_turns(tmp_path):
    """BunchMonitor and SliceMonitor histories over 12 turns of the linear
    ring against the oracle's statistics of the oracle bunch tracked beside
    them (refpath.bunch_statistics; the per-slice mean, std and emittance of
    stats.pyx:235-425 on refpath's uniform slicing, particles/slicing.py:
    502-518 for the normalised emittances)."""
    print("\n[case] monitors vs oracle: 3e4 MP, linear ring, 12 turns, 40 slices", flush=True)
    from scipy.constants import e as qe
    from oracle import refpath as R
    from pyheadtail_amd.monitors.monitors import (BunchMonitor, SliceMonitor, BUNCH_STATS,
                                                  SLICE_STATS)
    from pyheadtail_amd.particles.slicing import UniformBinSlicer
    ob = W.linear_bunch(30000, seed=8)
    pb = to_product(ob)
    mach, omach = linear_machine(), W.linear_map_oracle()
    n, S = 12, 40
    bm = BunchMonitor(str(tmp_path / 'b'), n, write_buffer_every=5, buffer_size=8)
    sm = SliceMonitor(str(tmp_path / 's'), n, UniformBinSlicer(S, n_sigma_z=3),
                      write_buffer_every=4, buffer_size=6)
    want_b = {s: np.empty(n) for s in BUNCH_STATS}
    want_s = {s: np.empty((S, n)) for s in SLICE_STATS}
    for step in range(n):
        for el in mach:
            el.track(pb)
        for el in omach:
            el.track(ob)
        bm.dump(pb)
        sm.dump(pb)
        st = R.bunch_statistics(ob)
        for s in BUNCH_STATS:
            want_b[s][step] = st[s] if s in st else ob.macroparticlenumber
        sd = R.slice_uniform(ob, S, n_sigma_z=3)
        for k in ('x', 'xp', 'y', 'yp', 'z', 'dp'):
            want_s['mean_' + k][:, step] = R.mean_per_slice(sd, getattr(ob, k))
        for k in ('x', 'y', 'z', 'dp'):
            want_s['sigma_' + k][:, step] = R.std_per_slice(sd, getattr(ob, k))
        want_s['epsn_x'][:, step] = R.emittance_per_slice(sd, ob.x, ob.xp, ob.dp) * ob.betagamma
        want_s['epsn_y'][:, step] = R.emittance_per_slice(sd, ob.y, ob.yp, ob.dp) * ob.betagamma
        want_s['epsn_z'][:, step] = (4 * np.pi * R.emittance_per_slice(sd, ob.z, ob.dp) *
                                     ob.p0 / qe)
        want_s['n_macroparticles_per_slice'][:, step] = sd.n_macroparticles_per_slice
    got_b, got_s = _load(bm._store.path), _load(sm._store.path)
    # scale of each statistic: the spread of its coordinate (means) or itself
    scale = {s: np.max(np.abs(want_b[s])) for s in BUNCH_STATS}
    for k in ('x', 'xp', 'y', 'yp', 'z', 'dp'):
        scale['mean_' + k] = float(np.std(getattr(ob, k)))
    worst = 0.
    for s in BUNCH_STATS:
        err = np.max(np.abs(got_b['Bunch/' + s] - want_b[s])) / scale[s]
        worst = max(worst, err)
        assert err < 1e-11, (s, err)
    np.testing.assert_array_equal(got_s['Slices/n_macroparticles_per_slice'],
                                  want_s['n_macroparticles_per_slice'])
    for s in SLICE_STATS[:-1]:
        err = np.max(np.abs(got_s['Slices/' + s] - want_s[s])) / scale[s]
        worst = max(worst, err)
        assert err < 1e-11, (s, err)
    print("[info] worst |monitor - oracle| / scale = %.3g" % worst, flush=True)
