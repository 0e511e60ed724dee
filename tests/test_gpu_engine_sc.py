"""Space-charge stages of the fused engine: TransverseGaussianSpaceCharge
after each transverse segment (spacecharge/spacecharge.py:202-246, installed
as Synchrotron.install_after_each_transverse_segment does), run by the
engine's stage passes with the space-charge slicing on the device.

Against the oracle's ring (refpath's segments, space-charge kicks, linear map
and resonator wake) over several turns; against the element-by-element path,
whose kicks use the same field arithmetic (the two differ only in the last
bits of the slice standard deviations: single-pass exact sums here, the
two-pass form there); calls split into chunks and graph replays give the
bits of one call; statistics records as the bunch's own statistics."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

COORDS = ('x', 'xp', 'y', 'yp', 'z', 'dp')


def _case(msg):
    print("\n[case] " + msg, flush=True)


def _host(b):
    return {k: getattr(b, k).cpu().numpy() for k in COORDS}


def _rel_err(a, b):
    """max over coordinates of max|a-b| / rms(b)."""
    worst = 0.
    for k in COORDS:
        rms = np.sqrt(np.mean(b[k] * b[k]))
        if rms > 0:
            worst = max(worst, float(np.max(np.abs(a[k] - b[k])) / rms))
    return worst


@pytest.mark.parametrize("n_seg,wake", [(4, True), (2, False), (1, True)])
def test_space_charge_stages_vs_oracle(n_seg, wake):
    _case("fused space charge vs oracle: %d segments, %s, 2e4 MP, 6 turns"
          % (n_seg, "resonator wake" if wake else "no wake"))
    from tests.machines import sc_ring, oracle_sc_ring, oracle_bunch_from
    from pyheadtail_amd.engine import FusedOneTurnMap
    m, b = sc_ring(n_seg, 3., 64, 20000, seed=11, wake_slices=100 if wake else 0)
    ob = oracle_bunch_from(b)
    eng = FusedOneTurnMap(m, graph_turns=4)
    assert len(eng.sc) == n_seg
    eng.track_turns(b, 6)
    for _ in range(6):
        for el in oracle_sc_ring(n_seg, 3., 64, 100 if wake else 0):
            el.track(ob)
    err = _rel_err(_host(b), {k: getattr(ob, k) for k in COORDS})
    print("[info] max |fused - oracle| / rms = %.3g" % err, flush=True)
    assert err < 1e-11


def test_space_charge_stages_vs_elements():
    _case("fused space charge vs element by element: 8 segments, 3e5 MP, 500 slices, 5 turns")
    import torch
    from tests.machines import sc_ring, clone_bunch
    from pyheadtail_amd import _phx
    from pyheadtail_amd.engine import FusedOneTurnMap
    from pyheadtail_amd.serve import ElementLoop
    m, b = sc_ring(8, 5., 500, 300000, seed=12)
    b2 = clone_bunch(b)
    # the element path's cuts come from the standalone slicing grid
    FusedOneTurnMap(m, graph_turns=0, pass_grid=_phx.STANDALONE_GRID).track_turns(b, 5)
    ElementLoop(m).track_turns(b2, 5)
    torch.cuda.synchronize()
    err = _rel_err(_host(b), _host(b2))
    print("[info] max |fused - elements| / rms = %.3g" % err, flush=True)
    assert err < 1e-12
    # z and dp never see the space-charge statistics: identical bits
    for k in ('z', 'dp'):
        np.testing.assert_array_equal(getattr(b, k).cpu().numpy(), getattr(b2, k).cpu().numpy())


def test_space_charge_chunks_and_graphs_equal_one_call():
    _case("fused space charge: 7 + 5 turns (graphs of 4) equal 12 turns in one call")
    from tests.machines import sc_ring, clone_bunch
    from pyheadtail_amd.engine import FusedOneTurnMap
    m1, b1 = sc_ring(4, 3., 128, 50000, seed=13)
    m2, _ = sc_ring(4, 3., 128, 10, seed=13)
    b2 = clone_bunch(b1)
    e1 = FusedOneTurnMap(m1, graph_turns=4)
    e1.track_turns(b1, 7)
    e1.track_turns(b1, 5)
    FusedOneTurnMap(m2, graph_turns=0).track_turns(b2, 12, use_graph=False)
    for k in COORDS:
        np.testing.assert_array_equal(getattr(b1, k).cpu().numpy(), getattr(b2, k).cpu().numpy(),
                                      err_msg=k)


def test_space_charge_statistics_records():
    _case("fused space charge with statistics records: same bits, records = bunch statistics")
    from tests.machines import sc_ring, clone_bunch
    from pyheadtail_amd.engine import FusedOneTurnMap
    m1, b1 = sc_ring(4, 3., 64, 40000, seed=14)
    m2, _ = sc_ring(4, 3., 64, 10, seed=14)
    b2 = clone_bunch(b1)
    rec = FusedOneTurnMap(m1, record_stats=True, graph_turns=4)
    rec.track_turns(b1, 9)
    b1._run_pending()          # the last wake kick and the final statistics row
    FusedOneTurnMap(m2, graph_turns=4).track_turns(b2, 9)
    for k in COORDS:
        np.testing.assert_array_equal(getattr(b1, k).cpu().numpy(), getattr(b2, k).cpu().numpy(),
                                      err_msg=k)
    st = rec.last_stats
    assert len(st['mean_x']) == 10
    # means against the spread of their coordinate (they are near zero)
    for name, scale in (('mean_x', 'sigma_x'), ('mean_yp', 'sigma_yp'), ('mean_dp', 'sigma_dp'),
                        ('sigma_x', 'sigma_x'), ('sigma_dp', 'sigma_dp'), ('epsn_x', 'epsn_x'),
                        ('epsn_z', 'epsn_z')):
        want = getattr(b1, name)()
        got = st[name][-1]
        assert abs(got - want) <= 1e-12 * abs(st[scale][-1]), (name, got, want)


def test_served_space_charge_job_full_size_vs_elements():
    _case("served space-charge job at 5e6 MP (8 segments, 700 slices): fused vs elements, 2 turns")
    import torch
    from tests.machines import sc_ring, clone_bunch
    from pyheadtail_amd import _phx
    from pyheadtail_amd.engine import FusedOneTurnMap
    from pyheadtail_amd.serve import ElementLoop
    m, b = sc_ring(8, 10., 700, 5000000, seed=15, wake_slices=700, intensity=2e11)
    b2 = clone_bunch(b)
    FusedOneTurnMap(m, graph_turns=0, pass_grid=_phx.STANDALONE_GRID).track_turns(b, 2)
    ElementLoop(m).track_turns(b2, 2)
    torch.cuda.synchronize()
    err = _rel_err(_host(b), _host(b2))
    print("[info] max |fused - elements| / rms = %.3g" % err, flush=True)
    assert err < 1e-12


def test_space_charge_engine_not_batched():
    _case("a space-charge job is refused by the batched scan")
    from tests.machines import sc_ring
    from pyheadtail_amd.batch import BatchedScan
    from pyheadtail_amd._phx import PhxError
    m, b = sc_ring(2, 3., 64, 20000, seed=16)
    scan = BatchedScan([(m, b)], graph_turns=4)
    with pytest.raises(PhxError, match="space-charge"):
        scan.track_turns(4)
