"""Batched scan: independent bunches advanced together -- one launch per
stage for all of them (fused mode) or per-job graphs on several HIP streams
-- end with exactly the coordinates each reaches alone with the same pass
grid; the autotuner's choice is stored and reused."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

COORDS = ('x', 'xp', 'y', 'yp', 'z', 'dp')
IDS = [0, 1, 2, 3, 5]


def _alone(ids, turns, n_scale, pass_grid=0):
    import bench
    from pyheadtail_amd.engine import FusedOneTurnMap
    out = []
    for i in ids:
        m, b = bench.scan_job(i, n_scale=n_scale)
        FusedOneTurnMap(m, graph_turns=4, pass_grid=pass_grid).track_turns(b, turns)
        out.append({k: getattr(b, k).cpu().numpy() for k in COORDS})
    return out


def _check(scan, alone, ids):
    for j, b in enumerate(scan.bunches):
        for k in COORDS:
            np.testing.assert_array_equal(getattr(b, k).cpu().numpy(), alone[j][k],
                                          err_msg="job %s %s" % (ids[j], k))


@pytest.mark.parametrize("mode,pass_grid", [("fused", 0), ("fused", 128), ("streams", 0)])
def test_batched_scan_equals_each_job_alone(mode, pass_grid):
    print("\n[case] 5 bunches (2 resonator, 3 resistive wall), %s, pass grid %d, 27 turns"
          % (mode, pass_grid), flush=True)
    import bench
    from pyheadtail_amd.batch import BatchedScan
    alone = _alone(IDS, 27, 0.02, pass_grid)
    scan = BatchedScan([bench.scan_job(i, n_scale=0.02) for i in IDS], n_streams=3,
                       chunk_turns=10, graph_turns=4, mode=mode, pass_grid=pass_grid)
    scan.track_turns(11)
    scan.track_turns(16)
    _check(scan, alone, IDS)


def test_batch_groups_of_different_passes():
    """Jobs whose maps compile to different passes (linear map vs RF system,
    detuner sets) form separate batches, each still bit-identical alone."""
    print("\n[case] batch of scan jobs and RF-ring jobs (two pass variants)", flush=True)
    import bench
    from oracle import workloads as W
    from tests.machines import nonlinear_machine, to_product
    from pyheadtail_amd.batch import BatchedScan
    from pyheadtail_amd.engine import FusedOneTurnMap
    obs = [W.nonlinear_bunch(30000, seed=s) for s in (5, 6)]
    alone = _alone([1, 3], 9, 0.02)
    for ob in obs:
        pb = to_product(ob)
        FusedOneTurnMap(nonlinear_machine(), graph_turns=4).track_turns(pb, 9)
        alone.append({k: getattr(pb, k).cpu().numpy() for k in COORDS})
    jobs = [bench.scan_job(i, n_scale=0.02) for i in (1, 3)]
    jobs += [(nonlinear_machine(), to_product(ob)) for ob in obs]
    scan = BatchedScan(jobs, graph_turns=4)
    scan.track_turns(9)
    assert len(scan._groups) == 2
    _check(scan, alone, [1, 3, 'rf0', 'rf1'])


def test_per_job_grids_equal_each_job_alone():
    """Pass grids sized per job (groups_per_block): every job bit-identical
    to its run alone with its own grid."""
    print("\n[case] 5 bunches, pass grids sized per job (2 groups per block), 21 turns",
          flush=True)
    import bench
    from pyheadtail_amd.batch import BatchedScan
    from pyheadtail_amd.engine import FusedOneTurnMap
    jobs = [bench.scan_job(i, n_scale=0.02) for i in IDS]
    grids = [BatchedScan.grid_for(b.macroparticlenumber, 2) for _, b in jobs]
    assert len(set(grids)) > 1
    alone = []
    for i, g in zip(IDS, grids):
        m, b = bench.scan_job(i, n_scale=0.02)
        FusedOneTurnMap(m, graph_turns=4, pass_grid=g).track_turns(b, 21)
        alone.append({k: getattr(b, k).cpu().numpy() for k in COORDS})
    scan = BatchedScan(jobs, graph_turns=4, groups_per_block=2)
    assert [e.pass_grid for e in scan.engines] == grids
    scan.track_turns(21)
    _check(scan, alone, IDS)


def test_autotuner_stores_and_reuses(tmp_path):
    print("\n[case] autotuner: measured choice stored, reused, results as alone", flush=True)
    import bench
    from pyheadtail_amd.batch import BatchedScan
    from pyheadtail_amd.tune import Tuner, tune_batch
    path = str(tmp_path / "tune.json")
    scan = BatchedScan([bench.scan_job(i, n_scale=0.02) for i in IDS], graph_turns=4)
    entry = tune_batch(scan, tuner=Tuner(path), turns=4, pass_grids=(0, 128),
                       groups_per_block=(3,), graph_turns=(4, 8))
    assert len(entry['candidates']) == 6
    assert scan.pass_grid == entry['pass_grid']
    assert scan.groups_per_block == entry['groups_per_block']
    again = tune_batch(scan, tuner=Tuner(path))
    assert again == entry
    # a fresh scan with the stored choice, against the jobs alone at its grids
    fresh = BatchedScan([bench.scan_job(i, n_scale=0.02) for i in IDS], graph_turns=4)
    tune_batch(fresh, tuner=Tuner(path))
    assert fresh.pass_grid == entry['pass_grid']
    grids = [e.pass_grid for e in fresh.engines]
    fresh.track_turns(13)
    alone = []
    from pyheadtail_amd.engine import FusedOneTurnMap
    for i, g in zip(IDS, grids):
        m, b = bench.scan_job(i, n_scale=0.02)
        FusedOneTurnMap(m, graph_turns=4, pass_grid=g).track_turns(b, 13)
        alone.append({k: getattr(b, k).cpu().numpy() for k in COORDS})
    _check(fresh, alone, IDS)
