"""Batched scan: independent bunches advanced together on several HIP
streams end with exactly the coordinates each reaches alone."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def test_batched_scan_equals_each_job_alone():
    print("\n[case] 5 bunches (2 resonator, 3 resistive wall), 3 streams, 27 turns", flush=True)
    import bench
    from pyheadtail_amd.batch import BatchedScan
    from pyheadtail_amd.engine import FusedOneTurnMap
    ids = [0, 1, 2, 3, 5]
    alone = []
    for i in ids:
        m, b = bench.scan_job(i, n_scale=0.02)
        eng = FusedOneTurnMap(m, graph_turns=4)
        eng.track_turns(b, 27)
        alone.append({k: getattr(b, k).cpu().numpy() for k in ('x', 'xp', 'y', 'yp', 'z', 'dp')})
    scan = BatchedScan([bench.scan_job(i, n_scale=0.02) for i in ids], n_streams=3,
                       chunk_turns=10, graph_turns=4)
    scan.track_turns(27)
    for j, b in enumerate(scan.bunches):
        for k in ('x', 'xp', 'y', 'yp', 'z', 'dp'):
            np.testing.assert_array_equal(getattr(b, k).cpu().numpy(), alone[j][k],
                                          err_msg="job %d %s" % (ids[j], k))
