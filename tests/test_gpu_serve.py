"""The scan server: jobs submitted at different times, queued behind an HBM
budget that admits only some of them at once, advanced in chunks on shared
streams, each end with exactly the coordinates they reach alone; the budget
is respected and returned."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

COORDS = ('x', 'xp', 'y', 'yp', 'z', 'dp')


def test_server_results_equal_each_job_alone():
    print("\n[case] scan server: 5 jobs, budget for about two, staggered submissions", flush=True)
    import bench
    from pyheadtail_amd.engine import FusedOneTurnMap
    from pyheadtail_amd.serve import ScanServer, job_bytes
    ids, turns = [0, 1, 2, 3, 5], [30, 55, 21, 70, 40]
    alone = []
    for i, n in zip(ids, turns):
        m, b = bench.scan_job(i, n_scale=0.01)
        FusedOneTurnMap(m, graph_turns=8).track_turns(b, n)
        alone.append({k: getattr(b, k).cpu().numpy() for k in COORDS})
    jobs = [bench.scan_job(i, n_scale=0.01) for i in ids]
    sizes = sorted(job_bytes(m, b) for m, b in jobs)
    srv = ScanServer(hbm_budget_bytes=sizes[-1] + sizes[-2], n_streams=3, chunk_turns=16,
                     graph_turns=8)
    handles = [srv.submit(jobs[0][0], jobs[0][1], turns[0], job_id=0),
               srv.submit(jobs[1][0], jobs[1][1], turns[1], job_id=1)]
    srv.step()
    handles.append(srv.submit(jobs[2][0], jobs[2][1], turns[2], job_id=2))
    srv.step()
    for q in (3, 4):
        handles.append(srv.submit(jobs[q][0], jobs[q][1], turns[q], job_id=q))
    assert srv.used <= srv.budget
    srv.drain()
    assert srv.used == 0 and not srv.active and not srv.queue
    assert len(srv.done) == 5
    waited = [h for h in handles if h.t_admit > h.t_submit + 1e-4]
    print("[info] %d jobs waited in the queue" % len(waited), flush=True)
    for j, h in enumerate(handles):
        for k in COORDS:
            np.testing.assert_array_equal(getattr(h.bunch, k).cpu().numpy(), alone[j][k],
                                          err_msg="job %d %s" % (j, k))
    p = srv.latency_percentiles()
    assert p[50] > 0 and p[99] >= p[50]
