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


def _sc_job(seed, n=20000, n_seg=4, qp=3.):
    import numpy as np
    from scipy.constants import e, m_p
    import bench
    from pyheadtail_amd.machines.synchrotron import Synchrotron
    from pyheadtail_amd.particles.slicing import UniformBinSlicer
    from pyheadtail_amd.impedances.wakes import WakeField, CircularResonator
    from pyheadtail_amd.spacecharge.spacecharge import TransverseGaussianSpaceCharge
    mach = Synchrotron(optics_mode='smooth', charge=e, mass=m_p, p0=bench.p0_of_gamma(27.7),
                       circumference=6911.5, n_segments=n_seg, beta_x=54.6, beta_y=54.6,
                       D_x=0., D_y=0., accQ_x=20.13, accQ_y=20.18, Qp_x=qp, Qp_y=qp,
                       alpha_mom_compaction=1.9e-3, longitudinal_mode='linear', Q_s=0.017)
    mach.install_after_each_transverse_segment(
        TransverseGaussianSpaceCharge(UniformBinSlicer(64, n_sigma_z=3), 6911.5 / n_seg))
    np.random.seed(seed)
    b = mach.generate_6D_Gaussian_bunch(n, 1e11, bench.EPSN, bench.EPSN, bench.SIGMA_Z)
    return list(mach.one_turn_map) + [WakeField(UniformBinSlicer(100, n_sigma_z=3),
                                                CircularResonator(1e6, 1e9, 1))], b


def test_server_two_element_jobs_on_two_streams():
    """Two space-charge jobs (element by element: slicing, statistics and
    space-charge kernels with the library's scratch workspace) on two
    streams at once, and a statistics call on the caller's stream between
    the server's steps: each job ends as it does alone."""
    print("\n[case] scan server: two space-charge jobs on two streams", flush=True)
    from pyheadtail_amd.serve import ScanServer, ElementLoop
    turns = [17, 13]
    alone = []
    for seed, n in zip((81, 82), turns):
        m, b = _sc_job(seed, n=30000 + seed)
        ElementLoop(m).track_turns(b, n)
        alone.append({k: getattr(b, k).cpu().numpy() for k in COORDS})
    srv = ScanServer(n_streams=3, chunk_turns=16, graph_turns=4, element_chunk_turns=3,
                     big_job_bytes=1 << 40)
    jobs = [_sc_job(seed, n=30000 + seed) for seed in (81, 82)]
    handles = [srv.submit(m, b, n, fuse=False) for (m, b), n in zip(jobs, turns)]
    assert handles[0].bytes < srv.big_job_bytes
    side = _sc_job(90, n=50000)[1]
    while srv.busy:
        srv.step()
        side.sigma_x()
        side.epsn_y()
    assert handles[0].stream is not handles[1].stream
    for j, h in enumerate(handles):
        for k in COORDS:
            np.testing.assert_array_equal(getattr(h.bunch, k).cpu().numpy(), alone[j][k],
                                          err_msg="job %d %s" % (j, k))


@pytest.mark.parametrize("fuse_sc", [False, True])
def test_server_mixes_fused_and_element_jobs(fuse_sc):
    """A job with space-charge kicks, run element by element (fuse_sc
    False) or by the fused engine's space-charge stages, beside fused jobs;
    every job ends as it does alone."""
    print("\n[case] scan server: space-charge job (%s) beside two fused jobs"
          % ("fused" if fuse_sc else "elements"), flush=True)
    import numpy as np
    from scipy.constants import e, m_p
    import bench
    from pyheadtail_amd.engine import FusedOneTurnMap
    from pyheadtail_amd.machines.synchrotron import Synchrotron
    from pyheadtail_amd.particles.slicing import UniformBinSlicer
    from pyheadtail_amd.impedances.wakes import WakeField, CircularResonator
    from pyheadtail_amd.spacecharge.spacecharge import TransverseGaussianSpaceCharge
    from pyheadtail_amd.serve import ScanServer, ElementLoop

    def sc_job():
        mach = Synchrotron(optics_mode='smooth', charge=e, mass=m_p, p0=bench.p0_of_gamma(27.7),
                           circumference=6911.5, n_segments=4, beta_x=54.6, beta_y=54.6,
                           D_x=0., D_y=0., accQ_x=20.13, accQ_y=20.18, Qp_x=3., Qp_y=3.,
                           alpha_mom_compaction=1.9e-3, longitudinal_mode='linear', Q_s=0.017)
        mach.install_after_each_transverse_segment(
            TransverseGaussianSpaceCharge(UniformBinSlicer(64, n_sigma_z=3), 6911.5 / 4))
        np.random.seed(77)
        b = mach.generate_6D_Gaussian_bunch(20000, 1e11, bench.EPSN, bench.EPSN, bench.SIGMA_Z)
        return list(mach.one_turn_map) + [WakeField(UniformBinSlicer(100, n_sigma_z=3),
                                                    CircularResonator(1e6, 1e9, 1))], b

    turns = [23, 19, 30]
    m, b = sc_job()
    if fuse_sc:
        FusedOneTurnMap(m, graph_turns=8).track_turns(b, turns[0])
    else:
        ElementLoop(m).track_turns(b, turns[0])
    alone = [{k: getattr(b, k).cpu().numpy() for k in COORDS}]
    for i, n in zip((1, 3), turns[1:]):
        m, b = bench.scan_job(i, n_scale=0.01)
        FusedOneTurnMap(m, graph_turns=8).track_turns(b, n)
        alone.append({k: getattr(b, k).cpu().numpy() for k in COORDS})
    srv = ScanServer(n_streams=2, chunk_turns=16, graph_turns=8, element_chunk_turns=5)
    jobs = [sc_job(), bench.scan_job(1, n_scale=0.01), bench.scan_job(3, n_scale=0.01)]
    handles = [srv.submit(m, b, n, fuse=(fuse_sc or i > 0))
               for i, ((m, b), n) in enumerate(zip(jobs, turns))]
    srv.drain()
    assert handles[0].fuse == fuse_sc
    assert [h.engine is None for h in handles] == [True] * 3
    for j, h in enumerate(handles):
        for k in COORDS:
            np.testing.assert_array_equal(getattr(h.bunch, k).cpu().numpy(), alone[j][k],
                                          err_msg="job %d %s" % (j, k))
