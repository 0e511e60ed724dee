"""Benchmark of the fused one-turn map on one MI355X.

Workload (the reference's "full plain one-turn map on GPU, FP64"): protons at
gamma = 27.7 on a ring of C = 6911.5 m, 1e6 macroparticles, intensity
1.5e11; one smooth-optics transverse segment (Q = 20.13 / 20.18,
beta = 54.6 m) with chromaticity Q' = 5 and amplitude detuning
app_x = app_y = 1e-10, app_xy = -5e-11 per m (times p0); non-linear
RFSystems h = 4620, V = 4.5 MV; UniformBinSlicer 500 slices over +-3 sigma_z;
CircularResonator R_s = 1e6 Ohm/m, f_r = 1 GHz, Q = 1, one turn of wake
memory.  A step is one turn of the whole bunch; value = macroparticle-turns/s
with the bunch resident in device memory.

The roofline entry is that of the dominant kernel (the fused particle pass):
96 algorithmic bytes per macroparticle (read and write of six FP64
coordinates) divided by its mean duration, measured with HIP events on its
launch stream; turn_roofline_frac is the same 96 bytes over the whole turn
as the timed (graph-replayed) turns take it.
A second leg ("large") runs the same map on 1e7 macroparticles, whose 480 MB
of coordinates exceed the 256 MB last-level cache.  The CPU baseline is the
oracle's multi-core C++ restatement of the reference CPU path on all host
threads; cpu_baseline_numpy_1core is its NumPy restatement on one core (the
reference's CPU path is single-threaded NumPy).  Both run on bounded
samples of the 1e6 workload.

Usage: python bench.py [--gpus 1] [--steps 1000] [--warmup 50]

--workload rw64 measures instead the multi-turn resistive-wall workload:
1e6 macroparticles, LinearMap (alpha_c = 1.9e-3, Q_s = 0.017), one smooth
transverse segment, 1000 slices over +-3 sigma_z and a parallel-plates
resistive wall (b = 0.02 m, L = 6911.5 m, sigma_c = 5.9e7 S/m, dt_min =
1e-12 s, dipolar and quadrupolar x/y kicks) with 64 turns of memory, in
every precision of the wall convolution (FP64 with the past-turn
expansion, FP64 pair by pair, FP32, FP32 MFMA, split-bfloat16 MFMA), and
each precision's kick error against FP64 pair by pair per kick component
on the bunch's own 64-turn history.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BYTES_PER_MP_PASS = 96
EPSN = 2.5e-6          # normalised emittance, both planes [m]
SIGMA_Z = 0.23         # bunch length [m]
INTENSITY = 1.5e11
SEED = 12345


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--n", type=int, default=1000000)
    ap.add_argument("--slices", type=int, default=500)
    ap.add_argument("--cpu-turns", type=int, default=30,
                    help="turns of the one-core NumPy baseline")
    ap.add_argument("--cpu-seconds", type=float, default=10.,
                    help="length of the all-core C++ baseline's sample")
    ap.add_argument("--n-large", type=int, default=10000000,
                    help="macroparticles of the second, cache-exceeding leg (0: none)")
    ap.add_argument("--steps-large", type=int, default=100)
    ap.add_argument("--cpu-n", type=int, default=1000000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--stage-turns", type=int, default=100)
    ap.add_argument("--graph-turns", type=int, default=10)
    ap.add_argument("--workload", choices=("resonator", "rw64", "scan64", "serve", "spacecharge"),
                    default="resonator")
    ap.add_argument("--n-sc", type=int, default=8000000,
                    help="macroparticles of the space-charge workload")
    ap.add_argument("--steps-sc", type=int, default=50)
    ap.add_argument("--jobs", type=int, default=48)
    ap.add_argument("--big-job-mib", type=float, default=128.,
                    help="batched and served jobs of this many MiB or more share one stream")
    ap.add_argument("--no-sc", action="store_true",
                    help="served stream without the space-charge kicks of the largest jobs")
    ap.add_argument("--streams", type=int, default=12)
    ap.add_argument("--tune-file", default=None,
                    help="autotuner table (default ~/.cache/pyheadtail_amd/tune.json)")
    ap.add_argument("--retune", action="store_true", help="measure again even if stored")
    ap.add_argument("--traffic-from", metavar="DIR",
                    help="only record profiles/pass_traffic.json from rocprofv3 --pmc output "
                         "under DIR: subdirectories fetch_pass, write_pass, fetch_copy and "
                         "write_copy (FETCH_SIZE / WRITE_SIZE of the pass at --n "
                         "macroparticles and of --copy-calibration at the same --n)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain device copy of the pass's algorithmic bytes "
                         "(48 B read + 48 B written per macroparticle), for scaling the "
                         "FETCH_SIZE / WRITE_SIZE counters")
    return ap.parse_args()


def copy_calibration(n, reps=5):
    """Plain device-to-device copy of 6 FP64 arrays of n elements (the bytes a
    turn pass must move), repeated reps times for the counters, then timed
    over 100 more copies (the rate a streaming kernel reaches)."""
    import torch
    torch.cuda.set_device(0)
    src = torch.rand((6, n), dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    # the copy's rate: 100 more copies back to back between HIP events
    timed = 100
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(timed):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    us = e0.elapsed_time(e1) * 1e3 / timed
    nbytes = 2 * src.numel() * 8
    emit(json.dumps({"copy_calibration": {"bytes_read": src.numel() * 8,
                                           "bytes_written": dst.numel() * 8,
                                           "reps": reps + timed, "us_per_copy": round(us, 2),
                                           "TB_per_s": round(nbytes / us / 1e6, 3)}}))


def _cpu_threads():
    """Host threads for the all-core baseline: the process's CPU set, capped
    by OMP_NUM_THREADS when that is set (the GPU machine's CPU share)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(1, n)


def cpu_baseline(n, seconds):
    """The oracle's multi-core C++ restatement of the reference CPU path on
    all host threads, on a bounded sample of the same workload: as many
    turns of the n-macroparticle bunch as take about `seconds`.  Runs before
    the GPU is touched."""
    from oracle import workloads as W
    from oracle import cpu_turn as CT
    ob = W.nonlinear_bunch(n, seed=SEED)
    ct = CT.CpuTurns(ob, n_threads=_cpu_threads())
    ct.track_turns(2)                         # untimed: pages, thread pool
    t0 = time.perf_counter()
    ct.track_turns(3)
    rate = 3 / (time.perf_counter() - t0)
    turns = max(5, int(seconds * rate))
    t0 = time.perf_counter()
    ct.track_turns(turns)
    ct.flush()
    dt = time.perf_counter() - t0
    return dict(value=n * turns / dt, unit="macroparticle-turns/s", cores=ct.threads,
                kind="port",
                sample="%d macroparticles x %d turns of the same one-turn map (C++ restatement "
                       "of the reference CPU path, oracle/cpu_turn.cpp, %d threads, %.1f s)"
                       % (n, turns, ct.threads, dt))


def cpu_baseline_numpy(n, turns):
    """The oracle's NumPy restatement of the reference CPU path (which is
    single-threaded NumPy) on one host core."""
    try:
        old = os.sched_getaffinity(0)
        os.sched_setaffinity(0, {sorted(old)[0]})
    except (AttributeError, OSError):
        old = None
    try:
        from oracle import workloads as W
        ob = W.nonlinear_bunch(n, seed=SEED)
        m = W.nonlinear_map_oracle()
        for el in m:          # one untimed turn (imports, caches)
            el.track(ob)
        t0 = time.perf_counter()
        for _ in range(turns):
            for el in m:
                el.track(ob)
        dt = time.perf_counter() - t0
    finally:
        if old is not None:
            os.sched_setaffinity(0, old)
    return dict(value=n * turns / dt, unit="macroparticle-turns/s", cores=1, kind="port",
                sample="%d macroparticles x %d turns of the same one-turn map "
                       "(NumPy restatement of the reference CPU path, 1 core, %.1f s)"
                       % (n, turns, dt))


TRAFFIC_SOURCES = ("pyheadtail_amd/csrc", "include")


def sources_sha256():
    """Digest of the HIP sources and the ABI header the library is built
    from, which a traffic record must carry to be used."""
    import hashlib
    h = hashlib.sha256()
    for d in TRAFFIC_SOURCES:
        base = os.path.join(ROOT, d)
        for name in sorted(os.listdir(base)):
            if name.endswith((".hip", ".h")):
                h.update(name.encode())
                with open(os.path.join(base, name), "rb") as f:
                    h.update(f.read())
    return h.hexdigest()


def pass_traffic(n):
    """HBM bytes per launch of the particle pass from the rocprofv3
    FETCH_SIZE / WRITE_SIZE runs recorded in profiles/pass_traffic.json, each
    scaled by what the counter reads for a plain copy of known size; None
    unless the record is for n macroparticles and for the sources this
    library is built from."""
    path = os.path.join(ROOT, "profiles", "pass_traffic.json")
    try:
        with open(path) as f:
            recs = json.load(f)
    except (OSError, ValueError):
        return None
    sha = sources_sha256()
    for t in recs if isinstance(recs, list) else [recs]:
        if t.get("macroparticles") != n or t.get("sources_sha256") != sha:
            continue
        read = t["fetch_kib"] * t["copy_bytes"] / t["copy_fetch_kib"]
        written = t["write_kib"] * t["copy_bytes"] / t["copy_write_kib"]
        return read + written
    return None


def _counter_means(d, counter, match):
    """Mean per dispatch of a rocprofv3 counter over the kernels whose name
    contains `match`, from the counter_collection CSV files under d."""
    import csv
    import glob
    vals = {}
    for path in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
        with open(path) as f:
            for row in csv.DictReader(f):
                if row.get("Counter_Name") != counter or match not in row.get("Kernel_Name", ""):
                    continue
                key = row.get("Dispatch_Id") or row.get("Correlation_Id") or len(vals)
                vals[key] = vals.get(key, 0.0) + float(row["Counter_Value"])
    if not vals:
        raise SystemExit("no %s values for kernels matching %r under %s" % (counter, match, d))
    return sum(vals.values()) / len(vals), len(vals)


def record_traffic(d, n):
    """profiles/pass_traffic.json: the pass's FETCH_SIZE and WRITE_SIZE (KiB
    per dispatch) at n macroparticles and the same counters for a plain copy
    of the six coordinate arrays, with the digest of the sources built."""
    fetch, nd = _counter_means(os.path.join(d, "fetch_pass"), "FETCH_SIZE", "k_engine_pass")
    write, _ = _counter_means(os.path.join(d, "write_pass"), "WRITE_SIZE", "k_engine_pass")
    cfetch, _ = _counter_means(os.path.join(d, "fetch_copy"), "FETCH_SIZE", "opy")
    cwrite, _ = _counter_means(os.path.join(d, "write_copy"), "WRITE_SIZE", "opy")
    rec = {"kernel": "k_engine_pass (headline map)", "macroparticles": n,
           "fetch_kib": fetch, "write_kib": write, "dispatches": nd,
           "copy_bytes": 48 * n, "copy_fetch_kib": cfetch, "copy_write_kib": cwrite,
           "sources_sha256": sources_sha256()}
    path = os.path.join(ROOT, "profiles", "pass_traffic.json")
    try:
        with open(path) as f:
            recs = json.load(f)
        recs = recs if isinstance(recs, list) else [recs]
    except (OSError, ValueError):
        recs = []
    recs = [r for r in recs if r.get("macroparticles") != n] + [rec]
    with open(path, "w") as f:
        json.dump(sorted(recs, key=lambda r: r["macroparticles"]), f, indent=1)
    emit(json.dumps(rec))


def p0_of_gamma(gamma):
    from scipy.constants import c, m_p
    import numpy as np
    return np.sqrt(gamma**2 - 1) * m_p * c


def build_machine(p0_ref):
    from scipy.constants import e, m_p
    from pyheadtail_amd.machines.synchrotron import Synchrotron
    mach = Synchrotron(optics_mode='smooth', charge=e, mass=m_p, p0=p0_ref,
                       circumference=6911.5, n_segments=1, beta_x=54.6, beta_y=54.6,
                       D_x=0., D_y=0., accQ_x=20.13, accQ_y=20.18, Qp_x=5., Qp_y=5.,
                       app_x=1e-10 * p0_ref, app_y=1e-10 * p0_ref, app_xy=-5e-11 * p0_ref,
                       longitudinal_mode='non-linear', alpha_mom_compaction=1.9e-3,
                       h_RF=4620, V_RF=4.5e6, dphi_RF=0., p_increment=0.)
    return mach


RW64_PRECISIONS = ('fp64', 'fp64_direct', 'fp32', 'mfma_fp32', 'mfma_bf16x3', 'mfma_bf16x6')


def rw64(args):
    """The multi-turn resistive-wall workload in every precision of the wall
    convolution (one JSON line): turn rate, stage times, and each
    precision's kick error per component against FP64 pair by pair on the
    bunch's own 64-turn history."""
    import numpy as np
    import torch
    torch.cuda.set_device(0)
    from scipy.constants import e, m_p
    from pyheadtail_amd import _phx
    from pyheadtail_amd.engine import FusedOneTurnMap
    from pyheadtail_amd.machines.synchrotron import Synchrotron
    from pyheadtail_amd.particles.slicing import UniformBinSlicer
    from pyheadtail_amd.impedances.wakes import WakeField, ParallelHorizontalPlatesResistiveWall
    n, S, T = args.n, 1000 if args.slices == 500 else args.slices, 64
    steps = min(args.steps, 500)
    mach = Synchrotron(optics_mode='smooth', charge=e, mass=m_p, p0=p0_of_gamma(27.7),
                       circumference=6911.5, n_segments=1, beta_x=54.6, beta_y=54.6,
                       D_x=0., D_y=0., accQ_x=20.13, accQ_y=20.18,
                       alpha_mom_compaction=1.9e-3, longitudinal_mode='linear', Q_s=0.017)
    res, engines = {}, {}
    for prec in RW64_PRECISIONS:
        wake = WakeField(UniformBinSlicer(S, n_sigma_z=3),
                         ParallelHorizontalPlatesResistiveWall(0.02, 6911.5, 5.9e7, 1e-12,
                                                               n_turns_wake=T),
                         precision=prec)
        np.random.seed(SEED)
        bunch = mach.generate_6D_Gaussian_bunch(n, INTENSITY, EPSN, EPSN, SIGMA_Z)
        eng = FusedOneTurnMap(list(mach.one_turn_map) + [wake], graph_turns=args.graph_turns)
        warm = T + max(args.warmup, args.graph_turns)
        eng.reserve(bunch, max(warm, steps, args.stage_turns))
        eng.track_turns(bunch, warm)  # history filled
        bunch.x
        torch.cuda.synchronize()
        stage_ms = eng.timed_turns(bunch, min(args.stage_turns, 50))
        bunch.x
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        eng.track_turns(bunch, steps)
        bunch.x
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        res[prec] = {"value": n * steps / dt, "ms_per_step": dt / steps * 1e3,
                     "stage_ms": {"pass": stage_ms[0], "hist": stage_ms[1],
                                  "conv": stage_ms[2]}}
        print("[rw64] %s: %.1f us/turn, convolution stage %.1f us"
              % (prec, dt / steps * 1e6, stage_ms[2] * 1e3), file=sys.stderr, flush=True)
        if prec == 'fp64':
            engines[prec] = (eng, wake, bunch)
        else:
            del eng, wake, bunch
    # every precision against FP64 pair by pair on the FP64 run's own history
    eng, wake, bunch = engines['fp64']
    ring = eng._bufs['history']
    ring.n_pushed = T
    recs = [k.record(bunch) for k in wake.wake_kicks]
    for r in recs:
        r.precision = _phx.PREC_FP64_DIRECT
    kref = ring.convolve(recs).cpu().numpy()
    errs = {}
    for prec in RW64_PRECISIONS:
        if prec == 'fp64_direct':
            continue
        for r in recs:
            r.precision = WakeField.PRECISIONS[prec]
        kp = ring.convolve(recs).cpu().numpy()
        errs[prec] = {}
        for i, k in enumerate(wake.wake_kicks):
            d = kp[i] - kref[i]
            errs[prec][type(k).__name__] = {
                "max_rel": float(np.max(np.abs(d)) / np.max(np.abs(kref[i]))),
                "rms_rel": float(np.sqrt(np.mean(d * d)) / np.sqrt(np.mean(kref[i] ** 2)))}
    flop = 2.0 * S * S * T * 3   # pairs x the three distinct moments, per turn
    out = {
        "metric": "macroparticle-turns/s, multi-turn resistive-wall wake",
        "value": res['fp64']['value'], "unit": "macroparticle-turns/s", "n_gpus": 1,
        "steps": steps, "warmup": args.warmup, "ms_per_step": res['fp64']['ms_per_step'],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded 6D Gaussian bunch)",
        "config": {"workload": "multi-turn resistive wall: %d macroparticles, LinearMap, %d "
                               "slices +-3 sigma_z, parallel-plates resistive wall (4 kicks), "
                               "%d turns of memory" % (n, S, T),
                   "macroparticles": n, "n_slices": S, "n_turns_wake": T},
        "precisions": res,
        "contraction_flop_per_turn": flop,
        "contraction_tflops": {p: flop / (res[p]['stage_ms']['conv'] * 1e-3) / 1e12
                               for p in RW64_PRECISIONS},
        "kick_error_vs_fp64_direct": errs,
    }
    emit(json.dumps(out))


def scan_params(i, seed_base=1000):
    """Parameters of batched-scan job i: macroparticles log-uniform in
    [2e5, 4e6], slices in {100, 250, 500, 1000}, intensity uniform in
    [1e10, 3e11], chromaticity uniform in [-5, 15], n_turns_wake in
    {1, 8, 32}; drawn from a generator seeded by seed_base + i."""
    import numpy as np
    rng = np.random.default_rng(seed_base + i)
    n = np.exp(rng.uniform(np.log(2e5), np.log(4e6)))
    S = int(rng.choice([100, 250, 500, 1000]))
    intensity = rng.uniform(1e10, 3e11)
    qp = rng.uniform(-5., 15.)
    T = int(rng.choice([1, 8, 32]))
    return dict(n=n, n_slices=S, intensity=intensity, qp=qp, n_turns_wake=T)


def scan_job(i, n_scale=1.0, seed_base=1000, bunch=True):
    """Job i of the batched scan (parameters from scan_params): the
    LinearMap ring with a broadband resonator (one turn of memory) or the
    parallel-plates resistive wall (8 or 32 turns), and its Gaussian bunch
    seeded by seed_base + i (None unless bunch)."""
    import numpy as np
    from scipy.constants import e, m_p
    from pyheadtail_amd.machines.synchrotron import Synchrotron
    from pyheadtail_amd.particles.slicing import UniformBinSlicer
    from pyheadtail_amd.impedances.wakes import (WakeField, CircularResonator,
                                                 ParallelHorizontalPlatesResistiveWall)
    p = scan_params(i, seed_base)
    n = int(p['n'] * n_scale)
    S, T, qp = p['n_slices'], p['n_turns_wake'], p['qp']
    mach = Synchrotron(optics_mode='smooth', charge=e, mass=m_p, p0=p0_of_gamma(27.7),
                       circumference=6911.5, n_segments=1, beta_x=54.6, beta_y=54.6, D_x=0.,
                       D_y=0., accQ_x=20.13, accQ_y=20.18, Qp_x=qp, Qp_y=qp,
                       alpha_mom_compaction=1.9e-3, longitudinal_mode='linear', Q_s=0.017)
    src = (CircularResonator(1e6, 1e9, 1) if T == 1 else
           ParallelHorizontalPlatesResistiveWall(0.02, 6911.5, 5.9e7, 1e-12, n_turns_wake=T))
    one_turn = list(mach.one_turn_map) + [WakeField(UniformBinSlicer(S, n_sigma_z=3), src)]
    if not bunch:
        return one_turn, None
    np.random.seed(seed_base + i)
    b = mach.generate_6D_Gaussian_bunch(n, p['intensity'], EPSN, EPSN, SIGMA_Z)
    return one_turn, b


def scan64(args):
    """64 independent bunches: one launch per stage for all of them (tuned
    launch shape) against the plain GPU path job by job and against the
    per-job graphs on several streams (one JSON line)."""
    import torch
    torch.cuda.set_device(0)
    from pyheadtail_amd.batch import BatchedScan
    from pyheadtail_amd.tune import Tuner, tune_batch
    jobs = [scan_job(i) for i in range(64)]
    scan = BatchedScan(jobs, mode='fused', graph_turns=args.graph_turns)
    n_mp = scan.macroparticles
    steps = min(args.steps, 200)
    tuner = Tuner(args.tune_file) if args.tune_file else Tuner()
    entry = tune_batch(scan, tuner=tuner, turns=20, force=args.retune)
    print("[scan64] tuned: pass_grid %d, groups_per_block %d, graph_turns %d (%.3f ms per turn)"
          % (entry['pass_grid'], entry.get('groups_per_block', 0), entry['graph_turns'],
             entry['ms_per_turn']), file=sys.stderr, flush=True)
    scan.track_turns(40)                      # histories (T <= 32) filled, graphs captured
    t0 = time.perf_counter()
    scan.track_turns(steps)
    t_bat = time.perf_counter() - t0
    # the plain GPU path job by job: each job's own engine at its default
    # pass grid, one after the other
    scan.configure(pass_grid=0, groups_per_block=0)
    scan.track_turns_job_by_job(40)
    t0 = time.perf_counter()
    scan.track_turns_job_by_job(steps)
    t_seq = time.perf_counter() - t0
    scan.configure(pass_grid=entry['pass_grid'], groups_per_block=entry.get('groups_per_block', 0))
    # per-job graphs on 12 streams
    streams = BatchedScan(jobs, mode='streams', n_streams=args.streams, chunk_turns=50,
                          graph_turns=args.graph_turns,
                          big_job_bytes=int(args.big_job_mib * (1 << 20)))
    streams.track_turns(40)
    t0 = time.perf_counter()
    streams.track_turns(steps)
    t_str = time.perf_counter() - t0
    out = {
        "metric": "aggregate macroparticle-turns/s, 64-bunch scan",
        "value": n_mp * steps / t_bat, "unit": "macroparticle-turns/s", "n_gpus": 1,
        "steps": steps, "warmup": 40, "ms_per_step": t_bat / steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (64 seeded Gaussian bunches)",
        "config": {"workload": "64 independent bunches: macroparticles log-uniform in "
                               "[2e5, 4e6] (%d in total), slices in {100,250,500,1000}, "
                               "chromaticity in [-5, 15], n_turns_wake in {1,8,32}; LinearMap "
                               "ring, resonator or resistive-wall wake" % n_mp,
                   "bunches": 64, "macroparticles": n_mp},
        "tuned": {k: entry.get(k, 0) for k in ('pass_grid', 'groups_per_block', 'graph_turns',
                                                 'ms_per_turn', 'candidates')},
        "job_by_job": {"value": n_mp * steps / t_seq, "ms_per_step": t_seq / steps * 1e3},
        "streams": {"value": n_mp * steps / t_str, "ms_per_step": t_str / steps * 1e3,
                    "streams": args.streams},
        "speedup_vs_job_by_job": t_seq / t_bat,
        "speedup_streams_vs_job_by_job": t_seq / t_str,
    }
    emit(json.dumps(out))


def sc_machine(qp, n_slices, n_seg=8):
    """The served ring with space charge: smooth optics in n_seg segments,
    chromaticity qp, the LinearMap (Q_s = 0.017), a 2.5D transverse
    Gaussian space-charge kick of length C/n_seg after each segment on its
    own n_slices-slice uniform slicing (+-3 sigma_z)."""
    from scipy.constants import e, m_p
    from pyheadtail_amd.machines.synchrotron import Synchrotron
    from pyheadtail_amd.particles.slicing import UniformBinSlicer
    from pyheadtail_amd.spacecharge.spacecharge import TransverseGaussianSpaceCharge
    mach = Synchrotron(optics_mode='smooth', charge=e, mass=m_p, p0=p0_of_gamma(27.7),
                       circumference=6911.5, n_segments=n_seg, beta_x=54.6, beta_y=54.6,
                       D_x=0., D_y=0., accQ_x=20.13, accQ_y=20.18, Qp_x=qp, Qp_y=qp,
                       alpha_mom_compaction=1.9e-3, longitudinal_mode='linear', Q_s=0.017)
    mach.install_after_each_transverse_segment(
        TransverseGaussianSpaceCharge(UniformBinSlicer(n_slices, n_sigma_z=3), 6911.5 / n_seg))
    return mach


def spacecharge(args):
    """A served job with space charge at full size: n_sc macroparticles,
    700 slices, 8 segments with a space-charge kick after each, the
    broadband resonator; turns of the fused engine (space-charge stages)
    against the same turns element by element."""
    import numpy as np
    import torch
    torch.cuda.set_device(0)
    from pyheadtail_amd.engine import FusedOneTurnMap
    from pyheadtail_amd.serve import ElementLoop
    from pyheadtail_amd.particles.slicing import UniformBinSlicer
    from pyheadtail_amd.impedances.wakes import WakeField, CircularResonator
    n, S = args.n_sc, 700
    mach = sc_machine(10., S)
    one_turn = list(mach.one_turn_map) + [WakeField(UniformBinSlicer(S, n_sigma_z=3),
                                                    CircularResonator(1e6, 1e9, 1))]
    np.random.seed(SEED)
    b = mach.generate_6D_Gaussian_bunch(n, 2e11, EPSN, EPSN, SIGMA_Z)
    eng = FusedOneTurnMap(one_turn, graph_turns=args.graph_turns)
    eng.reserve(b, args.steps_sc)
    eng.track_turns(b, max(args.warmup // 10, args.graph_turns))
    torch.cuda.synchronize()
    t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0.record()
    eng.track_turns(b, args.steps_sc)
    t1.record()
    torch.cuda.synchronize()
    ms = t0.elapsed_time(t1) / args.steps_sc
    b._run_pending()
    el = ElementLoop(one_turn)
    el.track_turns(b, 1)
    torch.cuda.synchronize()
    k = 3
    tt = time.perf_counter()
    el.track_turns(b, k)
    torch.cuda.synchronize()
    ms_el = (time.perf_counter() - tt) / k * 1e3
    out = {
        "metric": "space-charge job throughput (macroparticle-turns per second)",
        "value": n / ms * 1e3, "unit": "MP-turns/s", "n_gpus": 1, "steps": args.steps_sc,
        "warmup": max(args.warmup // 10, args.graph_turns), "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded Gaussian bunch)",
        "config": {"workload": "served job with space charge: %d macroparticles, 8 smooth "
                               "segments with a 2.5D Gaussian space-charge kick after each "
                               "(700 slices), LinearMap, broadband resonator (700 slices)" % n,
                   "macroparticles": n, "slices": S, "space_charge_kicks": 8},
        "element_by_element": {"value": n / ms_el * 1e3, "ms_per_step": ms_el},
        "speedup_vs_element_by_element": ms_el / ms,
    }
    emit(json.dumps(out))


def serve_jobs(n_jobs, seed=2000, space_charge=True):
    """Served instability-scan stream: Poisson arrivals at 4 jobs/s;
    macroparticles log-uniform in [1e5, 1e7], slices uniform in [100, 1000],
    turns log-uniform in [256, 4096], intensity and chromaticity as the
    batched scan; the LinearMap ring with the broadband resonator.  Jobs of
    5e6 macroparticles and more (the upper end of the range) also get a
    2.5D transverse Gaussian space-charge kick after each of 8 transverse
    segments (their own slicing, the job's slice count), run by the fused
    engine's space-charge stages.  Host-side bunches, generated up front."""
    import numpy as np
    from scipy.constants import e, m_p
    from pyheadtail_amd.machines.synchrotron import Synchrotron
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    for j in range(n_jobs):
        t += rng.exponential(1 / 4.0)
        n = int(np.exp(rng.uniform(np.log(1e5), np.log(1e7))))
        S = int(rng.integers(100, 1001))
        turns = int(np.exp(rng.uniform(np.log(256), np.log(4096))))
        intensity, qp = rng.uniform(1e10, 3e11), rng.uniform(-5., 15.)
        sc = space_charge and n >= 5e6
        if sc:
            mach = sc_machine(qp, S)
        else:
            mach = Synchrotron(optics_mode='smooth', charge=e, mass=m_p, p0=p0_of_gamma(27.7),
                               circumference=6911.5, n_segments=1, beta_x=54.6, beta_y=54.6,
                               D_x=0., D_y=0., accQ_x=20.13, accQ_y=20.18, Qp_x=qp, Qp_y=qp,
                               alpha_mom_compaction=1.9e-3, longitudinal_mode='linear',
                               Q_s=0.017)
        np.random.seed(seed + j)
        b = mach.generate_6D_Gaussian_bunch(n, intensity, EPSN, EPSN, SIGMA_Z)
        host = {k: getattr(b, k).cpu().numpy() for k in ('x', 'xp', 'y', 'yp', 'z', 'dp')}
        print("[serve] job %d generated: %d macroparticles, %d slices, %d turns%s"
              % (j, n, S, turns, ", space charge" if sc else ""), file=sys.stderr, flush=True)
        out.append(dict(t=t, mach=mach, S=S, turns=turns, n=n, host=host, sc=sc,
                        params=(b.particlenumber_per_mp, b.charge, b.mass, b.circumference,
                                b.gamma)))
        del b
    return out


def serve(args):
    """Served stream: p50 / p99 time-to-result and aggregate throughput."""
    import numpy as np
    import torch
    torch.cuda.set_device(0)
    from pyheadtail_amd.serve import ScanServer
    from pyheadtail_amd.particles.particles import Particles
    from pyheadtail_amd.particles.slicing import UniformBinSlicer
    from pyheadtail_amd.impedances.wakes import WakeField, CircularResonator
    jobs = serve_jobs(args.jobs, space_charge=not args.no_sc)
    srv = ScanServer(n_streams=args.streams, chunk_turns=64, graph_turns=8,
                     big_job_bytes=int(args.big_job_mib * (1 << 20)))
    quarter = {}
    t0 = time.perf_counter()
    i = 0
    while i < len(jobs) or srv.busy:
        now = time.perf_counter() - t0
        while i < len(jobs) and jobs[i]['t'] <= now:
            jd = jobs[i]
            ppm, q, m, C, g = jd['params']
            bunch = Particles(jd['n'], ppm, q, m, C, g, jd['host'])
            wake = WakeField(UniformBinSlicer(jd['S'], n_sigma_z=3), CircularResonator(1e6, 1e9, 1))
            srv.submit(list(jd['mach'].one_turn_map) + [wake], bunch, jd['turns'], job_id=i)
            jd['host'] = None
            i += 1
        if srv.busy:
            for j in srv.active:
                q = 4 * (j.n_turns - j.remaining) // j.n_turns
                if 0 < q < 4 and quarter.get(j.job_id, 0) < q:
                    quarter[j.job_id] = q
                    print("[serve] %.1f s: job %d at %d/4 of its turns"
                          % (time.perf_counter() - t0, j.job_id, q), file=sys.stderr, flush=True)
            for f in srv.step():
                print("[serve] %.1f s: job %d done after %.2f s (%d left)"
                      % (time.perf_counter() - t0, f.job_id, f.time_to_result,
                         len(jobs) - len(srv.done)), file=sys.stderr, flush=True)
        elif i < len(jobs):
            time.sleep(max(0.0, jobs[i]['t'] - (time.perf_counter() - t0)))
    makespan = time.perf_counter() - t0
    work = sum(j.bunch.macroparticlenumber * j.n_turns for j in srv.done)
    p = srv.latency_percentiles((50, 99))
    out = {
        "metric": "served scan: p99 time-to-result",
        "value": p[99], "unit": "s", "n_gpus": 1, "steps": len(srv.done), "warmup": 0,
        "ms_per_step": makespan / max(1, len(srv.done)) * 1e3, "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded Gaussian bunches, Poisson arrivals at 4 jobs/s)",
        "config": {"workload": "%d jobs: macroparticles log-uniform in [1e5, 1e7], slices in "
                               "[100, 1000], turns log-uniform in [256, 4096]; LinearMap ring "
                               "with a broadband resonator; jobs of >= 5e6 macroparticles with "
                               "8 space-charge kicks per turn" % len(jobs),
                   "jobs": len(jobs), "space_charge_jobs": sum(j['sc'] for j in jobs),
                   "streams": args.streams},
        "p50_s": p[50], "p99_s": p[99],
        "aggregate_mp_turns_per_s": work / makespan, "makespan_s": makespan,
    }
    emit(json.dumps(out))


def headline_turns(n, args, steps, stage_turns):
    """The fused one-turn map on an n-macroparticle bunch: warm-up (graph
    capture), per-stage kernel times, then `steps` timed turns.  Returns the
    timed seconds, the stage times and the host-to-device copy time."""
    import numpy as np
    import torch
    from pyheadtail_amd.engine import FusedOneTurnMap
    from pyheadtail_amd.particles.slicing import UniformBinSlicer
    from pyheadtail_amd.impedances.wakes import WakeField, CircularResonator

    mach = build_machine(p0_of_gamma(27.7))
    wake = WakeField(UniformBinSlicer(args.slices, n_sigma_z=3), CircularResonator(1e6, 1e9, 1))
    one_turn = list(mach.one_turn_map) + [wake]
    # the Gaussian bunch cut at the RF separatrix, drawn with NumPy's legacy
    # global RNG as PyHEADTAIL does (machines/synchrotron.py:309-360)
    np.random.seed(SEED)
    bunch = mach.generate_6D_Gaussian_bunch(n, INTENSITY, EPSN, EPSN, SIGMA_Z)
    # the one host-to-device copy of the bunch (pageable NumPy arrays, as the
    # Particles constructor receives them); not part of the timed region
    host_coords = {k: getattr(bunch, k).cpu().numpy() for k in
                   ('x', 'xp', 'y', 'yp', 'z', 'dp')}
    torch.cuda.synchronize()
    t_h2d = time.perf_counter()
    bunch.update(host_coords)
    torch.cuda.synchronize()
    h2d_ms = (time.perf_counter() - t_h2d) * 1e3
    del host_coords
    eng = FusedOneTurnMap(one_turn, graph_turns=args.graph_turns)
    # buffers for the longest call, so that no later call re-allocates them
    # or captures its graph again
    eng.reserve(bunch, max(steps, stage_turns, args.warmup, args.graph_turns))
    # warm-up (the graph is captured here)
    eng.track_turns(bunch, max(args.warmup, args.graph_turns))
    bunch.x  # applies the pending kick
    torch.cuda.synchronize()
    # per-stage kernel durations with HIP events on the launch stream
    stage_ms = eng.timed_turns(bunch, stage_turns)
    bunch.x
    torch.cuda.synchronize()
    # timed region: `steps` turns, graph-launched, then the final kick
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    eng.track_turns(bunch, steps)
    bunch.x
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    return dt, stage_ms, h2d_ms


def roofline(n, kernel_ms, kernel):
    achieved = BYTES_PER_MP_PASS * n / (kernel_ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": pass_traffic(n), "kernel": kernel,
            "algorithmic_bytes_per_launch": BYTES_PER_MP_PASS * n}


# The one result line goes to the process's standard output; everything
# else (the elements' warnings and notes, which the reference's classes
# print) goes to standard error while the workloads run.
_RESULT_OUT = sys.stdout


def emit(line):
    print(line, file=_RESULT_OUT, flush=True)


def main():
    sys.stdout = sys.stderr
    args = parse()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(independent bunches would shard across GPUs without communication). "
              "Use --gpus 1.", file=sys.stderr)
        sys.exit(2)
    if args.traffic_from:
        record_traffic(args.traffic_from, args.n)
        return
    if args.copy_calibration:
        copy_calibration(args.n)
        return
    if args.workload == "rw64":
        rw64(args)
        return
    if args.workload == "scan64":
        scan64(args)
        return
    if args.workload == "serve":
        serve(args)
        return
    if args.workload == "spacecharge":
        spacecharge(args)
        return
    cpu = cpu_np = None
    if not args.no_cpu:
        cpu = cpu_baseline(args.cpu_n, args.cpu_seconds)
        cpu_np = cpu_baseline_numpy(args.cpu_n, args.cpu_turns)

    import torch
    torch.cuda.set_device(0)
    dt, stage_ms, h2d_ms = headline_turns(args.n, args, args.steps, args.stage_turns)
    value = args.n * args.steps / dt
    turn_ms = dt / args.steps * 1e3        # a graph-replayed turn
    out = {
        "metric": "macroparticle-turns/s @1e6 MP+wake; %HBM roofline; batched-scan p99 latency",
        "value": value,
        "unit": "macroparticle-turns/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded 6D Gaussian bunch, separatrix-cut)",
        "config": {"workload": "full plain one-turn map, FP64: %d macroparticles, RFSystems "
                               "h=4620 V=4.5MV, transverse Q'=5 + amplitude detuning, %d slices "
                               "+-3 sigma_z, broadband resonator wake (1 turn)"
                               % (args.n, args.slices),
                   "macroparticles": args.n, "n_slices": args.slices, "turns": args.steps,
                   "graph_turns": args.graph_turns},
        "roofline": roofline(args.n, stage_ms[0], "k_engine_pass (fused particle pass)"),
        "turn_roofline_frac": BYTES_PER_MP_PASS * args.n / (turn_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "stage_ms": {"pass": stage_ms[0], "hist": stage_ms[1], "conv": stage_ms[2]},
        "h2d_bunch_ms": h2d_ms,
        "cpu_baseline": cpu,
        "cpu_baseline_numpy_1core": cpu_np,
    }
    if args.n_large > 0:
        # the same map on a bunch ten times larger than the last-level cache
        steps_l = min(args.steps, args.steps_large)
        dt_l, st_l, _ = headline_turns(args.n_large, args, steps_l, min(args.stage_turns, 50))
        out["large"] = {
            "macroparticles": args.n_large, "turns": steps_l,
            "value": args.n_large * steps_l / dt_l, "ms_per_step": dt_l / steps_l * 1e3,
            "stage_ms": {"pass": st_l[0], "hist": st_l[1], "conv": st_l[2]},
            "roofline": roofline(args.n_large, st_l[0], "k_engine_pass (fused particle pass)"),
            "turn_roofline_frac": (BYTES_PER_MP_PASS * args.n_large / (dt_l / steps_l) / 1e9 /
                                   HBM_PEAK_GBS),
        }
    emit(json.dumps(out))


if __name__ == "__main__":
    main()
