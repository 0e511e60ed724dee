"""Launch-shape autotuner for batched scans.

BatchedScan (fused mode) has these launch-shape settings:
  pass_grid         blocks of each job's particle pass (0: as many as stay
                    resident on the device, else at most this many); fewer
                    blocks per job mean fewer, fuller blocks in the batch's
                    pass launch.  The bunch sums are folded per block, so
                    this sets the last bits of every job's cuts: a job run
                    alone with the same pass grid gives the same bits.
  groups_per_block  instead of one pass grid for all jobs, each job's grid
                    sized so that its blocks walk about this many particle
                    groups (BatchedScan.grid_for), so that jobs of different
                    sizes finish their pass together.
  graph_turns       turns per replayed HIP graph (no effect on results).

tune_batch() times every candidate on the batch itself (HIP events around
`turns` turns after a warm-up) and keeps the fastest.  Results are
stored in a JSON file keyed by the device name and a signature of the batch
(each job's macroparticle count rounded to 2 significant digits, slice
count, wake memory and map variant), and reused when the same signature
comes again.  The tuned batch has already advanced its jobs by the turns
the measurement ran.
"""
import json
import os

import torch

DEFAULT_PATH = os.path.join(os.path.expanduser("~"), ".cache", "pyheadtail_amd", "tune.json")
PASS_GRIDS = (0, 1024, 512, 256, 128)
GROUPS_PER_BLOCK = (4, 8, 16, 32, 64)
GRAPH_TURNS = (10, 25)


def _round2(n):
    if n <= 0:
        return 0
    s = '%.1e' % n
    return int(float(s))


def signature(scan):
    jobs = []
    for eng, b in zip(scan.engines, scan.bunches):
        w = eng.wake
        jobs.append((_round2(b.macroparticlenumber), int(w.slicer.n_slices) if w else 0,
                     int(w.n_turns_wake_max) if w else 0, [type(e).__name__ for e in eng.local]))
    return json.dumps(sorted(jobs, key=str))


class Tuner(object):
    """Measured launch settings, persisted in a JSON file."""

    def __init__(self, path=None):
        self.path = path or os.environ.get("PHX_TUNE_FILE", DEFAULT_PATH)
        try:
            with open(self.path) as f:
                self.table = json.load(f)
        except (OSError, ValueError):
            self.table = {}

    def key(self, scan):
        dev = torch.cuda.get_device_name(torch.cuda.current_device())
        return dev + "|" + signature(scan)

    def lookup(self, scan):
        return self.table.get(self.key(scan))

    def store(self, scan, entry):
        self.table[self.key(scan)] = entry
        d = os.path.dirname(self.path)
        if d:
            os.makedirs(d, exist_ok=True)
        tmp = self.path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(self.table, f, indent=1, sort_keys=True)
        os.replace(tmp, self.path)


def _time_turns(scan, turns):
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start.record()
    scan.track_turns(turns)
    end.record()
    torch.cuda.synchronize()
    return start.elapsed_time(end) / turns


def tune_batch(scan, tuner=None, turns=20, pass_grids=PASS_GRIDS,
               groups_per_block=GROUPS_PER_BLOCK, graph_turns=GRAPH_TURNS, force=False):
    """Pick the pass grids (one pass_grid for all jobs, or groups_per_block)
    and graph_turns for a fused-mode BatchedScan by measurement (or take
    the stored choice), configure the scan with it and return the entry:
    {'pass_grid', 'groups_per_block', 'graph_turns', 'ms_per_turn',
    'candidates': [[pass_grid, groups_per_block, graph_turns, ms_per_turn],
    ...]} (groups_per_block 0: pass_grid for every job)."""
    tuner = tuner or Tuner()
    entry = None if force else tuner.lookup(scan)
    if entry is None:
        shapes = [(pg, 0) for pg in pass_grids] + [(0, k) for k in groups_per_block]
        cands = []
        for pg, k in shapes:
            for gt in graph_turns:
                scan.configure(pass_grid=pg, groups_per_block=k, graph_turns=gt)
                scan.track_turns(gt)          # batch built, graph captured
                cands.append([pg, k, gt, _time_turns(scan, max(turns, gt))])
        best = min(cands, key=lambda c: c[3])
        entry = {'pass_grid': best[0], 'groups_per_block': best[1], 'graph_turns': best[2],
                 'ms_per_turn': best[3], 'candidates': cands}
        tuner.store(scan, entry)
    scan.configure(pass_grid=entry['pass_grid'], groups_per_block=entry.get('groups_per_block', 0),
                   graph_turns=entry['graph_turns'])
    return entry
