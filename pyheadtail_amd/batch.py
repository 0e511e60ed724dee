"""Independent bunches advanced together on one device.

Each job is a one-turn map and its bunch, run by its own FusedOneTurnMap
(own buffers and wake history).  BatchedScan advances all jobs together in
one of two ways:

  mode='fused' (default): every stage of a turn is one kernel launch for all
      jobs (phx_batch_*: each launch's blocks are the concatenation of every
      job's blocks of that stage), a graph of graph_turns turns replayed;
      jobs whose maps compile to different pass variants (element order,
      detuner set, statistics) form separate groups, one batch each;
  mode='streams': each job's own graphs on one of n_streams HIP streams
      (default priority), chunk_turns turns at a time, jobs of at least
      big_job_bytes together on the first stream (two of them interleaved
      evict each other's coordinates from the last-level cache).

Either way every job's arithmetic, reduction order included, is the one it
has alone with the same pass grid: results are bit-identical to running
the job by itself.  The pass grid (blocks of the particle pass per job;
0: as many as stay resident) sets the bunch sums' fold order.  It is either
one value for every job (pass_grid) or sized per job so that every block
walks about the same number of particle groups (groups_per_block: jobs of
different sizes then end their pass together instead of the largest
job's blocks running on alone).  The autotuner (tune.py) picks between
these, and the number of turns per graph, by measurement.
"""
import ctypes as C

import torch

from . import _phx
from .engine import FusedOneTurnMap
from .serve import job_bytes


class BatchedScan(object):

    def __init__(self, jobs, n_streams=8, chunk_turns=100, graph_turns=10, record_stats=False,
                 big_job_bytes=128 << 20, mode='fused', pass_grid=0, groups_per_block=0):
        """jobs: list of (one_turn_map, bunch); the one-turn maps must be
        distinct objects (each keeps its wake history)."""
        if mode not in ('fused', 'streams'):
            raise ValueError("mode must be 'fused' or 'streams'")
        self.mode = mode
        self.bunches = [b for _, b in jobs]
        self.engines = [FusedOneTurnMap(m, graph_turns=graph_turns, record_stats=record_stats,
                                        pass_grid=pass_grid) for m, _ in jobs]
        self.graph_turns = int(graph_turns)
        self.chunk_turns = int(chunk_turns)
        self._groups = None
        self._pass_grid = int(pass_grid)
        self.groups_per_block = 0
        if groups_per_block:
            self.configure(groups_per_block=groups_per_block)
        if mode == 'streams':
            self.streams = [torch.cuda.Stream() for _ in range(max(1, n_streams))]
            ns, k = len(self.streams), 0
            self.job_stream = []
            for m, b in jobs:
                if ns == 1 or job_bytes(m, b) >= big_job_bytes:
                    self.job_stream.append(self.streams[0])
                else:
                    self.job_stream.append(self.streams[1 + k % (ns - 1)])
                    k += 1

    # -- settings the tuner changes ------------------------------------------
    @property
    def pass_grid(self):
        """The pass grid of every job (0: as many blocks as stay resident);
        unused while groups_per_block is set."""
        return self._pass_grid

    @staticmethod
    def grid_for(n, groups_per_block):
        """The pass grid that gives each block about groups_per_block of an
        n-macroparticle bunch's particle groups."""
        group = _phx.lib().phx_pass_group_size()
        groups = -(-int(n) // group)
        return max(1, min(_phx.MAX_GRID, -(-groups // int(groups_per_block))))

    def configure(self, pass_grid=None, graph_turns=None, groups_per_block=None):
        """Change the pass grids (one value for every job, or sized per job
        by groups_per_block > 0) or the turns per graph; the batches and
        graphs are built again at the next call."""
        if pass_grid is not None:
            self._pass_grid = int(pass_grid)
        if groups_per_block is not None:
            self.groups_per_block = int(groups_per_block)
        if pass_grid is not None or groups_per_block is not None:
            for eng, b in zip(self.engines, self.bunches):
                eng.pass_grid = (self.grid_for(b.macroparticlenumber, self.groups_per_block)
                                 if self.groups_per_block > 0 else self._pass_grid)
        if graph_turns is not None:
            self.graph_turns = int(graph_turns)
            for eng in self.engines:
                eng.graph_turns = int(graph_turns)
        self._drop_batches()

    # -- fused mode ----------------------------------------------------------
    def _drop_batches(self):
        if self._groups:
            L = _phx.load_library()
            for grp in self._groups.values():
                for h in grp['graphs'].values():
                    L.phx_graph_destroy(h)
                if grp['batch'] is not None:
                    L.phx_batch_destroy(grp['batch'])
        self._groups = {}

    def __del__(self):
        try:
            self._drop_batches()
        except Exception:
            pass

    def _fused_turns(self, n_turns):
        L = _phx.lib()
        s = _phx.stream()
        for eng, b in zip(self.engines, self.bunches):
            eng.reserve(b, max(n_turns, self.graph_turns))
        recs = [eng._prepare(b, n_turns) for eng, b in zip(self.engines, self.bunches)]
        groups = {}
        for i, r in enumerate(recs):
            groups.setdefault((r.record_stats, r.program_shape, r.det_shape), []).append(i)
        if self._groups is None:
            self._groups = {}
        for key, idx in groups.items():
            arr = (_phx.Engine * len(idx))(*[recs[i] for i in idx])
            sig = bytes(memoryview(arr).cast("B"))
            grp = self._groups.get(key)
            if grp is None or grp['sig'] != sig:
                if grp is not None:
                    for h in grp['graphs'].values():
                        L.phx_graph_destroy(h)
                    L.phx_batch_destroy(grp['batch'])
                nb = L.phx_batch_desc_bytes(len(idx))
                desc = torch.empty(nb, dtype=torch.uint8, device=_phx.device())
                h = C.c_void_p()
                _phx.call(L.phx_batch_create, arr, len(idx), desc.data_ptr(), nb, s, C.byref(h))
                grp = dict(sig=sig, batch=h, desc=desc, arr=arr, graphs={})
                self._groups[key] = grp
            G = self.graph_turns
            done = 0
            if G > 0 and n_turns >= G:
                g = grp['graphs'].get(G)
                if g is None:
                    g = C.c_void_p()
                    _phx.call(L.phx_batch_capture, grp['batch'], G, C.byref(g))
                    grp['graphs'][G] = g
                _phx.call(L.phx_batch_snapshot, grp['batch'], s)
                while n_turns - done >= G:
                    _phx.call(L.phx_graph_launch, g, s)
                    done += G
            if n_turns > done:
                _phx.call(L.phx_batch_turns, grp['batch'], n_turns - done, s)
        for eng, b in zip(self.engines, self.bunches):
            eng._finish(b)

    # -- both modes ----------------------------------------------------------
    def track_turns(self, n_turns):
        """Advance every job by n_turns turns; returns when all are done."""
        if n_turns <= 0:
            return
        if self.mode == 'fused':
            self._fused_turns(n_turns)
            for eng, b in zip(self.engines, self.bunches):
                if b._pending is eng:
                    b._run_pending()
            torch.cuda.current_stream().synchronize()
            return
        caller = torch.cuda.current_stream()
        for s in self.streams:           # the bunches were made on the caller's stream
            s.wait_stream(caller)
        done = 0
        while done < n_turns:
            k = min(self.chunk_turns, n_turns - done)
            for i, (eng, b) in enumerate(zip(self.engines, self.bunches)):
                with torch.cuda.stream(self.job_stream[i]):
                    eng.track_turns(b, k)
            done += k
        for i, (eng, b) in enumerate(zip(self.engines, self.bunches)):
            with torch.cuda.stream(self.job_stream[i]):
                if b._pending is eng:
                    b._run_pending()
        for s in self.streams:
            caller.wait_stream(s)
            s.synchronize()

    def track_turns_job_by_job(self, n_turns):
        """The reference point: each job runs all its turns alone, one job
        after the other, on the current stream."""
        for eng, b in zip(self.engines, self.bunches):
            eng.track_turns(b, n_turns)
            if b._pending is eng:
                b._run_pending()
        torch.cuda.current_stream().synchronize()

    @property
    def macroparticles(self):
        return sum(b.macroparticlenumber for b in self.bunches)
