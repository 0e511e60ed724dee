"""Continuous batching of instability-scan jobs on one device.

A job is a one-turn map, its bunch (device-resident from submission) and a
number of turns.  Jobs wait in arrival order until the HBM budget admits
them; every step then advances each admitted job by a chunk of turns, the
jobs spread over a few HIP streams so that their kernels overlap, and a job
leaves as soon as its turns are done, releasing its share of the budget to
the queue.  Jobs share no device state: each result is the one the job gets
alone, bit for bit.
"""
import collections
import time

import numpy as np
import torch

from . import _phx
from .engine import FusedOneTurnMap


class Job(object):
    __slots__ = ('job_id', 'one_turn_map', 'bunch', 'n_turns', 'remaining', 'bytes', 'engine',
                 'stream', 't_submit', 't_admit', 't_done')

    @property
    def time_to_result(self):
        return self.t_done - self.t_submit


def job_bytes(one_turn_map, bunch):
    """Device bytes a job holds: coordinates and ids, and the engine's
    workspace, kick table, histogram and wake history."""
    n = bunch.macroparticlenumber
    wake = one_turn_map[-1] if hasattr(one_turn_map[-1], 'wake_kicks') else None
    S = int(wake.slicer.n_slices) if wake is not None else 0
    nk = len(wake.wake_kicks) if wake is not None else 0
    T = wake.n_turns_wake_max if wake is not None else 1
    ws = max(0, _phx.lib().phx_engine_ws_bytes(n, S, nk, T))
    return 6 * 8 * n + 4 * n + ws + (S + 1) * 8 * (3 + 4 * T + nk)


class ScanServer(object):

    def __init__(self, hbm_budget_bytes=None, n_streams=4, chunk_turns=64, graph_turns=8):
        if hbm_budget_bytes is None:
            hbm_budget_bytes = int(0.8 * torch.cuda.get_device_properties(0).total_memory)
        self.budget = int(hbm_budget_bytes)
        self.used = 0
        self.streams = [torch.cuda.Stream() for _ in range(max(1, n_streams))]
        self.chunk_turns, self.graph_turns = int(chunk_turns), int(graph_turns)
        self.queue = collections.deque()
        self.active = []
        self.done = []
        self._next_id = 0
        self._slot = 0

    @property
    def busy(self):
        return bool(self.queue or self.active)

    def submit(self, one_turn_map, bunch, n_turns, job_id=None):
        """Queue a job; bunch must already be on the device."""
        j = Job()
        j.job_id = self._next_id if job_id is None else job_id
        self._next_id += 1
        j.one_turn_map, j.bunch, j.n_turns = one_turn_map, bunch, int(n_turns)
        j.remaining = j.n_turns
        j.bytes = job_bytes(one_turn_map, bunch)
        if j.bytes > self.budget:
            raise ValueError("job %r needs %d bytes, more than the budget %d"
                             % (j.job_id, j.bytes, self.budget))
        j.engine = j.stream = j.t_admit = j.t_done = None
        j.t_submit = time.perf_counter()
        self.queue.append(j)
        return j

    def _admit(self):
        caller = torch.cuda.current_stream()
        while self.queue and self.used + self.queue[0].bytes <= self.budget:
            j = self.queue.popleft()
            j.engine = FusedOneTurnMap(j.one_turn_map, graph_turns=self.graph_turns)
            j.stream = self.streams[self._slot % len(self.streams)]
            self._slot += 1
            j.stream.wait_stream(caller)      # the bunch was uploaded on the caller's stream
            j.t_admit = time.perf_counter()
            self.used += j.bytes
            self.active.append(j)

    def step(self):
        """Admit what the budget allows, advance every active job by one
        chunk, retire the finished ones; returns them."""
        self._admit()
        if not self.active:
            return []
        for j in self.active:
            k = min(self.chunk_turns, j.remaining)
            with torch.cuda.stream(j.stream):
                j.engine.track_turns(j.bunch, k)
                j.remaining -= k
                if j.remaining == 0 and j.bunch._pending is j.engine:
                    j.bunch._run_pending()
        for s in self.streams:
            s.synchronize()
        now = time.perf_counter()
        finished = [j for j in self.active if j.remaining == 0]
        caller = torch.cuda.current_stream()
        for j in finished:
            caller.wait_stream(j.stream)
            j.t_done = now
            self.used -= j.bytes
            j.engine = None                   # its buffers go back to the allocator
        self.active = [j for j in self.active if j.remaining > 0]
        self.done.extend(finished)
        return finished

    def drain(self):
        while self.busy:
            self.step()

    def latency_percentiles(self, q=(50, 99)):
        t = np.array([j.time_to_result for j in self.done])
        return {p: float(np.percentile(t, p)) for p in q} if len(t) else {}
