"""Continuous batching of instability-scan jobs on one device.

A job is a one-turn map, its bunch (device-resident from submission) and a
number of turns.  Jobs wait in arrival order until the HBM budget admits
them; an admitted job then advances by chunks of turns, each chunk enqueued
as soon as the job's previous one has completed, the jobs spread over a few
HIP streams so that their kernels overlap, and a job leaves as soon as its
turns are done, releasing its share of the budget to
the queue.  A one-turn map the fused engine takes (space-charge kicks
included) runs fused; one it does not take runs element by element, in
smaller chunks.  Jobs share no device state: each result is the one the job gets
alone, bit for bit.
"""
import collections
import time

import numpy as np
import torch

from . import _phx
from .engine import FusedOneTurnMap


class ElementLoop(object):
    """Turns of a one-turn map element by element on the current stream
    (the elements' own device kernels)."""

    fused = False

    def __init__(self, one_turn_map):
        self.one_turn_map = list(one_turn_map)

    def track_turns(self, beam, n_turns):
        for _ in range(n_turns):
            for el in self.one_turn_map:
                el.track(beam)


def make_runner(one_turn_map, graph_turns, fuse=True):
    """The fused engine for one_turn_map if it takes every element (and
    fuse is set), else an element loop."""
    if not fuse:
        return ElementLoop(one_turn_map)
    try:
        eng = FusedOneTurnMap(one_turn_map, graph_turns=graph_turns)
    except (TypeError, ValueError):
        return ElementLoop(one_turn_map)
    eng.fused = True
    return eng


class Job(object):
    __slots__ = ('job_id', 'one_turn_map', 'bunch', 'n_turns', 'remaining', 'bytes', 'engine',
                 'stream', 'event', 't_submit', 't_admit', 't_done', 'fuse')

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
    from .spacecharge.spacecharge import TransverseGaussianSpaceCharge
    S_sc = max([int(el.slicer.n_slices) for el in one_turn_map
                if isinstance(el, TransverseGaussianSpaceCharge)] + [0])
    if S_sc > 0:
        ws += max(0, _phx.lib().phx_engine_sc_ws_bytes(n, S_sc))
    return 6 * 8 * n + 4 * n + ws + (S + 1) * 8 * (3 + 4 * T + nk)


class ScanServer(object):

    def __init__(self, hbm_budget_bytes=None, n_streams=8, chunk_turns=64, graph_turns=8,
                 element_chunk_turns=8, pause=2e-4, big_job_bytes=128 << 20):
        if hbm_budget_bytes is None:
            hbm_budget_bytes = int(0.8 * torch.cuda.get_device_properties(0).total_memory)
        self.budget = int(hbm_budget_bytes)
        self.used = 0
        self.streams = [torch.cuda.Stream() for _ in range(max(1, n_streams))]
        self.chunk_turns, self.graph_turns = int(chunk_turns), int(graph_turns)
        self.element_chunk_turns = int(element_chunk_turns)
        self.pause = float(pause)
        self.big_job_bytes = int(big_job_bytes)
        self.queue = collections.deque()
        self.active = []
        self.done = []
        self._next_id = 0
        self._slot = 0

    @property
    def busy(self):
        return bool(self.queue or self.active)

    def submit(self, one_turn_map, bunch, n_turns, job_id=None, fuse=True):
        """Queue a job; bunch must already be on the device.  fuse=False
        runs it element by element even where the fused engine would take
        its map."""
        j = Job()
        j.fuse = bool(fuse)
        j.job_id = self._next_id if job_id is None else job_id
        self._next_id += 1
        j.one_turn_map, j.bunch, j.n_turns = one_turn_map, bunch, int(n_turns)
        j.remaining = j.n_turns
        j.bytes = job_bytes(one_turn_map, bunch)
        if j.bytes > self.budget:
            raise ValueError("job %r needs %d bytes, more than the budget %d"
                             % (j.job_id, j.bytes, self.budget))
        j.engine = j.stream = j.event = j.t_admit = j.t_done = None
        j.t_submit = time.perf_counter()
        self.queue.append(j)
        return j

    def _stream_for(self, j):
        """Jobs of at least big_job_bytes share the first stream, so that two
        of them never interleave: one such job's coordinates stay in the
        device's last-level cache across the kernels of its turns, two
        interleaved ones evict each other (two space-charge jobs of 6e6
        macroparticles: 6.5 ms per job-turn one after the other, 30-50 ms
        interleaved).  The others take the remaining streams in turn."""
        if len(self.streams) == 1 or j.bytes >= self.big_job_bytes:
            return self.streams[0]
        st = self.streams[1 + self._slot % (len(self.streams) - 1)]
        self._slot += 1
        return st

    def _admit(self):
        caller = torch.cuda.current_stream()
        while self.queue and self.used + self.queue[0].bytes <= self.budget:
            j = self.queue.popleft()
            j.engine = make_runner(j.one_turn_map, self.graph_turns, j.fuse)
            j.stream = self._stream_for(j)
            j.stream.wait_stream(caller)      # the bunch was uploaded on the caller's stream
            j.t_admit = time.perf_counter()
            self.used += j.bytes
            self.active.append(j)

    def step(self):
        """Admit what the budget allows; enqueue the next chunk of every
        active job whose previous chunk has completed; retire the jobs whose
        last chunk has completed and return them.  A slow job holds no other
        job back: when no job can move, the host sleeps for pause seconds
        and returns (waiting on one job's chunk would stall the jobs whose
        chunks end first)."""
        self._admit()
        finished, moved = [], False
        for j in self.active:
            if j.event is not None and not j.event.query():
                continue
            if j.remaining == 0:
                finished.append(j)
                continue
            chunk = self.chunk_turns if j.engine.fused else self.element_chunk_turns
            k = min(chunk, j.remaining)
            with torch.cuda.stream(j.stream):
                j.engine.track_turns(j.bunch, k)
                j.remaining -= k
                if j.remaining == 0 and j.bunch._pending is j.engine:
                    j.bunch._run_pending()
                j.event = torch.cuda.Event()
                j.event.record(j.stream)
            moved = True
        if finished:
            now = time.perf_counter()
            caller = torch.cuda.current_stream()
            for j in finished:
                caller.wait_stream(j.stream)
                j.t_done = now
                self.used -= j.bytes
                j.engine = j.event = None         # its buffers go back to the allocator
            gone = set(map(id, finished))
            self.active = [j for j in self.active if id(j) not in gone]
            self.done.extend(finished)
        elif not moved and self.active:
            time.sleep(self.pause)
        return finished

    def drain(self):
        while self.busy:
            self.step()

    def latency_percentiles(self, q=(50, 99)):
        t = np.array([j.time_to_result for j in self.done])
        return {p: float(np.percentile(t, p)) for p in q} if len(t) else {}
