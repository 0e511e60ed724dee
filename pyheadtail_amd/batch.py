"""Independent bunches advanced together on one device.

Each job is a one-turn map and its bunch, run by its own FusedOneTurnMap
(own buffers, wake history and HIP graphs).  BatchedScan advances all jobs
in chunks of turns: each job's chunk is replayed from its graph on one of
n_streams HIP streams (default priority; each first waits for the caller's
stream, which made the bunches), so the short, latency-bound
per-turn kernels of different bunches overlap on the device instead of
running one bunch after another.  Jobs of at least big_job_bytes share
the first stream (two of them interleaved evict each other's coordinates
from the last-level cache); the others take the remaining streams in turn.
Jobs share no device state, so every job's result is the one it gets alone,
bit for bit.
"""
import torch

from .engine import FusedOneTurnMap
from .serve import job_bytes


class BatchedScan(object):

    def __init__(self, jobs, n_streams=8, chunk_turns=100, graph_turns=10, record_stats=False,
                 big_job_bytes=128 << 20):
        """jobs: list of (one_turn_map, bunch); the one-turn maps must be
        distinct objects (each keeps its wake history)."""
        self.bunches = [b for _, b in jobs]
        self.engines = [FusedOneTurnMap(m, graph_turns=graph_turns, record_stats=record_stats)
                        for m, _ in jobs]
        self.streams = [torch.cuda.Stream() for _ in range(max(1, n_streams))]
        self.chunk_turns = int(chunk_turns)
        ns, k = len(self.streams), 0
        self.job_stream = []
        for m, b in jobs:
            if ns == 1 or job_bytes(m, b) >= big_job_bytes:
                self.job_stream.append(self.streams[0])
            else:
                self.job_stream.append(self.streams[1 + k % (ns - 1)])
                k += 1

    def track_turns(self, n_turns):
        """Advance every job by n_turns turns; returns when all are done."""
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
