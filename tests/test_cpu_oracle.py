"""The multi-core C++ restatement (oracle/cpu_turn.cpp, the all-core CPU
baseline) against the NumPy oracle, which the reference's golden vectors
pin: coordinates within 1e-12 of their rms after 20 turns of the
benchmark's one-turn map, on 1 and on several threads."""
import os
import subprocess

import numpy as np
import pytest

from oracle import workloads as W
from tests.machines import coord_err

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def cpu_turn():
    from oracle import cpu_turn as CT
    if not os.path.exists(CT.LIB):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True)
    return CT


@pytest.mark.parametrize("threads", [1, 4])
def test_cpu_restatement_matches_numpy_oracle(cpu_turn, threads):
    n, turns = 20000, 20
    ob = W.nonlinear_bunch(n, seed=31)
    ct = cpu_turn.CpuTurns(ob, n_threads=threads)
    m = W.nonlinear_map_oracle()
    for _ in range(turns):
        for el in m:
            el.track(ob)
    ct.track_turns(turns)
    res = ct.result()
    assert coord_err(res, ob) <= 1e-12


def test_cpu_restatement_kick_is_applied(cpu_turn):
    """A planted slip (the wake kick left out) fails the comparison: the
    check is sensitive to the wake part of the turn."""
    n = 20000
    ob = W.nonlinear_bunch(n, seed=32)
    ct = cpu_turn.CpuTurns(ob, n_threads=2)
    m = W.nonlinear_map_oracle()
    for _ in range(3):
        for el in m:
            el.track(ob)
    ct.track_turns(3)
    ct.state.pending = 0          # drop the last turn's kick
    res = {k: a.copy() for k, a in zip(('x', 'xp', 'y', 'yp', 'z', 'dp'), ct.coords)}
    err = max(np.max(np.abs(res[k] - getattr(ob, k))) / np.sqrt(np.mean(getattr(ob, k)**2))
              for k in ('xp', 'yp'))
    assert err > 1e-12
