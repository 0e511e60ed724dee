"""phx_engine_turns_noalloc, the engine call without device allocation (its
workspace is the library's static scratch), called through the C ABI: two
calls in a row, the second on a smaller bunch, each bit-identical to
phx_engine_turns with a caller-allocated workspace followed by
phx_engine_flush; and an engine beyond the scratch is refused."""
import ctypes as C

import numpy as np
import pytest
import torch

from oracle import workloads as W
from tests.machines import COORDS, nonlinear_machine, to_product

pytestmark = pytest.mark.gpu


def _with_workspace(ob, n_turns, n_turns_wake):
    from pyheadtail_amd.engine import FusedOneTurnMap
    pb = to_product(ob)
    eng = FusedOneTurnMap(nonlinear_machine(n_turns_wake=n_turns_wake), graph_turns=0)
    eng.track_turns(pb, n_turns, use_graph=False)
    return {k: getattr(pb, k).cpu().numpy() for k in COORDS}   # reading flushes


def _noalloc(ob, n_turns, n_turns_wake):
    from pyheadtail_amd import _phx
    from pyheadtail_amd.engine import FusedOneTurnMap
    L = _phx.lib()
    pb = to_product(ob)
    eng = FusedOneTurnMap(nonlinear_machine(n_turns_wake=n_turns_wake), graph_turns=0)
    rec = eng._prepare(pb, n_turns)     # programs, turn table and state buffers
    rec.ws = None                       # the call brings its own workspace
    _phx.call(L.phx_engine_turns_noalloc, C.byref(rec), n_turns,
              _phx.NOALLOC_RESET | _phx.NOALLOC_FLUSH, _phx.stream())
    torch.cuda.synchronize()
    return {k: pb._raw(k).cpu().numpy() for k in COORDS}


def test_noalloc_equals_workspace_call_twice():
    print("\n[case] noalloc engine: 2e5 then 1e5 macroparticles, 7 and 5 turns", flush=True)
    for n, turns, T in ((200000, 7, 2), (100000, 5, 1)):
        ob = W.nonlinear_bunch(n, seed=n // 1000)
        want = _with_workspace(ob, turns, T)
        got = _noalloc(ob, turns, T)
        for k in COORDS:
            np.testing.assert_array_equal(got[k], want[k], err_msg="n=%d %s" % (n, k))
        print("[info] n=%d bit-identical" % n, flush=True)


def test_noalloc_rejects_engine_beyond_scratch():
    print("\n[case] noalloc engine: refused beyond the scratch size", flush=True)
    from pyheadtail_amd import _phx
    from pyheadtail_amd.engine import FusedOneTurnMap
    from pyheadtail_amd.particles.slicing import UniformBinSlicer
    from pyheadtail_amd.impedances.wakes import WakeField, ParallelHorizontalPlatesResistiveWall
    L = _phx.lib()
    # 1000 slices x 4 kicks x 800 turns of memory: 25.6 MB of convolution
    # partials alone
    m = nonlinear_machine(wake=False) + [WakeField(
        UniformBinSlicer(1000, n_sigma_z=3),
        ParallelHorizontalPlatesResistiveWall(0.02, W.CIRCUMFERENCE, 5.9e7, 1e-12,
                                              n_turns_wake=800))]
    assert L.phx_engine_ws_bytes(20000, 1000, 4, 800) > _phx.NOALLOC_SCRATCH_BYTES
    pb = to_product(W.nonlinear_bunch(20000, seed=3))
    eng = FusedOneTurnMap(m, graph_turns=0)
    rec = eng._prepare(pb, 1)
    rec.ws = None
    rc = L.phx_engine_turns_noalloc(C.byref(rec), 1, 0, _phx.stream())
    assert rc == -2
    assert b"scratch" in L.phx_last_error()
