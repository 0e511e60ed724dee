"""ctypes binding of libphx.so, the HIP (gfx950) library behind this package.

The records below mirror ``include/phx.h`` field by field; ``check_layouts``
compares every ``sizeof`` with the library's own.  Device buffers are torch
tensors on the current HIP device; every call is enqueued on torch's current
stream, so library kernels and torch operations stay stream-ordered.

There is no fallback: if the library or a GPU is missing, ``lib()`` raises.
This plays the role of the reference's ctypes binding of its Thrust library
(PyHEADTAIL/gpu/thrust_interface.py:5-269).
"""
import ctypes as C
import os

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("PHX_LIB", os.path.join(_HERE, "libphx.so"))

MAX_DETUNERS = 4
MAX_CHROMA = 8
MAX_LONG_OPS = 10
MAX_KICKS = 16
MAX_TERMS = 8
MAX_PROGRAM = 16
MAX_SC_STAGES = 16       # space-charge kicks per turn of the fused engine
STANDALONE_GRID = 768
MAX_GRID = 2048

CUTS_FIXED, CUTS_NSIGMA, CUTS_MINMAX = 0, 1, 2
DET_CHROMA, DET_AMPLITUDE = 1, 2
LONG_DRIFT, LONG_KICK = 1, 2
WAKE_RES_TRANSVERSE, WAKE_RES_LONGITUDINAL, WAKE_RESISTIVE_WALL, WAKE_TABLE = 1, 2, 3, 4
TABLE_TRANS_UR, TABLE_TRANS_LOWBETA, TABLE_LONG_BL, TABLE_LONG = 0, 1, 2, 3


def conv_kind(kick):
    """The convolution kernel a phx_wake_kick record needs (CONV_*)."""
    if kick.dot_product and kick.fn == WAKE_RESISTIVE_WALL:
        return CONV_RW_MFMA if kick.precision in MFMA_PRECISIONS else CONV_RW_DOT
    return CONV_GENERAL
MOM_N, MOM_X, MOM_Y = 0, 1, 2
PLANE_XP, PLANE_YP, PLANE_DP = 0, 1, 2
EL_TSEG, EL_LONG, EL_LINMAP = 1, 2, 3
(SHAPE_GENERIC, SHAPE_RF1_T, SHAPE_RF2_T, SHAPE_T_RF1, SHAPE_T_RF2, SHAPE_LIN_T, SHAPE_T_LIN,
 SHAPE_T, SHAPE_NONE) = range(9)
KICKS_GENERIC, KICKS_DXDY, KICKS_PLATES = 0, 1, 2
DETS_GENERIC, DETS_NONE, DETS_C1, DETS_C1_AMP = 0, 1, 2, 3
DETS_SMOOTH = 8          # flag: every segment smooth (PHX_DETS_SMOOTH)
SHAPE_STATIONARY = 16    # flag on SHAPE_RF1_T: stationary RF (PHX_SHAPE_STATIONARY)
ENGINE_MAX_N = 536870911
CONV_GENERAL, CONV_RW_DOT, CONV_RW_MFMA = 1, 2, 4
(PREC_FP64, PREC_FP32, PREC_FP64_DIRECT, PREC_MFMA_F32, PREC_MFMA_BF16X3,
 PREC_MFMA_BF16X6) = range(6)
MFMA_PRECISIONS = (PREC_MFMA_F32, PREC_MFMA_BF16X3, PREC_MFMA_BF16X6)
STAT_NV = 33
NOALLOC_SCRATCH_BYTES = 24 << 20
NOALLOC_RESET, NOALLOC_FLUSH = 1, 2


class PhxError(RuntimeError):
    """A libphx call returned an error code."""


class Slicer(C.Structure):
    _fields_ = [("n_slices", C.c_int32), ("cuts_mode", C.c_int32),
                ("n_sigma_z", C.c_double), ("z_cut_tail", C.c_double),
                ("z_cut_head", C.c_double)]


class SliceState(C.Structure):
    _fields_ = [("tail", C.c_double), ("head", C.c_double), ("width", C.c_double),
                ("mean_z", C.c_double), ("sigma_z", C.c_double),
                ("absmax", C.c_double * 6), ("n", C.c_int64),
                ("n_slices", C.c_int32), ("turn", C.c_int32),
                ("valid", C.c_int32), ("pad", C.c_int32)]


class Term(C.Structure):
    _fields_ = [("a", C.c_int32), ("b", C.c_int32),
                ("center_a", C.c_void_p), ("center_b", C.c_void_p)]


class Detuner(C.Structure):
    _fields_ = [("type", C.c_int32), ("order", C.c_int32),
                ("cx", C.c_double * MAX_CHROMA), ("cy", C.c_double * MAX_CHROMA),
                ("gx", C.c_double), ("ax2", C.c_double), ("bx", C.c_double),
                ("gy", C.c_double), ("ay2", C.c_double), ("by", C.c_double),
                ("dapp_x", C.c_double), ("dapp_y", C.c_double),
                ("dapp_xy", C.c_double), ("dapp_yx", C.c_double),
                ("p0", C.c_double), ("inv_p0", C.c_double)]


_M = ["00", "01", "10", "11", "22", "23", "32", "33"]


class TSeg(C.Structure):
    _fields_ = ([("I" + m, C.c_double) for m in _M] +
                [("J" + m, C.c_double) for m in _M] +
                [("M" + m, C.c_double) for m in _M] +
                [("dQ_x", C.c_double), ("dQ_y", C.c_double),
                 ("D_x_s0", C.c_double), ("D_x_s1", C.c_double),
                 ("D_y_s0", C.c_double), ("D_y_s1", C.c_double),
                 ("with_dispersion", C.c_int32), ("n_detuners", C.c_int32),
                 ("det", Detuner * MAX_DETUNERS)])


class LongOp(C.Structure):
    _fields_ = [("type", C.c_int32), ("with_dispersion", C.c_int32),
                ("beta_ratio", C.c_double), ("eta0", C.c_double), ("length", C.c_double),
                ("harmonic", C.c_double), ("two_pi", C.c_double),
                ("circumference", C.c_double), ("phi_offset", C.c_double),
                ("phi_lock", C.c_double), ("amplitude", C.c_double),
                ("p_increment", C.c_double), ("p0_before", C.c_double),
                ("p0_after", C.c_double), ("D_x", C.c_double), ("D_y", C.c_double),
                ("inv_circumference", C.c_double), ("inv_p0_after", C.c_double)]


class LongSeq(C.Structure):
    _fields_ = [("n_ops", C.c_int32), ("shrink", C.c_int32),
                ("shrink_factor", C.c_double), ("op", LongOp * MAX_LONG_OPS)]


class LinMap(C.Structure):
    _fields_ = [("cosQ", C.c_double), ("sinQ", C.c_double), ("longfac", C.c_double),
                ("D_x", C.c_double), ("D_y", C.c_double)]


class SCSlices(C.Structure):
    _fields_ = [("n_slices", C.c_int32), ("pad", C.c_int32), ("counts", C.c_void_p),
                ("z_bins", C.c_void_p), ("mean_x", C.c_void_p), ("mean_y", C.c_void_p),
                ("sigma_x", C.c_void_p), ("sigma_y", C.c_void_p)]


class SpaceCharge(C.Structure):
    _fields_ = [("prefactor", C.c_double), ("charge_per_mp", C.c_double),
                ("denom_pre", C.c_double), ("round_pre", C.c_double), ("rtol", C.c_double),
                ("atol", C.c_double), ("sig_check", C.c_int32), ("pad", C.c_int32)]


class Damper(C.Structure):
    _fields_ = [("coef_xp", C.c_double), ("coef_yp", C.c_double), ("coef_x", C.c_double),
                ("coef_y", C.c_double), ("beta", C.c_double), ("planes", C.c_int32),
                ("pad", C.c_int32)]


class WakeKick(C.Structure):
    _fields_ = [("fn", C.c_int32), ("moment", C.c_int32), ("plane", C.c_int32),
                ("multiply", C.c_int32), ("n_turns_wake", C.c_int32),
                ("dot_product", C.c_int32), ("precision", C.c_int32),
                ("reserved", C.c_int32),
                ("Q", C.c_double), ("prefix", C.c_double), ("alpha", C.c_double),
                ("omegabar", C.c_double), ("alpha_over_omegabar", C.c_double),
                ("yokoya", C.c_double), ("L", C.c_double), ("b3", C.c_double),
                ("dt_min_abs", C.c_double), ("k_sigma", C.c_double),
                ("sqrt_z0c", C.c_double), ("wake_factor", C.c_double),
                ("table_t", C.c_void_p), ("table_w", C.c_void_p), ("table_n", C.c_int32),
                ("table_kind", C.c_int32)]


class WakeHist(C.Structure):
    _fields_ = [("n_slices", C.c_int32), ("capacity", C.c_int32),
                ("n_filled", C.c_int32), ("head", C.c_int32),
                ("times", C.c_void_p), ("mom", C.c_void_p), ("age", C.c_void_p),
                ("beta", C.c_void_p)]


class TurnProgram(C.Structure):
    _fields_ = [("n_el", C.c_int32), ("el_type", C.c_int32 * MAX_PROGRAM),
                ("el_index", C.c_int32 * MAX_PROGRAM), ("n_tseg", C.c_int32),
                ("n_long", C.c_int32), ("n_lin", C.c_int32), ("has_wake", C.c_int32),
                ("tseg", TSeg * 4), ("lseq", LongSeq * 2), ("lin", LinMap * 2),
                ("slicer", Slicer), ("beta", C.c_double), ("beta_c", C.c_double),
                ("age_increment", C.c_double), ("n_kicks", C.c_int32), ("pad", C.c_int32),
                ("kick", WakeKick * MAX_KICKS), ("has_sc", C.c_int32),
                ("pad_sc", C.c_int32), ("sc_slicer", Slicer), ("sc", SpaceCharge)]


class Engine(C.Structure):
    _fields_ = [("coords", C.c_void_p * 6), ("n", C.c_int64), ("n_slices", C.c_int32),
                ("n_kicks", C.c_int32), ("hist_capacity", C.c_int32),
                ("record_stats", C.c_int32), ("program_shape", C.c_int32),
                ("kick_shape", C.c_int32), ("pass_grid", C.c_int32),
                ("conv_kinds", C.c_int32), ("det_shape", C.c_int32),
                ("sc_stages", C.c_int32), ("programs", C.c_void_p),
                ("turn_program", C.c_void_p), ("stats", C.c_void_p),
                ("turn_counter", C.c_void_p), ("pending", C.c_void_p),
                ("state", C.c_void_p), ("hist", C.c_void_p), ("z_bins", C.c_void_p),
                ("kicks", C.c_void_p), ("counts", C.c_void_p),
                ("slice_means", C.c_void_p), ("ws", C.c_void_p),
                ("sc_slices", C.c_int32), ("pad_sc", C.c_int32), ("sc_ws", C.c_void_p),
                ("sc_stage_shape", C.c_int32 * (MAX_SC_STAGES + 1)), ("pad_sc2", C.c_int32)]


RECORDS = {
    "phx_slicer": Slicer, "phx_slice_state": SliceState, "phx_term": Term,
    "phx_detuner": Detuner, "phx_tseg": TSeg, "phx_long_op": LongOp,
    "phx_long_seq": LongSeq, "phx_linmap": LinMap, "phx_damper": Damper,
    "phx_sc_slices": SCSlices, "phx_spacecharge": SpaceCharge,
    "phx_wake_kick": WakeKick,
    "phx_wake_hist": WakeHist, "phx_turn_program": TurnProgram, "phx_engine": Engine,
}

P = C.c_void_p
I32, I64, D = C.c_int32, C.c_int64, C.c_double
PTR6 = C.c_void_p * 6

# name -> (restype, argtypes); every symbol declared in include/phx.h
SIGNATURES = {
    "phx_last_error": (C.c_char_p, []),
    "phx_abi_version": (C.c_int, []),
    "phx_init": (C.c_int, [C.c_int]),
    "phx_sizeof": (I64, [C.c_char_p]),
    "phx_pass_group_size": (I32, []),
    "phx_workspace_bytes": (I64, [I64, I32]),
    "phx_bunch_sums_len": (I64, [I32]),
    "phx_bunch_sums": (C.c_int, [P, I32, I64, P, P, P]),
    "phx_slice_cuts": (C.c_int, [P, I64, P, P, P, P, P]),
    "phx_slice_accumulate": (C.c_int, [P, I64, P, I32, P, I32, P, P, P, P, P, P, P]),
    "phx_slice_finalize": (C.c_int, [I32, I32, P, P, I32, D, D, D, P, P]),
    "phx_slice_times": (C.c_int, [P, I32, D, P, P, P]),
    "phx_transverse_segment": (C.c_int, [P, P, P, P, P, I64, P, P]),
    "phx_longitudinal": (C.c_int, [P, P, P, P, P, P, I64, P, P]),
    "phx_linear_map": (C.c_int, [P, P, P, P, I64, P, P]),
    "phx_transverse_damper": (C.c_int, [P, P, I64, P, P, P]),
    "phx_spacecharge_kick": (C.c_int, [P, I64, P, P, P, P, P]),
    "phx_math_wofz": (C.c_int, [P, P, I64, P, P, P]),
    "phx_ucs_workspace_bytes": (I64, [I64]),
    "phx_ucs_sort": (C.c_int, [P, I64, P, P, P, P, P, P]),
    "phx_ucs_assign": (C.c_int, [P, P, I64, C.c_int32, P, P, P, P, P]),
    "phx_math_sincos": (C.c_int, [P, I64, P, P, P]),
    "phx_wake_hist_push": (C.c_int, [P, P, P, P, P, D, D, P]),
    "phx_wake_workspace_bytes": (I64, [I32, I32, I32]),
    "phx_wake_convolve": (C.c_int, [P, P, I32, I32, I32, P, P, P]),
    "phx_wake_apply": (C.c_int, [P, I64, P, P, P, I32, I32, P, P]),
    "phx_engine_ws_bytes": (I64, [I64, I32, I32, I32]),
    "phx_engine_sc_ws_bytes": (I64, [I64, I32]),
    "phx_engine_reset": (C.c_int, [P, P]),
    "phx_engine_turns": (C.c_int, [P, I32, P]),
    "phx_engine_flush": (C.c_int, [P, P]),
    "phx_engine_stage": (C.c_int, [P, I32, P]),
    "phx_engine_snapshot": (C.c_int, [P, P]),
    "phx_engine_turns_noalloc": (C.c_int, [P, I32, I32, P]),
    "phx_batch_desc_bytes": (I64, [I32]),
    "phx_batch_create": (C.c_int, [P, I32, P, I64, P, C.POINTER(C.c_void_p)]),
    "phx_batch_turns": (C.c_int, [P, I32, P]),
    "phx_batch_snapshot": (C.c_int, [P, P]),
    "phx_batch_capture": (C.c_int, [P, I32, C.POINTER(C.c_void_p)]),
    "phx_batch_destroy": (C.c_int, [P]),
    "phx_graph_capture": (C.c_int, [P, I32, C.POINTER(C.c_void_p)]),
    "phx_graph_launch": (C.c_int, [P, P]),
    "phx_graph_destroy": (C.c_int, [P]),
}

_lib = None
_initialized = set()


def load_library(path=LIB_PATH):
    """Load libphx.so and declare every entry point (no GPU needed)."""
    global _lib
    if _lib is None:
        if not os.path.exists(path):
            raise PhxError(
                "libphx.so not found at %s: build it with "
                "`python -c 'import __graft_entry__ as g; g.build()'` or "
                "pyheadtail_amd/csrc/build.sh" % path)
        lib = C.CDLL(path)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def check_layouts(lib=None):
    """Compare the ctypes record sizes with the library's sizeof."""
    lib = lib or load_library()
    bad = {}
    for name, rec in RECORDS.items():
        n = lib.phx_sizeof(name.encode())
        if n != C.sizeof(rec):
            bad[name] = (n, C.sizeof(rec))
    return bad


def lib():
    """The library, initialised on the current torch device.  Raises if the
    library or a HIP device is missing: there is no CPU fallback."""
    L = load_library()
    if not torch.cuda.is_available():
        raise PhxError("pyheadtail_amd needs a HIP device: torch.cuda.is_available() is False")
    dev = torch.cuda.current_device()
    if dev not in _initialized:
        bad = check_layouts(L)
        if bad:
            raise PhxError("ctypes record layouts differ from libphx.so: %r" % bad)
        call(L.phx_init, dev)
        _initialized.add(dev)
    return L


def call(fn, *args):
    rc = fn(*args)
    if rc != 0:
        msg = load_library().phx_last_error()
        raise PhxError("%s failed (%d): %s" % (fn.__name__, rc, msg.decode(errors="replace")))
    return rc


def stream():
    """Handle of torch's current HIP stream."""
    return torch.cuda.current_stream().cuda_stream


def ptr(t):
    """Device address of a tensor (None -> NULL)."""
    if t is None:
        return None
    return t.data_ptr()


def ptr6(ts):
    arr = PTR6()
    for i, t in enumerate(ts):
        arr[i] = ptr(t) if t is not None else None
    return arr


def device():
    return torch.device("cuda", torch.cuda.current_device())


class Workspace:
    """Zero-filled device scratch that the library leaves zero-filled again;
    grown (re-allocated zeroed) on demand, one per (device, stream).

    Calls enqueued on one stream run one after the other, so they can share
    one buffer; calls on different streams may run at the same time and
    each stream has its own.  A buffer is allocated while its stream is
    current, so when a grown buffer replaces it the caching allocator
    reuses the old memory only in that stream's order, after the calls
    already enqueued there."""
    _cache = {}

    @classmethod
    def get(cls, nbytes):
        key = (torch.cuda.current_device(), torch.cuda.current_stream().cuda_stream)
        buf = cls._cache.get(key)
        if buf is None or buf.numel() < nbytes:
            nb = max(int(nbytes), 1 << 20)
            if buf is not None:
                nb = max(nb, 2 * buf.numel())
            buf = torch.zeros(nb, dtype=torch.uint8, device=device())
            cls._cache[key] = buf
        return buf


def workspace(n, n_slices=0):
    L = lib()
    return Workspace.get(L.phx_workspace_bytes(int(n), int(n_slices)))


def device_record(rec):
    """Copy a ctypes record into a new device buffer (uint8 tensor)."""
    raw = bytes(memoryview(rec).cast("B"))
    host = torch.frombuffer(bytearray(raw), dtype=torch.uint8)
    return host.to(device())


def read_record(buf, rec_type):
    """Synchronously read a ctypes record back from a device buffer."""
    host = buf.cpu().numpy().tobytes()
    return rec_type.from_buffer_copy(host[:C.sizeof(rec_type)])
