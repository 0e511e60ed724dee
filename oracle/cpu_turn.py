"""ctypes binding of liborc.so (cpu_turn.cpp), the multi-core C++
restatement of the benchmark's one-turn map: [RFSystems (h = 4620,
V = 4.5 MV), TransverseSegmentMap with chromaticity and amplitude detuning,
WakeField (500 slices over +-3 sigma_z, CircularResonator 1 MOhm/m, 1 GHz,
Q = 1)], the non-linear workload of oracle/workloads.py.

Test infrastructure only: the all-core CPU baseline of bench.py, checked
against refpath.py by tests/test_cpu_oracle.py.  Build: make -C oracle.
"""
import ctypes as C
import os

import numpy as np
from scipy.constants import c, e

from . import workloads as W

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liborc.so")


class Ring(C.Structure):
    _fields_ = ([(k, C.c_double) for k in
                 ("eta0", "drift_length", "amplitude", "harmonic", "circumference",
                  "phi_offset", "phi_lock", "p0", "dQ_x", "dQ_y",
                  "I00", "I01", "I10", "I11", "I22", "I23", "I32", "I33",
                  "J00", "J01", "J10", "J11", "J22", "J23", "J32", "J33",
                  "chroma_x", "chroma_y", "gx", "ax2", "bx", "gy", "ay2", "by",
                  "dapp_x", "dapp_y", "dapp_xy", "dapp_yx", "det_p0")] +
                [("has_chroma", C.c_int32), ("has_amplitude", C.c_int32),
                 ("n_slices", C.c_int32), ("pad", C.c_int32),
                 ("n_sigma_z", C.c_double), ("beta_c", C.c_double),
                 ("wake_factor", C.c_double), ("res_prefix", C.c_double),
                 ("res_alpha", C.c_double), ("res_omegabar", C.c_double)])


class State(C.Structure):
    _fields_ = [("tail", C.c_double), ("width", C.c_double), ("pending", C.c_int32),
                ("pad", C.c_int32), ("kick_x", C.c_void_p), ("kick_y", C.c_void_p)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError("%s is missing: make -C oracle" % LIB)
        L = C.CDLL(LIB)
        P, I64 = C.c_void_p, C.c_int64
        L.orc_threads.argtypes, L.orc_threads.restype = [C.c_int], C.c_int
        L.orc_turns.argtypes = [P] * 6 + [I64, P, P, C.c_int32]
        L.orc_turns.restype = C.c_int
        L.orc_flush.argtypes = [P] * 6 + [I64, P, P]
        L.orc_flush.restype = None
        _lib = L
    return _lib


def nonlinear_ring(ob, n_slices=500):
    """The ring record of workloads.nonlinear_map_oracle for bunch ob, with
    the scalars formed by the reference's expressions (the same ones
    refpath.py evaluates)."""
    r = Ring()
    p0 = W.P0
    # RFSystems (refpath.drift / rf_kick)
    r.eta0 = W.ALPHA_C - ob.gamma**-2
    r.drift_length = 0.5 * W.CIRCUMFERENCE
    r.amplitude = np.abs(ob.charge) * W.V_RF / (ob.beta * c)
    r.harmonic = float(W.H_RF)
    r.circumference = W.CIRCUMFERENCE
    r.phi_offset = 0.
    r.phi_lock = 0.
    r.p0 = ob.p0
    # transverse segment (refpath.smooth_segment)
    from . import refpath as R
    seg = R.smooth_segment(W.Q_X, W.Q_Y, W.BETA_X, W.BETA_Y, Qp_x=5., Qp_y=5.,
                           app_x=1e-10 * p0, app_y=1e-10 * p0, app_xy=-5e-11 * p0)
    r.dQ_x, r.dQ_y = seg.dQ_x, seg.dQ_y
    for i, j in ((0, 0), (0, 1), (1, 0), (1, 1), (2, 2), (2, 3), (3, 2), (3, 3)):
        setattr(r, "I%d%d" % (i, j), seg.I[i, j])
        setattr(r, "J%d%d" % (i, j), seg.J[i, j])
    ch, am = seg.detuners
    r.has_chroma, r.chroma_x, r.chroma_y = 1, ch.cx[0], ch.cy[0]
    ax, bx, ay, by = am.o
    r.has_amplitude = 1
    r.gx, r.ax2, r.bx = (1 + ax**2) / bx, 2 * ax, bx
    r.gy, r.ay2, r.by = (1 + ay**2) / by, 2 * ay, by
    r.dapp_x, r.dapp_y, r.dapp_xy, r.dapp_yx = am.d
    r.det_p0 = ob.p0
    # wake field (refpath.WakeFieldOracle, resonator_transverse, Kick)
    r.n_slices = n_slices
    r.n_sigma_z = 3.
    r.beta_c = ob.beta * c
    r.wake_factor = (-(ob.charge)**2 / (ob.mass * ob.gamma * (ob.beta * c)**2) *
                     ob.particlenumber_per_mp)
    omega = 2 * np.pi * 1e9
    alpha = omega / (2 * 1.)
    omegabar = np.sqrt(np.abs(omega**2 - alpha**2))
    r.res_prefix = 1. * 1e6 * omega**2 / (1. * omegabar)
    r.res_alpha, r.res_omegabar = alpha, omegabar
    return r


class CpuTurns(object):
    """Turns of the non-linear workload on host arrays (NumPy float64,
    modified in place) with the C++ restatement on n_threads threads."""

    def __init__(self, ob, n_slices=500, n_threads=0):
        self.L = lib()
        self.threads = self.L.orc_threads(int(n_threads))
        self.ring = nonlinear_ring(ob, n_slices)
        self.kx = np.zeros(n_slices)
        self.ky = np.zeros(n_slices)
        self.state = State(0., 0., 0, 0, self.kx.ctypes.data, self.ky.ctypes.data)
        self.coords = [np.ascontiguousarray(getattr(ob, k), dtype=np.float64)
                       for k in ('x', 'xp', 'y', 'yp', 'z', 'dp')]
        self.n = ob.macroparticlenumber

    def _ptrs(self):
        return [a.ctypes.data for a in self.coords]

    def track_turns(self, n_turns):
        rc = self.L.orc_turns(*self._ptrs(), self.n, C.byref(self.ring), C.byref(self.state),
                              int(n_turns))
        if rc != 0:
            raise RuntimeError("orc_turns failed")

    def flush(self):
        self.L.orc_flush(*self._ptrs(), self.n, C.byref(self.ring), C.byref(self.state))

    def result(self):
        self.flush()
        return dict(zip(('x', 'xp', 'y', 'yp', 'z', 'dp'), self.coords))
