"""Fused one-turn map: many turns of [local elements..., WakeField] with four
kernel launches per turn and no host synchronisation.

FusedOneTurnMap(one_turn_map) accepts the element list that a user would
otherwise track one by one (as Synchrotron.one_turn_map,
machines/synchrotron.py:294-298, or the loop of
tests/simulation_tests/test_bb_resonator.py:91-96): TransverseSegmentMap,
LinearMap, RFSystems, Drift and Kick elements, TransverseGaussianSpaceCharge
kicks between them (as Synchrotron.install_after_each_transverse_segment
places them), optionally ended by one WakeField.  It is itself an Element,
so it can replace that list.

A map with m space-charge kicks runs each turn as m + 1 stages: stage 0
is the elements before the first kick, stage k the kick k, then the
elements up to the next kick; after each of the first m stages the
space-charge slicing of the current state (counts, means, standard
deviations per slice) is computed on the device for the kick that follows.

Host scalars (p0 updates of the RF kicks, wake ages, wake factors) are
advanced on the host for every turn with the elements' own arithmetic
(the same code the element-by-element path runs) and uploaded as per-turn
programs; identical programs are stored once.  The wake kick of the last
turn stays pending on the bunch until a coordinate is read (or flush()).
"""
import ctypes as C
import math

import numpy as np
import torch
from scipy.constants import c, e

from . import _phx
from .general.element import Element
from .impedances.wakes import WakeField
from .particles.particles import ScalarState, COORDS
from .spacecharge.spacecharge import TransverseGaussianSpaceCharge
from .trackers.longitudinal_tracking import Drift, Kick, LinearMap, RFSystems
from .trackers.transverse_tracking import TransverseSegmentMap

STAT_NAMES = ['mean_x', 'mean_xp', 'mean_y', 'mean_yp', 'mean_z', 'mean_dp',
              'sigma_x', 'sigma_y', 'sigma_z', 'sigma_dp', 'epsn_x', 'epsn_y', 'epsn_z']


def stats_from_sums(rows, n, betagamma, p0):
    """Bunch statistics from the engine's raw per-turn sums, with the
    formulas of Particles.mean_* / sigma_* / epsn_* (np.mean,
    cobra_functions/stats.pyx:cov and emittance)."""
    rows = np.asarray(rows, dtype=np.float64)
    k = 6

    def cov(i, j):
        lo, hi = min(i, j), max(i, j)
        p = lo * k - lo * (lo - 1) // 2 + (hi - lo)
        if n < 2:
            return np.zeros(len(rows))
        return (rows[:, 12 + p] - rows[:, 6 + i] * rows[:, 6 + j] / n) / (n - 1)

    def emit(u, up, dp):
        s11, s12, s22 = cov(u, u), cov(up, u), cov(up, up)
        if dp is not None:
            cdp2 = cov(dp, dp)
            nz = cdp2 != 0
            cudp, cupdp = cov(u, dp), cov(up, dp)
            safe = np.where(nz, cdp2, 1.0)
            s11 = s11 - np.where(nz, cudp * cudp / safe, 0.)
            s12 = s12 - np.where(nz, cudp * cupdp / safe, 0.)
            s22 = s22 - np.where(nz, cupdp * cupdp / safe, 0.)
        return np.sqrt(s11 * s22 - s12 * s12)

    out = {}
    for i, name in enumerate(COORDS):
        out['mean_' + name] = rows[:, i] / n
    for name, i in (('x', 0), ('y', 2), ('z', 4), ('dp', 5), ('xp', 1), ('yp', 3)):
        out['sigma_' + name] = np.sqrt(cov(i, i))
    out['epsn_x'] = emit(0, 1, 5) * betagamma
    out['epsn_y'] = emit(2, 3, 5) * betagamma
    out['epsn_z'] = 4 * np.pi * emit(4, 5, None) * p0 / e
    return out


class FusedOneTurnMap(Element):
    """Runs turns of a one-turn map with the fused device engine."""

    LOCAL = (TransverseSegmentMap, LinearMap, RFSystems, Drift, Kick)

    def __init__(self, one_turn_map, record_stats=False, graph_turns=8, specialise=True,
                 pass_grid=0, *args, **kwargs):
        elements = list(one_turn_map)
        self.wake = None
        if elements and isinstance(elements[-1], WakeField):
            self.wake = elements.pop()
        for el in elements:
            if not isinstance(el, self.LOCAL + (TransverseGaussianSpaceCharge,)):
                raise TypeError("FusedOneTurnMap cannot fuse %s: supported are "
                                "TransverseSegmentMap, LinearMap, RFSystems, Drift, Kick, "
                                "TransverseGaussianSpaceCharge and one final WakeField"
                                % type(el).__name__)
        # stages: the elements before the first space-charge kick, then
        # each kick with the elements up to the next one
        self.sc = [el for el in elements if isinstance(el, TransverseGaussianSpaceCharge)]
        self.stages = [[]]
        for el in elements:
            if isinstance(el, TransverseGaussianSpaceCharge):
                self.stages.append([])
            else:
                self.stages[-1].append(el)
        if len(self.sc) > _phx.MAX_SC_STAGES:
            raise ValueError("the fused engine takes at most %d space-charge kicks per turn"
                             % _phx.MAX_SC_STAGES)
        for sc in self.sc:
            if sc.slicer.mode != 'uniform_bin':
                raise ValueError("the fused engine slices with uniform bins; track a "
                                 "space-charge kick with a %s slicer as an element"
                                 % sc.slicer.mode)
        for st in self.stages:
            n_t = sum(isinstance(el, TransverseSegmentMap) for el in st)
            n_l = sum(isinstance(el, (RFSystems, Drift, Kick)) for el in st)
            n_m = sum(isinstance(el, LinearMap) for el in st)
            if n_t > 4 or n_l > 2 or n_m > 2 or len(st) > _phx.MAX_PROGRAM:
                raise ValueError("too many elements for one fused pass")
        if self.wake is not None and not self.wake.wake_kicks:
            raise ValueError("WakeField without kicks")
        if self.wake is not None and self.wake.slicer.mode != 'uniform_bin':
            raise ValueError("the fused engine slices with uniform bins; track a WakeField "
                             "with a %s slicer as an element" % self.wake.slicer.mode)
        if self.wake is not None and len(self.wake.wake_kicks) > _phx.MAX_KICKS:
            raise ValueError("the fused engine takes at most %d wake kicks; track this "
                             "WakeField as an element" % _phx.MAX_KICKS)
        self.local = [el for el in elements if not isinstance(el, TransverseGaussianSpaceCharge)]
        self.record_stats = bool(record_stats)
        self.graph_turns = int(graph_turns)
        self._bufs = None
        self._graphs = {}
        self._engine = None
        # the bunch whose last wake kick this engine still holds (its device
        # state: pending flag, cuts and kick table, is the engine's)
        self._pending_beam = None
        self.last_stats = None
        self._shape = (_phx.SHAPE_GENERIC, _phx.KICKS_GENERIC, _phx.DETS_GENERIC)
        self._sc_shapes = [_phx.SHAPE_GENERIC] * len(self.sc)
        self._conv_kinds = 0
        # False: always the generic pass (the same arithmetic, for checks)
        self.specialise = bool(specialise)
        # blocks of the particle pass (0: as many as stay resident); the cuts
        # equal those of the element-by-element slicing bit for bit at
        # _phx.STANDALONE_GRID
        self.pass_grid = int(pass_grid)

    # -- per-turn programs -------------------------------------------------------
    def _program(self, state):
        """The programs of one turn's stages (one without space charge),
        advancing the host scalars as the elements do, in map order."""
        progs = []
        for k, st in enumerate(self.stages):
            if k > 0:
                # the kick that ends the previous stage: its parameters at
                # this point of the turn (spacecharge.py:209-210)
                prev = progs[-1]
                sc = self.sc[k - 1]
                prev.has_sc = 1
                prev.sc_slicer = sc.slicer.phx_config()
                prev.sc = sc._params(state)
            progs.append(self._stage_program(st, state))
        if self.wake is not None:
            w = self.wake
            for prog in progs:
                prog.has_wake = 1
                prog.age_increment = state.circumference / (state.beta * c)
                prog.slicer = w.slicer.phx_config()
                prog.beta = state.beta
                prog.beta_c = state.beta * c
                prog.n_kicks = len(w.wake_kicks)
            for i, k in enumerate(w.wake_kicks):
                rec = k.record(state)
                for prog in progs:
                    prog.kick[i] = rec
        return progs

    def _stage_program(self, elements, state):
        prog = _phx.TurnProgram()
        n_t = n_l = n_m = 0
        for i, el in enumerate(elements):
            if isinstance(el, TransverseSegmentMap):
                prog.el_type[i], prog.el_index[i] = _phx.EL_TSEG, n_t
                prog.tseg[n_t] = el._advance(state)
                n_t += 1
            elif isinstance(el, LinearMap):
                prog.el_type[i], prog.el_index[i] = _phx.EL_LINMAP, n_m
                prog.lin[n_m] = el._advance(state)
                n_m += 1
            else:
                prog.el_type[i], prog.el_index[i] = _phx.EL_LONG, n_l
                if isinstance(el, RFSystems):
                    prog.lseq[n_l] = el._advance(state)
                else:
                    seq = _phx.LongSeq()
                    seq.n_ops = 1
                    seq.op[0] = el._op(state)
                    prog.lseq[n_l] = seq
                n_l += 1
        prog.n_el = len(elements)
        prog.n_tseg, prog.n_long, prog.n_lin = n_t, n_l, n_m
        return prog

    @staticmethod
    def _shapes(prog, wake):
        """(PHX_SHAPE_*, PHX_KICKS_*) of a turn program: its element order,
        when the library has a pass compiled for it, and its kick set.  All
        programs of a map share them (the same elements in every turn)."""
        kinds = []
        for i in range(prog.n_el):
            t, k = prog.el_type[i], prog.el_index[i]
            if t == _phx.EL_TSEG:
                kinds.append('T')
            elif t == _phx.EL_LINMAP:
                kinds.append('LIN')
            else:
                seq = prog.lseq[k]
                ops = [seq.op[j].type for j in range(seq.n_ops)]
                nh = len(ops) - 2
                rf = (nh in (1, 2) and ops[0] == _phx.LONG_DRIFT and ops[-1] == _phx.LONG_DRIFT
                      and all(o == _phx.LONG_KICK for o in ops[1:-1]))
                kinds.append('RF%d' % nh if rf else '?')
        shape = {('RF1', 'T'): _phx.SHAPE_RF1_T, ('RF2', 'T'): _phx.SHAPE_RF2_T,
                 ('T', 'RF1'): _phx.SHAPE_T_RF1, ('T', 'RF2'): _phx.SHAPE_T_RF2,
                 ('LIN', 'T'): _phx.SHAPE_LIN_T, ('T', 'LIN'): _phx.SHAPE_T_LIN,
                 ('T',): _phx.SHAPE_T}.get(tuple(kinds), _phx.SHAPE_GENERIC)
        kicks = _phx.KICKS_GENERIC
        if wake is not None:
            sig = tuple((prog.kick[i].plane, prog.kick[i].multiply)
                        for i in range(prog.n_kicks))
            if sig == ((_phx.PLANE_XP, -1), (_phx.PLANE_YP, -1)):
                kicks = _phx.KICKS_DXDY
            elif sig == ((_phx.PLANE_XP, -1), (_phx.PLANE_XP, 0), (_phx.PLANE_YP, -1),
                         (_phx.PLANE_YP, 2)):
                kicks = _phx.KICKS_PLATES
        dets = _phx.DETS_GENERIC
        if prog.n_tseg >= 1 and all(FusedOneTurnMap._det_key(prog.tseg[i]) ==
                                    FusedOneTurnMap._det_key(prog.tseg[0])
                                    for i in range(prog.n_tseg)):
            dets = {(): _phx.DETS_NONE, ((_phx.DET_CHROMA, 1),): _phx.DETS_C1,
                    ((_phx.DET_CHROMA, 1), (_phx.DET_AMPLITUDE, 0)): _phx.DETS_C1_AMP
                    }.get(FusedOneTurnMap._det_key(prog.tseg[0]), _phx.DETS_GENERIC)
            if dets in (_phx.DETS_C1, _phx.DETS_C1_AMP) and all(
                    FusedOneTurnMap._smooth(prog.tseg[i]) for i in range(prog.n_tseg)):
                dets |= _phx.DETS_SMOOTH
        if shape == _phx.SHAPE_RF1_T and FusedOneTurnMap._stationary(prog):
            shape |= _phx.SHAPE_STATIONARY
        return shape, kicks, dets

    @staticmethod
    def _stationary(prog):
        """The program's RF sequences have drifts with beta_ratio = 1 and
        kicks with phi_offset = phi_lock = p_increment = +0 (sign bit clear)
        and no dispersion: the terms PHX_SHAPE_STATIONARY drops leave the
        values unchanged."""
        for i in range(prog.n_long):
            seq = prog.lseq[i]
            for j in range(seq.n_ops):
                o = seq.op[j]
                if o.type == _phx.LONG_DRIFT:
                    if o.beta_ratio != 1.0:
                        return False
                elif o.with_dispersion != 0 or any(
                        v != 0.0 or math.copysign(1.0, v) < 0.0
                        for v in (o.phi_offset, o.phi_lock, o.p_increment)):
                    return False
        return True

    @staticmethod
    def _smooth(tseg):
        """A segment record whose matrices give M00 = cos etc. exactly: I the
        identity, J00 = J11 = J22 = J33 = 0 and no dispersion (alpha = 0 and
        equal beta at both ends, as in smooth optics); its amplitude
        detuners with ax2 = ay2 = 0 and gx, gy > 0 (their 2 alpha u u' term
        a zero added to gx u^2 >= +0)."""
        amp_ok = all(tseg.det[i].ax2 == 0.0 and tseg.det[i].ay2 == 0.0 and
                     tseg.det[i].gx > 0.0 and tseg.det[i].gy > 0.0
                     for i in range(tseg.n_detuners)
                     if tseg.det[i].type == _phx.DET_AMPLITUDE)
        return (tseg.I00 == 1.0 and tseg.I11 == 1.0 and tseg.I22 == 1.0 and tseg.I33 == 1.0 and
                tseg.I01 == 0.0 and tseg.I10 == 0.0 and tseg.I23 == 0.0 and tseg.I32 == 0.0 and
                tseg.J00 == 0.0 and tseg.J11 == 0.0 and tseg.J22 == 0.0 and tseg.J33 == 0.0 and
                tseg.with_dispersion == 0 and amp_ok)

    @classmethod
    def _stage_shapes(cls, progs, wake):
        """Space-charge engines: ((stage 0's PHX_SHAPE_*, the kick set, the
        detuner set shared by every stage's segments), [PHX_SHAPE_* of
        stages 1..m]); a stage without elements is PHX_SHAPE_NONE."""
        shapes, dets = [], set()
        for q in progs:
            sh, _, ds = cls._shapes(q, None)
            shapes.append(_phx.SHAPE_NONE if q.n_el == 0 else sh)
            if q.n_tseg > 0:
                dets.add(ds)
        kicks = cls._shapes(progs[0], wake)[1]
        ds = dets.pop() if len(dets) == 1 else _phx.DETS_GENERIC
        # the stage passes run the general RF code (no stationary variant)
        return (shapes[0], kicks, ds), [sh & ~_phx.SHAPE_STATIONARY for sh in shapes[1:]]

    @staticmethod
    def _det_key(tseg):
        """(type, order) of each detuner of a segment record (order 0 for
        amplitude detuning)."""
        return tuple((tseg.det[i].type,
                      tseg.det[i].order if tseg.det[i].type == _phx.DET_CHROMA else 0)
                     for i in range(tseg.n_detuners))

    @staticmethod
    def _state_key(state):
        return (state._gamma, state._beta, state._betagamma, state._p0)

    def _schedule(self, beam, n_turns):
        """Per-turn programs from the host-scalar recurrence.  Once a turn
        leaves the scalars unchanged, every later turn has the same program
        and the same scalars (the recurrence is deterministic), so the rest
        of the table is filled without evaluating it.  A turn's stage
        programs are stored together; the table holds the first's index."""
        state = ScalarState(beam)
        progs, table, index = [], np.empty(n_turns, dtype=np.int32), {}
        n_st = len(self.stages)
        t = 0
        while t < n_turns:
            before = self._state_key(state)
            ps = self._program(state)
            p = ps[0]
            if t == 0 and self.specialise:
                self._shape = self._shapes(p, self.wake)
                if n_st > 1:
                    self._shape, self._sc_shapes = self._stage_shapes(ps, self.wake)
            for k in range(p.n_kicks):
                self._conv_kinds |= _phx.conv_kind(p.kick[k])
            key = b"".join(bytes(memoryview(q).cast("B")) for q in ps)
            if key not in index:
                index[key] = len(progs)
                progs.append(key)
                # every program must be stationary for the stationary pass
                if self._shape[0] & _phx.SHAPE_STATIONARY and not self._stationary(p):
                    self._shape = (self._shape[0] & ~_phx.SHAPE_STATIONARY,) + self._shape[1:]
            table[t] = index[key] * n_st
            t += 1
            if self._state_key(state) == before:
                table[t:] = index[key]
                break
        state.write_back(beam)
        return progs, table

    # -- device buffers ----------------------------------------------------------
    def _ensure_buffers(self, beam, n_turns, n_programs):
        n = beam.macroparticlenumber
        dev = _phx.device()
        b = self._bufs
        if b is None or b['n'] != n:
            self._drop_graphs()
            S_sc = max([int(sc.slicer.n_slices) for sc in self.sc] + [0])
            S = int(self.wake.slicer.n_slices) if self.wake is not None else 0
            n_kicks = len(self.wake.wake_kicks) if self.wake is not None else 0
            T = self.wake.n_turns_wake_max if self.wake is not None else 1
            b = dict(n=n, S=S, n_kicks=n_kicks, T=T, turn_cap=0, prog_cap=0, S_sc=S_sc)
            if S_sc > 0:
                nsc = _phx.lib().phx_engine_sc_ws_bytes(n, S_sc)
                b['sc_ws'] = torch.zeros(nsc, dtype=torch.uint8, device=dev)
            b['counter'] = torch.zeros(1, dtype=torch.int32, device=dev)
            b['pending'] = torch.zeros(1, dtype=torch.int32, device=dev)
            b['state'] = torch.zeros(C.sizeof(_phx.SliceState), dtype=torch.uint8, device=dev)
            b['z_bins'] = torch.zeros(max(S + 1, 1), dtype=torch.float64, device=dev)
            b['kicks'] = torch.zeros((max(n_kicks, 1), max(S, 1)), dtype=torch.float64,
                                     device=dev)
            b['counts'] = torch.zeros(max(S, 1), dtype=torch.int32, device=dev)
            b['means'] = torch.zeros((2, max(S, 1)), dtype=torch.float64, device=dev)
            nb = _phx.lib().phx_engine_ws_bytes(n, S, n_kicks, T)
            b['ws'] = torch.zeros(nb, dtype=torch.uint8, device=dev)
            if self.wake is not None:
                from .impedances.wake_kicks import WakeHistory
                b['history'] = WakeHistory(S, T)
                b['hist'] = b['history'].record
            else:
                b['history'] = b['hist'] = None
            b['needs_reset'] = True
            self._bufs = b
        if b['turn_cap'] < n_turns or b['prog_cap'] < n_programs:
            self._drop_graphs()
            b['turn_cap'] = max(n_turns, 2 * b['turn_cap'], 64)
            b['prog_cap'] = max(n_programs, 2 * b['prog_cap'], 4)
            psz = C.sizeof(_phx.TurnProgram) * len(self.stages)
            b['programs'] = torch.zeros(b['prog_cap'] * psz, dtype=torch.uint8, device=dev)
            b['table'] = torch.zeros(b['turn_cap'], dtype=torch.int32, device=dev)
            # host mirrors of the device programs and table: a call uploads
            # only what differs from them
            b['host_programs'] = b''
            b['host_table'] = np.zeros(b['turn_cap'], dtype=np.int32)
            b['stats'] = torch.zeros((b['turn_cap'] + 1, _phx.STAT_NV), dtype=torch.float64,
                                     device=dev)
        return b

    def _engine_record(self, beam, b, n_turns):
        eng = _phx.Engine()
        for i, k in enumerate(COORDS):
            eng.coords[i] = beam._raw(k).data_ptr()
        eng.n = beam.macroparticlenumber
        eng.n_slices = b['S']
        eng.n_kicks = b['n_kicks']
        eng.hist_capacity = b['T']
        eng.record_stats = 1 if self.record_stats else 0
        eng.program_shape, eng.kick_shape, eng.det_shape = self._shape
        eng.pass_grid = self.pass_grid
        eng.conv_kinds = self._conv_kinds
        eng.programs = b['programs'].data_ptr()
        eng.turn_program = b['table'].data_ptr()
        eng.stats = b['stats'].data_ptr()
        eng.turn_counter = b['counter'].data_ptr()
        eng.pending = b['pending'].data_ptr()
        eng.state = b['state'].data_ptr()
        eng.hist = _phx.ptr(b['hist'])
        eng.z_bins = b['z_bins'].data_ptr()
        eng.kicks = b['kicks'].data_ptr()
        eng.counts = b['counts'].data_ptr()
        eng.slice_means = b['means'].data_ptr()
        eng.ws = b['ws'].data_ptr()
        eng.sc_stages = len(self.sc)
        for k, sh in enumerate(self._sc_shapes):
            eng.sc_stage_shape[k + 1] = sh
        eng.sc_slices = b['S_sc']
        eng.sc_ws = b['sc_ws'].data_ptr() if b['S_sc'] > 0 else None
        return eng

    def _drop_graphs(self):
        L = _phx.load_library()
        for h in self._graphs.values():
            L.phx_graph_destroy(h)
        self._graphs = {}

    def __del__(self):
        try:
            self._drop_graphs()
        except Exception:
            pass

    # -- tracking ---------------------------------------------------------------
    def reset(self):
        """Forget the wake history (the next turn starts without memory)."""
        if self._bufs is not None:
            self._bufs['needs_reset'] = True

    def track(self, beam):
        self.track_turns(beam, 1)

    def reserve(self, beam, n_turns, n_programs=4):
        """Size the per-turn device buffers (programs, turn table, statistics
        rows) for calls of up to n_turns turns, so that later calls neither
        re-allocate them nor capture their graphs again."""
        self._ensure_buffers(beam, int(n_turns), int(n_programs))

    def _prepare(self, beam, n_turns):
        """Schedule n_turns, upload the programs, reset the turn counter;
        returns the engine record."""
        if beam._pending is not None and beam._pending is not self:
            beam._run_pending()
        held = self._pending_beam
        resetting = self._bufs is not None and self._bufs.get('needs_reset')
        if held is not None and (held is not beam or resetting):
            # the kick of the held bunch's last turn: applied to that bunch
            # before its cuts and kicks are replaced by another bunch's, or
            # cleared by a reset
            if held._pending is self:
                held._run_pending()
            self._pending_beam = None
        for k in COORDS:
            beam._raw(k)  # all six coordinates must exist
        L = _phx.lib()
        progs, table = self._schedule(beam, n_turns)
        b = self._ensure_buffers(beam, n_turns, len(progs))
        blob = b"".join(progs)
        # copies from pageable memory wait for the stream: only what differs
        # from the device's current programs and table is uploaded
        if b['host_programs'][:len(blob)] != blob:
            host = torch.frombuffer(bytearray(blob), dtype=torch.uint8)
            b['programs'][:len(blob)].copy_(host)
            b['host_programs'] = blob + b['host_programs'][len(blob):]
        if not np.array_equal(b['host_table'][:n_turns], table):
            b['table'][:n_turns].copy_(torch.from_numpy(table))
            b['host_table'][:n_turns] = table
        eng = self._engine_record(beam, b, n_turns)
        self._engine = eng
        if b.get('needs_reset'):
            _phx.call(L.phx_engine_reset, C.byref(eng), _phx.stream())
            b['needs_reset'] = False
        b['counter'].zero_()
        self._n_turns_last = n_turns
        return eng

    def _finish(self, beam):
        beam.clean_slices()
        if self.wake is not None:
            beam._set_pending(self)
            self._pending_beam = beam
        elif self.record_stats:
            self._finish_stats(beam)

    def track_turns(self, beam, n_turns, use_graph=True):
        """Advance beam by n_turns turns."""
        if n_turns <= 0:
            return
        L = _phx.lib()
        eng = self._prepare(beam, n_turns)
        s = _phx.stream()
        G = self.graph_turns
        done = 0
        if use_graph and G > 0 and n_turns >= G:
            key = (tuple(eng.coords), self.record_stats, G)
            h = self._graphs.get(key)
            if h is None:
                h = C.c_void_p()
                _phx.call(L.phx_graph_capture, C.byref(eng), G, C.byref(h))
                self._graphs[key] = h
            _phx.call(L.phx_engine_snapshot, C.byref(eng), s)  # the graph does not
            while n_turns - done >= G:
                _phx.call(L.phx_graph_launch, h, s)
                done += G
        if n_turns > done:
            _phx.call(L.phx_engine_turns, C.byref(eng), n_turns - done, s)
        self._finish(beam)

    def timed_turns(self, beam, n_turns):
        """Run n_turns with one launch per stage, bracketing every stage with
        HIP events on the launch stream; returns the mean duration of each
        stage in ms (pass, histogram, convolution)."""
        L = _phx.lib()
        eng = self._prepare(beam, n_turns)
        s = _phx.stream()
        n_st = 3 if self.wake is not None else 1
        _phx.call(L.phx_engine_snapshot, C.byref(eng), s)
        ev = [[torch.cuda.Event(enable_timing=True) for _ in range(n_st + 1)]
              for _ in range(n_turns)]
        # a spin kernel holds the stream while the host enqueues every stage
        # and event, so that the stages then run back to back and the events
        # bracket the kernels rather than the host's launch latency
        torch.cuda._sleep(int(n_turns) * 400000)
        for t in range(n_turns):
            ev[t][0].record()
            for st in range(n_st):
                _phx.call(L.phx_engine_stage, C.byref(eng), st, s)
                ev[t][st + 1].record()
        torch.cuda.synchronize()
        self._finish(beam)
        return [float(np.mean([ev[t][st].elapsed_time(ev[t][st + 1]) for t in range(n_turns)]))
                for st in range(n_st)]

    def flush(self, beam):
        """Apply the pending wake kick of the last turn (and record the
        final statistics row)."""
        L = _phx.lib()
        _phx.call(L.phx_engine_flush, C.byref(self._engine), _phx.stream())
        if self._pending_beam is beam:
            self._pending_beam = None
        if self.record_stats:
            self._finish_stats(beam)

    def _finish_stats(self, beam):
        n = self._n_turns_last
        rows = self._bufs['stats'][:n + 1].cpu().numpy()
        self.last_stats = stats_from_sums(rows, beam.macroparticlenumber, beam.betagamma,
                                          beam.p0)
        self.last_stats['raw'] = rows

    @property
    def slice_data(self):
        """(z_bins, counts, mean_x, mean_y) of the last slicing (device)."""
        b = self._bufs
        return b['z_bins'], b['counts'], b['means'][0], b['means'][1]

    @property
    def pending_kicks(self):
        return self._bufs['kicks']
