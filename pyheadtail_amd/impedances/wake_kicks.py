"""Wake kicks: convolution of per-slice source moments with a wake function,
and the kick on the particles inside the cuts.

Contract of PyHEADTAIL impedances/wake_kicks.py:
  - WakeKick(wake_function, slicer, n_turns_wake) picks the np.convolve form
    for uniform-bin slicing with one turn of memory or fixed cuts, else the
    dense dot-product form (:41-53);
  - _wake_factor(bunch) = -q^2/(m gamma (beta c)^2) * N_per_mp (:66-73);
  - apply(bunch, slice_set_list, slice_set_age_list) sums the convolutions
    over min(len(list), n_turns_wake) stored turns, scales by the wake factor
    of the current bunch and adds the kick to the particles of the newest
    SliceSet (:118-139, :144-361);
  - eleven kick classes differing in source moment (n, n*mean_x, n*mean_y),
    kicked momentum (xp, yp, dp) and an optional factor x or y (quadrupolar).

Each kick is a phx_wake_kick record; the convolution and the particle kick
run on the device (phx_wake_convolve, phx_wake_apply).  WakeField batches
all of its kicks into one convolution launch and one particle pass, with the
same per-kick arithmetic and the kicks applied in list order.
"""
import numpy as np
import torch
from scipy.constants import c

from .. import _phx
from ..general.element import Printing


class WakeKick(Printing):
    """Base class; subclasses fix (moment, plane, multiply)."""
    moment = _phx.MOM_N
    plane = _phx.PLANE_XP
    multiply = -1
    # arithmetic of a multi-turn resistive-wall convolution (_phx.PREC_*)
    precision = _phx.PREC_FP64

    def __init__(self, wake_function, slicer, n_turns_wake, *args, **kwargs):
        self.wake_function = wake_function
        self.n_turns_wake = n_turns_wake
        self.dot_product = not (slicer.mode == 'uniform_bin' and
                                (n_turns_wake == 1 or slicer.z_cuts))
        if self.dot_product and n_turns_wake > 1:
            self.warns('You use multiturn wakes. The stored '
                       'SliceSet instances from previous turns are converted '
                       'from z to dt according to the current turn\'s gamma. '
                       'If you accelerate this is only an approximation which '
                       'becomes invalid when gamma changes much over '
                       'n_turns_wake.')

    @staticmethod
    def _wake_factor(bunch):
        return (-(bunch.charge)**2 / (bunch.mass * bunch.gamma * (bunch.beta * c)**2) *
                bunch.particlenumber_per_mp)

    def record(self, bunch):
        """phx_wake_kick record for the current bunch."""
        k = _phx.WakeKick()
        self.wake_function.fill(k)
        k.moment = self.moment
        k.plane = self.plane
        k.multiply = self.multiply
        k.n_turns_wake = int(self.n_turns_wake)
        k.dot_product = 1 if self.dot_product else 0
        k.precision = self.precision
        k.wake_factor = self._wake_factor(bunch)
        return k

    def apply(self, bunch, slice_set_list, slice_set_age_list):
        """Kick bunch from the stored SliceSets (newest first)."""
        n_turns = min(len(slice_set_age_list), self.n_turns_wake)
        ring = WakeHistory(slice_set_list[0].n_slices, max(1, n_turns))
        for i in reversed(range(n_turns)):
            ring.push_sliceset(slice_set_list[i], 0.0)
        ring.set_ages([slice_set_age_list[i] for i in range(n_turns)])
        kicks = ring.convolve([self.record(bunch)])
        apply_kicks(bunch, slice_set_list[0], [self.record(bunch)], kicks)


class WakeHistory(object):
    """Device ring of the slice data of past turns (phx_wake_hist)."""

    def __init__(self, n_slices, capacity):
        dev = _phx.device()
        self.n_slices = n_slices
        self.capacity = capacity
        self.times = torch.zeros((capacity, n_slices), dtype=torch.float64, device=dev)
        self.mom = torch.zeros((capacity, 3, n_slices), dtype=torch.float64, device=dev)
        self.age = torch.zeros(capacity, dtype=torch.float64, device=dev)
        self.beta = torch.zeros(capacity, dtype=torch.float64, device=dev)
        rec = _phx.WakeHist()
        rec.n_slices, rec.capacity = n_slices, capacity
        rec.n_filled, rec.head = 0, capacity - 1
        rec.times, rec.mom = self.times.data_ptr(), self.mom.data_ptr()
        rec.age, rec.beta = self.age.data_ptr(), self.beta.data_ptr()
        self.record = _phx.device_record(rec)
        self.n_pushed = 0
        self._ws = None

    def push_sliceset(self, sliceset, age_increment):
        """Ages += age_increment, then add sliceset with age 0."""
        L = _phx.lib()
        mx = getattr(sliceset, 'mean_x', None)
        my = getattr(sliceset, 'mean_y', None)
        _phx.call(L.phx_wake_hist_push, self.record.data_ptr(), sliceset.times.data_ptr(),
                  sliceset.n_macroparticles_per_slice.data_ptr(), _phx.ptr(mx), _phx.ptr(my),
                  float(age_increment), float(sliceset.beta), _phx.stream())
        self.n_pushed += 1

    def set_ages(self, ages_newest_first):
        """Overwrite the ages of the stored entries (host values)."""
        n = len(ages_newest_first)
        head = (self.n_pushed - 1) % self.capacity
        slots = [(head - i) % self.capacity for i in range(n)]
        a = self.age.cpu()
        for s, v in zip(slots, ages_newest_first):
            a[s] = float(v)
        self.age.copy_(a)

    def convolve(self, records):
        L = _phx.lib()
        n_kicks = len(records)
        S = self.n_slices
        out = torch.empty((n_kicks, S), dtype=torch.float64, device=_phx.device())
        recs = (_phx.WakeKick * n_kicks)(*records)
        max_turns = max(1, min(self.n_pushed, self.capacity))
        # own zero-filled workspace: its tickets must start (and stay) zero
        nb = L.phx_wake_workspace_bytes(_phx.MAX_KICKS, self.capacity, S)
        if self._ws is None:
            self._ws = torch.zeros(nb, dtype=torch.uint8, device=_phx.device())
        ws = self._ws
        _phx.call(L.phx_wake_convolve, self.record.data_ptr(), recs, n_kicks, S, max_turns,
                  out.data_ptr(), ws.data_ptr(), _phx.stream())
        return out


def apply_kicks(bunch, sliceset, records, kicks):
    L = _phx.lib()
    n_kicks = len(records)
    recs = (_phx.WakeKick * n_kicks)(*records)
    coords = _phx.ptr6([bunch.x, bunch.xp, bunch.y, bunch.yp, bunch.z, bunch.dp])
    _phx.call(L.phx_wake_apply, coords, bunch.macroparticlenumber,
              sliceset.slice_index_of_particle.data_ptr(), sliceset._state.data_ptr(), recs,
              n_kicks, sliceset.n_slices, kicks.data_ptr(), _phx.stream())


def _kick_class(name, moment, plane, multiply, doc):
    return type(name, (WakeKick,), dict(moment=moment, plane=plane, multiply=multiply,
                                        __doc__=doc))


ConstantWakeKickX = _kick_class('ConstantWakeKickX', _phx.MOM_N, _phx.PLANE_XP, -1,
                                'xp += W * n  (wake_kicks.py:144-162)')
ConstantWakeKickY = _kick_class('ConstantWakeKickY', _phx.MOM_N, _phx.PLANE_YP, -1,
                                'yp += W * n  (wake_kicks.py:165-182)')
ConstantWakeKickZ = _kick_class('ConstantWakeKickZ', _phx.MOM_N, _phx.PLANE_DP, -1,
                                'dp += W * n  (wake_kicks.py:185-202)')
DipoleWakeKickX = _kick_class('DipoleWakeKickX', _phx.MOM_X, _phx.PLANE_XP, -1,
                              'xp += W * n mean_x  (wake_kicks.py:207-223)')
DipoleWakeKickXY = _kick_class('DipoleWakeKickXY', _phx.MOM_Y, _phx.PLANE_XP, -1,
                               'xp += W * n mean_y  (wake_kicks.py:225-241)')
DipoleWakeKickY = _kick_class('DipoleWakeKickY', _phx.MOM_Y, _phx.PLANE_YP, -1,
                              'yp += W * n mean_y  (wake_kicks.py:244-260)')
DipoleWakeKickYX = _kick_class('DipoleWakeKickYX', _phx.MOM_X, _phx.PLANE_YP, -1,
                               'yp += W * n mean_x  (wake_kicks.py:262-279)')
QuadrupoleWakeKickX = _kick_class('QuadrupoleWakeKickX', _phx.MOM_N, _phx.PLANE_XP, 0,
                                  'xp += (W * n) x  (wake_kicks.py:284-301)')
QuadrupoleWakeKickXY = _kick_class('QuadrupoleWakeKickXY', _phx.MOM_N, _phx.PLANE_XP, 2,
                                   'xp += (W * n) y  (wake_kicks.py:304-321)')
QuadrupoleWakeKickY = _kick_class('QuadrupoleWakeKickY', _phx.MOM_N, _phx.PLANE_YP, 2,
                                  'yp += (W * n) y  (wake_kicks.py:324-341)')
QuadrupoleWakeKickYX = _kick_class('QuadrupoleWakeKickYX', _phx.MOM_N, _phx.PLANE_YP, 0,
                                   'yp += (W * n) x  (wake_kicks.py:344-361)')
