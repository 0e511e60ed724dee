"""Ideal transverse damper (PyHEADTAIL feedback/transverse_damper.py:15-94).

The bunch means come from one device reduction (phx_bunch_sums over xp, yp,
x, y) and the kick from phx_transverse_damper, which reads those sums on
the device: no host synchronisation per turn.
"""
import ctypes

import numpy as np
import torch

from .. import _phx
from ..general.element import Element


class TransverseDamper(Element):

    def __init__(self, dampingrate_x, dampingrate_y, phase=90, local_beta_function=None,
                 *args, **kwargs):
        """dampingrate_x, dampingrate_y: damping times in turns (gain =
        2 / rate; 0 or None switches the plane off); phase: of the kick
        with respect to the pick-up, in degrees (90: resistive);
        local_beta_function: needed unless the phase is 90 or 270."""
        planes = 0
        self.gain_x = self.gain_y = 0.
        if dampingrate_x:
            self.gain_x = 2 / dampingrate_x
            planes |= 1
        if dampingrate_y:
            self.gain_y = 2 / dampingrate_y
            planes |= 2
        self.prints({0: 'Dampers not active', 1: 'Damper in horizontal plane active',
                     2: 'Damper in vertical plane active', 3: 'Dampers active'}[planes])
        if phase != 90 and phase != 270 and not local_beta_function:
            raise TypeError(
                'TransverseDamper: numeric local_beta_function value at '
                'position of damper missing! (Required because of non-zero '
                'reactive damper component.)')
        self.phase_in_2pi = phase / 360. * 2 * np.pi
        self.local_beta_function = local_beta_function
        self._planes = planes
        self._sums = None

    @classmethod
    def horizontal(cls, dampingrate_x, *args, **kwargs):
        return cls(dampingrate_x, 0, *args, **kwargs)

    @classmethod
    def vertical(cls, dampingrate_y, *args, **kwargs):
        return cls(0, dampingrate_y, *args, **kwargs)

    def _record(self):
        p = _phx.Damper()
        s, co = np.sin(self.phase_in_2pi), np.cos(self.phase_in_2pi)
        p.coef_xp, p.coef_yp = self.gain_x * s, self.gain_y * s
        p.coef_x, p.coef_y = self.gain_x * co, self.gain_y * co
        p.beta = float(self.local_beta_function) if self.local_beta_function else 0.
        p.planes = self._planes
        return p

    def track(self, beam):
        if not self._planes:
            return
        L = _phx.lib()
        n = beam.macroparticlenumber
        if self._sums is None:
            self._sums = torch.empty(L.phx_bunch_sums_len(4), dtype=torch.float64,
                                     device=_phx.device())
        arrays = [beam.xp, beam.yp, beam.x, beam.y]
        ptrs = (_phx.P * 4)(*[a.data_ptr() for a in arrays])
        ws = _phx.workspace(n)
        _phx.call(L.phx_bunch_sums, ptrs, 4, n, self._sums.data_ptr(), ws.data_ptr(),
                  _phx.stream())
        p = self._record()
        _phx.call(L.phx_transverse_damper, beam.xp.data_ptr(), beam.yp.data_ptr(), n,
                  self._sums.data_ptr(), ctypes.byref(p), _phx.stream())
