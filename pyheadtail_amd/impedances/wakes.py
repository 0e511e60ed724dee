"""Wake fields and wake sources.

Contract of PyHEADTAIL impedances/wakes.py:
  - WakeField(slicer, *wake_sources) (:77-145): collects the kicks of all
    sources; track(bunch) ages the stored slicings by C/(beta c), slices the
    bunch with statistics mean_x, mean_y, stores the SliceSet (deque of
    maxlen = max n_turns_wake) and applies every kick;
  - Resonator(R_shunt, frequency, Q, Yokoya_X1, Yokoya_Y1, Yokoya_X2,
    Yokoya_Y2, switch_Z, n_turns_wake) with CircularResonator and the
    parallel-plates variants (:379-548); Alex Chao's resonator wake
    functions (:450-493);
  - ResistiveWall(pipe_radius, resistive_wall_length, conductivity, dt_min,
    Yokoya_X1, ..., n_turns_wake) with the circular and parallel-plates
    variants (:551-678) and its clipped 1/sqrt(-dt) wake (:611-624).

Wake functions are objects: called with (dt, beta=...) they evaluate the
reference's formula on the host (numpy); fill(record) writes the host-side
constants of that formula into a phx_wake_kick record for the device.
The WakeField keeps its SliceSet history on the device (WakeHistory).
"""
from collections import deque

import numpy as np
from scipy.constants import c

from .. import _phx
from ..general.element import Element, Printing
from .wake_kicks import (WakeHistory, apply_kicks, ConstantWakeKickZ, DipoleWakeKickX,
                         DipoleWakeKickY, QuadrupoleWakeKickX, QuadrupoleWakeKickY)
from .wake_kicks import *  # noqa: F401,F403  (kick classes, as the reference re-exports)


class ResonatorTransverseWake(object):
    """W(dt) = Y R w^2/(Q wbar) e^{a dt} sin(wbar dt), dt clipped at 0
    (Q > 0.5; Q == 0.5 and Q < 0.5 branches as in wakes.py:458-469)."""

    def __init__(self, yokoya, R_shunt, frequency, Q):
        self.yokoya, self.R_shunt, self.frequency, self.Q = yokoya, R_shunt, frequency, Q
        omega = 2 * np.pi * frequency
        self.alpha = omega / (2 * Q)
        self.omegabar = np.sqrt(np.abs(omega**2 - self.alpha**2))
        if Q == 0.5:
            self.prefix = yokoya * R_shunt * omega**2 / Q
        else:
            self.prefix = yokoya * R_shunt * omega**2 / (Q * self.omegabar)

    def __call__(self, dt, *args, **kwargs):
        dt = np.asarray(dt, dtype=np.float64).clip(max=0)
        ea = self.prefix * np.exp(self.alpha * dt)
        if self.Q > 0.5:
            return ea * np.sin(self.omegabar * dt)
        if self.Q == 0.5:
            return ea * dt
        return ea * np.sinh(self.omegabar * dt)

    def fill(self, k):
        k.fn = _phx.WAKE_RES_TRANSVERSE
        k.Q, k.prefix, k.alpha, k.omegabar = self.Q, self.prefix, self.alpha, self.omegabar


class ResonatorLongitudinalWake(object):
    """W(dt) = -(sign(dt)-1) R a e^{a dt} (cos(wbar dt) + a/wbar sin(wbar dt))
    (wakes.py:472-493)."""

    def __init__(self, R_shunt, frequency, Q):
        self.R_shunt, self.Q = R_shunt, Q
        omega = 2 * np.pi * frequency
        self.alpha = omega / (2 * Q)
        self.omegabar = np.sqrt(np.abs(omega**2 - self.alpha**2))

    def __call__(self, dt, *args, **kwargs):
        dt = np.asarray(dt, dtype=np.float64)
        a, wb = self.alpha, self.omegabar
        t = -(np.sign(dt) - 1) * self.R_shunt * a * np.exp(a * dt)
        if self.Q > 0.5:
            return t * (np.cos(wb * dt) + a / wb * np.sin(wb * dt))
        if self.Q == 0.5:
            return t * (1. + a * dt)
        return t * (np.cosh(wb * dt) + a / wb * np.sinh(wb * dt))

    def fill(self, k):
        k.fn = _phx.WAKE_RES_LONGITUDINAL
        k.Q, k.prefix, k.alpha, k.omegabar = self.Q, self.R_shunt, self.alpha, self.omegabar
        k.alpha_over_omegabar = self.alpha / self.omegabar


class ResistiveWallWake(object):
    """W(dt) = Y (sign(dt+|dt_min|)-1)/2 sqrt(beta) L/pi/b^3
    sqrt(-mu_r/(pi sigma dt_c)) sqrt(Z0 c), dt_c = min(dt, -|dt_min|),
    with beta of the source slicing (wakes.py:611-624)."""

    mu_r = 1
    Z_0 = 119.9169832 * np.pi

    def __init__(self, yokoya, pipe_radius, length, conductivity, dt_min):
        self.yokoya = yokoya
        self.pipe_radius = np.array([pipe_radius]).flatten()
        self.length, self.conductivity, self.dt_min = length, conductivity, dt_min

    def __call__(self, dt, *args, **kwargs):
        dt = np.asarray(dt, dtype=np.float64)
        return (self.yokoya * (np.sign(dt + np.abs(self.dt_min)) - 1) / 2. *
                np.sqrt(kwargs['beta']) * self.length / np.pi / self.pipe_radius**3 *
                np.sqrt(-self.mu_r / np.pi / self.conductivity /
                        dt.clip(max=-abs(self.dt_min)))) * np.sqrt(self.Z_0 * c)

    def fill(self, k):
        k.fn = _phx.WAKE_RESISTIVE_WALL
        k.yokoya = self.yokoya
        k.L = self.length
        k.b3 = float((self.pipe_radius**3)[0])
        k.dt_min_abs = np.abs(self.dt_min)
        k.k_sigma = -self.mu_r / np.pi / self.conductivity
        k.sqrt_z0c = np.sqrt(self.Z_0 * c)


# Yokoya factors (X1, Y1, X2, Y2) of the standard chamber geometries
_PI2 = np.pi**2
YOKOYA_FACTORS = {
    'circular': (1., 1., 0., 0.),
    'horizontal plates': (_PI2 / 24., _PI2 / 12., -_PI2 / 24., _PI2 / 24.),
    'vertical plates': (_PI2 / 12., _PI2 / 24., _PI2 / 24., -_PI2 / 24.),
}

# (Yokoya attribute, kick class), in the order the kicks are produced
_TRANSVERSE_KICKS = (('Yokoya_X1', DipoleWakeKickX), ('Yokoya_X2', QuadrupoleWakeKickX),
                     ('Yokoya_Y1', DipoleWakeKickY), ('Yokoya_Y2', QuadrupoleWakeKickY))


class WakeSource(Printing):
    """A source of wake kicks: one kick per non-zero Yokoya factor, each
    with its own wake function object."""

    def get_wake_kicks(self, slicer):
        kicks = [kick_class(self.function_transverse(getattr(self, factor)), slicer,
                            self.n_turns_wake)
                 for factor, kick_class in _TRANSVERSE_KICKS if getattr(self, factor)]
        if getattr(self, 'switch_Z', False):
            kicks.append(ConstantWakeKickZ(self.function_longitudinal(), slicer,
                                           self.n_turns_wake))
        return kicks

    def _set_yokoya(self, X1, Y1, X2, Y2):
        self.Yokoya_X1, self.Yokoya_Y1, self.Yokoya_X2, self.Yokoya_Y2 = X1, Y1, X2, Y2


class Resonator(WakeSource):
    def __init__(self, R_shunt, frequency, Q, Yokoya_X1, Yokoya_Y1, Yokoya_X2, Yokoya_Y2,
                 switch_Z, n_turns_wake=1, *args, **kwargs):
        self.R_shunt, self.frequency, self.Q = R_shunt, frequency, Q
        self._set_yokoya(Yokoya_X1, Yokoya_Y1, Yokoya_X2, Yokoya_Y2)
        self.switch_Z, self.n_turns_wake = switch_Z, n_turns_wake

    def function_transverse(self, Yokoya_factor):
        return ResonatorTransverseWake(Yokoya_factor, self.R_shunt, self.frequency, self.Q)

    def function_longitudinal(self):
        return ResonatorLongitudinalWake(self.R_shunt, self.frequency, self.Q)


class ResistiveWall(WakeSource):
    def __init__(self, pipe_radius, resistive_wall_length, conductivity, dt_min, Yokoya_X1,
                 Yokoya_Y1, Yokoya_X2, Yokoya_Y2, n_turns_wake=1, *args, **kwargs):
        self.pipe_radius = np.array([pipe_radius]).flatten()
        self.resistive_wall_length, self.conductivity = resistive_wall_length, conductivity
        self.dt_min, self.n_turns_wake = dt_min, n_turns_wake
        self._set_yokoya(Yokoya_X1, Yokoya_Y1, Yokoya_X2, Yokoya_Y2)

    def function_transverse(self, Yokoya_factor):
        return ResistiveWallWake(Yokoya_factor, self.pipe_radius, self.resistive_wall_length,
                                 self.conductivity, self.dt_min)


class CircularResonator(Resonator):
    def __init__(self, R_shunt, frequency, Q, n_turns_wake=1, *args, **kwargs):
        Resonator.__init__(self, R_shunt, frequency, Q, *YOKOYA_FACTORS['circular'], False,
                           n_turns_wake, *args, **kwargs)


class ParallelHorizontalPlatesResonator(Resonator):
    def __init__(self, R_shunt, frequency, Q, n_turns_wake=1, *args, **kwargs):
        Resonator.__init__(self, R_shunt, frequency, Q, *YOKOYA_FACTORS['horizontal plates'],
                           False, n_turns_wake, *args, **kwargs)


class ParallelVerticalPlatesResonator(Resonator):
    def __init__(self, R_shunt, frequency, Q, n_turns_wake=1, *args, **kwargs):
        Resonator.__init__(self, R_shunt, frequency, Q, *YOKOYA_FACTORS['vertical plates'],
                           False, n_turns_wake, *args, **kwargs)


class CircularResistiveWall(ResistiveWall):
    def __init__(self, pipe_radius, resistive_wall_length, conductivity, dt_min,
                 n_turns_wake=1, *args, **kwargs):
        ResistiveWall.__init__(self, pipe_radius, resistive_wall_length, conductivity, dt_min,
                               *YOKOYA_FACTORS['circular'], n_turns_wake, *args, **kwargs)


class ParallelHorizontalPlatesResistiveWall(ResistiveWall):
    def __init__(self, pipe_radius, resistive_wall_length, conductivity, dt_min,
                 n_turns_wake=1, *args, **kwargs):
        ResistiveWall.__init__(self, pipe_radius, resistive_wall_length, conductivity, dt_min,
                               *YOKOYA_FACTORS['horizontal plates'], n_turns_wake, *args,
                               **kwargs)


class ParallelVerticalPlatesResistiveWall(ResistiveWall):
    """Vertical plates.  (The reference derives this class from Resonator,
    wakes.py:665, which cannot be constructed with these arguments; here it
    is a resistive wall with the vertical-plates Yokoya factors.)"""

    def __init__(self, pipe_radius, resistive_wall_length, conductivity, dt_min,
                 n_turns_wake=1, *args, **kwargs):
        ResistiveWall.__init__(self, pipe_radius, resistive_wall_length, conductivity, dt_min,
                               *YOKOYA_FACTORS['vertical plates'], n_turns_wake, *args,
                               **kwargs)


ParallelPlatesResonator = ParallelHorizontalPlatesResonator
ParallelPlatesResistiveWall = ParallelHorizontalPlatesResistiveWall


class WakeField(Element):
    """Collective kick from the wake of the current and past slicings."""

    PRECISIONS = {'fp64': _phx.PREC_FP64, 'fp32': _phx.PREC_FP32}

    def __init__(self, slicer, *wake_sources, precision='fp64'):
        """precision (not in PyHEADTAIL): 'fp32' evaluates the multi-turn
        resistive-wall convolutions with FP32 wake values and products
        (summed in FP32 over runs of n_slices/8 sources, in FP64 beyond);
        every other convolution stays FP64."""
        if precision not in self.PRECISIONS:
            raise ValueError("precision must be one of %s" % sorted(self.PRECISIONS))
        self.slicer = slicer
        self.precision = precision
        self.wake_kicks = [kick for src in wake_sources for kick in src.get_wake_kicks(slicer)]
        for kick in self.wake_kicks:
            kick.precision = self.PRECISIONS[precision]
        self.n_turns_wake_max = max(src.n_turns_wake for src in wake_sources)
        # the remembered slicings, newest first, and their ages in seconds
        self.slice_set_deque = deque(maxlen=self.n_turns_wake_max)
        self.slice_set_age_deque = deque(maxlen=self.n_turns_wake_max)
        self._history = None

    def _remember(self, slice_set, turn_time):
        older = [age + turn_time for age in self.slice_set_age_deque]
        self.slice_set_age_deque.clear()
        self.slice_set_age_deque.extend(older)
        self.slice_set_age_deque.appendleft(0.)
        self.slice_set_deque.appendleft(slice_set)

    def track(self, bunch):
        age_increment = bunch.circumference / (bunch.beta * c)
        slice_set = bunch.get_slices(self.slicer, statistics=['mean_x', 'mean_y'])
        self._remember(slice_set, age_increment)
        if not self.wake_kicks:
            return
        if self._history is None:
            self._history = WakeHistory(slice_set.n_slices, self.n_turns_wake_max)
        self._history.push_sliceset(slice_set, age_increment)
        records = [k.record(bunch) for k in self.wake_kicks]
        # every kick convolves the moments of this slicing, so groups of at
        # most MAX_KICKS applied in list order give the same result as one
        for i in range(0, len(records), _phx.MAX_KICKS):
            group = records[i:i + _phx.MAX_KICKS]
            kicks = self._history.convolve(group)
            apply_kicks(bunch, slice_set, group, kicks)
