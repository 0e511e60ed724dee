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
from abc import ABCMeta, abstractmethod
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


class WakeSource(Printing, metaclass=ABCMeta):
    @abstractmethod
    def get_wake_kicks(self, slicer):
        pass


class Resonator(WakeSource):
    def __init__(self, R_shunt, frequency, Q, Yokoya_X1, Yokoya_Y1, Yokoya_X2, Yokoya_Y2,
                 switch_Z, n_turns_wake=1, *args, **kwargs):
        self.R_shunt, self.frequency, self.Q = R_shunt, frequency, Q
        self.Yokoya_X1, self.Yokoya_X2 = Yokoya_X1, Yokoya_X2
        self.Yokoya_Y1, self.Yokoya_Y2 = Yokoya_Y1, Yokoya_Y2
        self.switch_Z = switch_Z
        self.n_turns_wake = n_turns_wake

    def function_transverse(self, Yokoya_factor):
        return ResonatorTransverseWake(Yokoya_factor, self.R_shunt, self.frequency, self.Q)

    def function_longitudinal(self):
        return ResonatorLongitudinalWake(self.R_shunt, self.frequency, self.Q)

    def get_wake_kicks(self, slicer):
        kicks = []
        if self.Yokoya_X1:
            kicks.append(DipoleWakeKickX(self.function_transverse(self.Yokoya_X1), slicer,
                                         self.n_turns_wake))
        if self.Yokoya_X2:
            kicks.append(QuadrupoleWakeKickX(self.function_transverse(self.Yokoya_X2), slicer,
                                             self.n_turns_wake))
        if self.Yokoya_Y1:
            kicks.append(DipoleWakeKickY(self.function_transverse(self.Yokoya_Y1), slicer,
                                         self.n_turns_wake))
        if self.Yokoya_Y2:
            kicks.append(QuadrupoleWakeKickY(self.function_transverse(self.Yokoya_Y2), slicer,
                                             self.n_turns_wake))
        if self.switch_Z:
            kicks.append(ConstantWakeKickZ(self.function_longitudinal(), slicer,
                                           self.n_turns_wake))
        return kicks


class CircularResonator(Resonator):
    def __init__(self, R_shunt, frequency, Q, n_turns_wake=1, *args, **kwargs):
        super(CircularResonator, self).__init__(R_shunt, frequency, Q, 1., 1., 0., 0., False,
                                                n_turns_wake, *args, **kwargs)


class ParallelHorizontalPlatesResonator(Resonator):
    def __init__(self, R_shunt, frequency, Q, n_turns_wake=1, *args, **kwargs):
        super(ParallelHorizontalPlatesResonator, self).__init__(
            R_shunt, frequency, Q, np.pi**2 / 24., np.pi**2 / 12., -np.pi**2 / 24.,
            np.pi**2 / 24., False, n_turns_wake, *args, **kwargs)


ParallelPlatesResonator = ParallelHorizontalPlatesResonator


class ParallelVerticalPlatesResonator(Resonator):
    def __init__(self, R_shunt, frequency, Q, n_turns_wake=1, *args, **kwargs):
        super(ParallelVerticalPlatesResonator, self).__init__(
            R_shunt, frequency, Q, np.pi**2 / 12., np.pi**2 / 24., np.pi**2 / 24.,
            -np.pi**2 / 24., False, n_turns_wake, *args, **kwargs)


class ResistiveWall(WakeSource):
    def __init__(self, pipe_radius, resistive_wall_length, conductivity, dt_min, Yokoya_X1,
                 Yokoya_Y1, Yokoya_X2, Yokoya_Y2, n_turns_wake=1, *args, **kwargs):
        self.pipe_radius = np.array([pipe_radius]).flatten()
        self.resistive_wall_length = resistive_wall_length
        self.conductivity = conductivity
        self.dt_min = dt_min
        self.Yokoya_X1, self.Yokoya_Y1 = Yokoya_X1, Yokoya_Y1
        self.Yokoya_X2, self.Yokoya_Y2 = Yokoya_X2, Yokoya_Y2
        self.n_turns_wake = n_turns_wake

    def function_transverse(self, Yokoya_factor):
        return ResistiveWallWake(Yokoya_factor, self.pipe_radius, self.resistive_wall_length,
                                 self.conductivity, self.dt_min)

    def get_wake_kicks(self, slicer):
        kicks = []
        if self.Yokoya_X1:
            kicks.append(DipoleWakeKickX(self.function_transverse(self.Yokoya_X1), slicer,
                                         self.n_turns_wake))
        if self.Yokoya_X2:
            kicks.append(QuadrupoleWakeKickX(self.function_transverse(self.Yokoya_X2), slicer,
                                             self.n_turns_wake))
        if self.Yokoya_Y1:
            kicks.append(DipoleWakeKickY(self.function_transverse(self.Yokoya_Y1), slicer,
                                         self.n_turns_wake))
        if self.Yokoya_Y2:
            kicks.append(QuadrupoleWakeKickY(self.function_transverse(self.Yokoya_Y2), slicer,
                                             self.n_turns_wake))
        return kicks


class CircularResistiveWall(ResistiveWall):
    def __init__(self, pipe_radius, resistive_wall_length, conductivity, dt_min,
                 n_turns_wake=1, *args, **kwargs):
        super(CircularResistiveWall, self).__init__(
            pipe_radius, resistive_wall_length, conductivity, dt_min, 1., 1., 0., 0.,
            n_turns_wake, *args, **kwargs)


class ParallelHorizontalPlatesResistiveWall(ResistiveWall):
    def __init__(self, pipe_radius, resistive_wall_length, conductivity, dt_min,
                 n_turns_wake=1, *args, **kwargs):
        super(ParallelHorizontalPlatesResistiveWall, self).__init__(
            pipe_radius, resistive_wall_length, conductivity, dt_min, np.pi**2 / 24.,
            np.pi**2 / 12., -np.pi**2 / 24., np.pi**2 / 24., n_turns_wake, *args, **kwargs)


ParallelPlatesResistiveWall = ParallelHorizontalPlatesResistiveWall


class ParallelVerticalPlatesResistiveWall(ResistiveWall):
    """Vertical plates.  (The reference derives this class from Resonator,
    wakes.py:665, which cannot be constructed with these arguments; here it
    is a resistive wall with the vertical-plates Yokoya factors.)"""

    def __init__(self, pipe_radius, resistive_wall_length, conductivity, dt_min,
                 n_turns_wake=1, *args, **kwargs):
        super(ParallelVerticalPlatesResistiveWall, self).__init__(
            pipe_radius, resistive_wall_length, conductivity, dt_min, np.pi**2 / 12.,
            np.pi**2 / 24., np.pi**2 / 24., -np.pi**2 / 24., n_turns_wake, *args, **kwargs)


class WakeField(Element):
    """Collective kick from the wake of the current and past slicings."""

    def __init__(self, slicer, *wake_sources):
        self.slicer = slicer
        self.wake_kicks = []
        for source in wake_sources:
            self.wake_kicks.extend(source.get_wake_kicks(self.slicer))
        n_turns_wake_max = max(source.n_turns_wake for source in wake_sources)
        self.n_turns_wake_max = n_turns_wake_max
        self.slice_set_deque = deque([], maxlen=n_turns_wake_max)
        self.slice_set_age_deque = deque([], maxlen=n_turns_wake_max)
        self._history = None

    def track(self, bunch):
        age_increment = bunch.circumference / (bunch.beta * c)
        for i in range(len(self.slice_set_age_deque)):
            self.slice_set_age_deque[i] += age_increment
        slice_set = bunch.get_slices(self.slicer, statistics=['mean_x', 'mean_y'])
        self.slice_set_deque.appendleft(slice_set)
        self.slice_set_age_deque.appendleft(0.)
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
