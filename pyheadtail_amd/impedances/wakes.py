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
import torch
from scipy.constants import c

from .. import _phx
from ..general.element import Element, Printing
from .wake_kicks import (WakeHistory, apply_kicks, ConstantWakeKickZ, DipoleWakeKickX,
                         DipoleWakeKickY, QuadrupoleWakeKickX, QuadrupoleWakeKickY)
from .wake_kicks import *  # noqa: F401,F403  (kick classes, as the reference re-exports)


class TableWake(object):
    """Linear interpolation of a wake table at -dt, as interp1d (which hands
    FP64 data to np.interp) in WakeTable.function_transverse /
    function_longitudinal (wakes.py:290-376).  time [s] and strength [SI]
    are the converted table columns; kind is one of _phx.TABLE_*."""

    def __init__(self, time, strength, kind, name):
        self.time = np.ascontiguousarray(time, dtype=np.float64)
        self.strength = np.ascontiguousarray(strength, dtype=np.float64)
        self.kind, self.name = kind, name
        self._dev = None

    def _interp(self, x, outside):
        y = np.interp(x, self.time, self.strength)
        out = (x < self.time[0]) | (x > self.time[-1])
        if outside is None:
            if np.any(out):
                raise ValueError("%s: a value of -dt is outside the wake table's range "
                                 "[%g, %g] s" % (self.name, self.time[0], self.time[-1]))
        else:
            y = np.where(out, outside, y)
        return y

    def __call__(self, dt, *args, **kwargs):
        dt = np.asarray(dt, dtype=np.float64)
        if self.kind == _phx.TABLE_TRANS_UR:
            return self._interp(-dt.clip(max=0), None)
        if self.kind == _phx.TABLE_TRANS_LOWBETA:
            return self._interp(-dt, None)
        w = self._interp(-dt, 0.)
        if self.kind == _phx.TABLE_LONG_BL:
            return (np.sign(-dt) + 1.) / 2. * w
        return w

    def fill(self, k):
        if self._dev is None:
            dev = _phx.device()
            self._dev = (torch.from_numpy(self.time).to(dev),
                         torch.from_numpy(self.strength).to(dev))
        k.fn = _phx.WAKE_TABLE
        k.table_t, k.table_w = self._dev[0].data_ptr(), self._dev[1].data_ptr()
        k.table_n, k.table_kind = len(self.time), self.kind


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


class WakeTable(WakeSource):
    """Wake functions from a table (wakes.py:167-376): columns named by
    wake_file_columns, 'time' [ns] among them; transverse components
    [V/pC/mm], 'longitudinal' [V/pC]; HEADTAIL signs.  The component names
    and the order of the kicks are the reference's.  quadrupole_xy gives
    its kick like every other component (the reference appends it to a
    list that does not exist)."""

    COMPONENTS = ('constant_x', 'constant_y', 'longitudinal', 'dipole_x', 'dipole_y',
                  'dipole_xy', 'dipole_yx', 'quadrupole_x', 'quadrupole_y', 'quadrupole_xy',
                  'quadrupole_yx')

    def __init__(self, wake_file, wake_file_columns, n_turns_wake=1, *args, **kwargs):
        self.wake_table = {}
        wake_data = np.loadtxt(wake_file)
        if len(wake_file_columns) != wake_data.shape[1]:
            raise ValueError("Length of wake_file_columns list does not correspond to the "
                             "number of columns in the specified wake_file.")
        if 'time' not in wake_file_columns:
            raise ValueError("No wake_file_column with name 'time' has been specified.")
        for i, name in enumerate(wake_file_columns):
            self.wake_table[name] = wake_data[:, i]
        self.n_turns_wake = n_turns_wake

    def _is_provided(self, wake_component):
        return wake_component in self.wake_table

    def get_wake_kicks(self, slicer):
        kinds = {'constant_x': ConstantWakeKickX, 'constant_y': ConstantWakeKickY,
                 'longitudinal': ConstantWakeKickZ, 'dipole_x': DipoleWakeKickX,
                 'dipole_y': DipoleWakeKickY, 'dipole_xy': DipoleWakeKickXY,
                 'dipole_yx': DipoleWakeKickYX, 'quadrupole_x': QuadrupoleWakeKickX,
                 'quadrupole_y': QuadrupoleWakeKickY, 'quadrupole_xy': QuadrupoleWakeKickXY,
                 'quadrupole_yx': QuadrupoleWakeKickYX}
        kicks = []
        for comp in self.COMPONENTS:
            if not self._is_provided(comp):
                continue
            fn = (self.function_longitudinal() if comp == 'longitudinal'
                  else self.function_transverse(comp))
            kicks.append(kinds[comp](fn, slicer, self.n_turns_wake))
        return kicks

    def function_transverse(self, wake_component):
        """ns -> s, V/pC/mm -> V/C/m with the sign flip; ultrarelativistic
        (table starts at (0, 0): dt clipped at 0) or low-beta (table starts
        below 0) wake (wakes.py:290-339)."""
        time = 1e-9 * self.wake_table['time']
        strength = -1e15 * self.wake_table[wake_component]
        if time[0] == 0 and strength[0] == 0:
            self.prints(wake_component + ' Assuming ultrarelativistic wake.')
            return TableWake(time, strength, _phx.TABLE_TRANS_UR, wake_component)
        if time[0] < 0:
            self.prints(wake_component + ' Found low beta wake.')
            return TableWake(time, strength, _phx.TABLE_TRANS_LOWBETA, wake_component)
        raise ValueError(wake_component + ' does not meet requirements.')

    def function_longitudinal(self):
        """ns -> s, V/pC -> V/C; zero outside the table; half the value at
        dt = 0 when the table starts at 0 (wakes.py:341-376)."""
        time = 1e-9 * self.wake_table['time']
        strength = 1e12 * self.wake_table['longitudinal']
        if time[0] == 0:
            return TableWake(time, strength, _phx.TABLE_LONG_BL, 'longitudinal')
        if time[0] < 0:
            return TableWake(time, strength, _phx.TABLE_LONG, 'longitudinal')
        raise ValueError('Longitudinal wake component does not meet requirements.')


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

    PRECISIONS = {'fp64': _phx.PREC_FP64, 'fp32': _phx.PREC_FP32,
                  'fp64_direct': _phx.PREC_FP64_DIRECT, 'mfma_fp32': _phx.PREC_MFMA_F32,
                  'mfma_bf16x3': _phx.PREC_MFMA_BF16X3, 'mfma_bf16x6': _phx.PREC_MFMA_BF16X6}

    def __init__(self, slicer, *wake_sources, precision='fp64'):
        """precision (not in PyHEADTAIL) of the multi-turn resistive-wall
        convolutions; every other convolution stays FP64.
          'fp64'        FP64; past turns by a separable expansion of the
                        wall's (-dt)^(-1/2), exact to below 1e-18 relative,
                        the current turn pair by pair;
          'fp64_direct' FP64, every pair (the reference's sums);
          'fp32'        FP32 wake values and products (summed in FP32 over
                        runs of n_slices/8 sources, in FP64 beyond);
          'mfma_fp32'   every pair on the matrix cores, FP32 MFMA with FP32
                        accumulation over 256 sources (FP64 beyond);
          'mfma_bf16x3' the same with bfloat16 MFMA on wake values and
                        moments split in two bfloat16 parts (three
                        products, about 16 significant bits);
          'mfma_bf16x6' three parts each, six products."""
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
            self._check_table_range(group, kicks)
            apply_kicks(bunch, slice_set, group, kicks)

    def _check_table_range(self, records, kicks):
        """Transverse wake tables give NaN outside their range, where the
        reference's interp1d raises ValueError: raise it here (one host
        synchronisation, only for such tables)."""
        rows = [i for i, r in enumerate(records) if r.fn == _phx.WAKE_TABLE and
                r.table_kind in (_phx.TABLE_TRANS_UR, _phx.TABLE_TRANS_LOWBETA)]
        if rows and not bool(torch.isfinite(kicks[rows]).all()):
            raise ValueError("wake table: a value of -dt is outside the table's range")
