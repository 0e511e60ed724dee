"""CPU restatement of PyHEADTAIL 1.16.5's per-turn path (test infrastructure).

This module is the oracle of the parity tests: only tests/, the smoke check
and the CPU-baseline leg of bench.py use it, and only as the checker.  It
restates the reference's CPU arithmetic operation by operation in NumPy:

  bunch statistics   np.mean (general/pmath.py:183); cov / std / emittance
                     (cobra_functions/stats.pyx:17-122) as sequential
                     shifted sums (np.cumsum runs in index order, like the
                     Cython loop)
  slicing            get_long_cuts + compute_sliceset_kwargs
                     (particles/slicing.py:395-413, 599-615)
  per-slice stats    count / mean / std / cov / emittance per slice
                     (stats.pyx:192-425); np.bincount accumulates in
                     particle order, like the Cython loops
  transverse map     TransverseSegmentMap.track (transverse_tracking.py:
                     71-196) with ChromaticitySegment / AmplitudeDetuning
                     (detuners.py:64-151)
  longitudinal       Drift, Kick, RFSystems (longitudinal_tracking.py:
                     84-203, 362-369, 556-574), LinearMap (:720-759)
  wakes              resonator and resistive-wall wake functions
                     (impedances/wakes.py:450-493, 611-624); WakeKick
                     convolution (np.convolve 'valid' or np.dot) and kick
                     (wake_kicks.py:66-361); WakeField.track (wakes.py:121-145)

Every function cites the reference lines it follows.  Nothing here calls the
product package.
"""
from collections import deque
from math import factorial

import numpy as np
from scipy.constants import c, e


# ---------------------------------------------------------------------------
# bunch and statistics
# ---------------------------------------------------------------------------
class Bunch(object):
    """Host bunch: six float64 columns and the reference's scalar setters
    (particles/particles.py:21-103)."""

    def __init__(self, macroparticlenumber, particlenumber_per_mp, charge, mass,
                 circumference, gamma, coords):
        self.macroparticlenumber = macroparticlenumber
        self.particlenumber_per_mp = particlenumber_per_mp
        self.charge, self.mass, self.circumference = charge, mass, circumference
        self.gamma = gamma
        for k, v in coords.items():
            setattr(self, k, np.array(v, dtype=np.float64, copy=True))

    @property
    def gamma(self):
        return self._gamma

    @gamma.setter
    def gamma(self, value):
        self._gamma = value
        self._beta = np.sqrt(1 - self.gamma**-2)
        self._betagamma = np.sqrt(self.gamma**2 - 1)
        self._p0 = self.betagamma * self.mass * c

    @property
    def beta(self):
        return self._beta

    @property
    def betagamma(self):
        return self._betagamma

    @property
    def p0(self):
        return self._p0

    @p0.setter
    def p0(self, value):
        self.gamma = np.sqrt(1 + (value / (self.mass * c))**2)

    @property
    def intensity(self):
        return self.particlenumber_per_mp * self.macroparticlenumber

    @property
    def charge_per_mp(self):
        return self.particlenumber_per_mp * self.charge

    def copy(self):
        return Bunch(self.macroparticlenumber, self.particlenumber_per_mp, self.charge,
                     self.mass, self.circumference, self.gamma,
                     {k: getattr(self, k) for k in ('x', 'xp', 'y', 'yp', 'z', 'dp')})


def cov(a, b):
    """Shifted single-pass covariance, shift = first element, /(n-1)
    (stats.pyx:17-45)."""
    n = len(a)
    if n < 2:
        return 0.
    da = a - a[0]
    db = b - b[0]
    a_sum = np.cumsum(da)[-1]
    b_sum = np.cumsum(db)[-1]
    ab_sum = np.cumsum(da * db)[-1]
    return (ab_sum - a_sum * b_sum / n) / (n - 1)


def std(u):
    return np.sqrt(cov(u, u))


def emittance(u, up, dp):
    """stats.pyx:84-122."""
    cov_u2, cov_up2, cov_u_up = cov(u, u), cov(up, up), cov(up, u)
    t11 = t12 = t22 = 0.
    if dp is not None:
        cov_dp2 = cov(dp, dp)
        if cov_dp2 != 0:
            cov_u_dp, cov_up_dp = cov(u, dp), cov(up, dp)
            t11 = cov_u_dp * cov_u_dp / cov_dp2
            t12 = cov_u_dp * cov_up_dp / cov_dp2
            t22 = cov_up_dp * cov_up_dp / cov_dp2
    s11, s12, s22 = cov_u2 - t11, cov_u_up - t12, cov_up2 - t22
    return np.sqrt(s11 * s22 - s12 * s12)


def bunch_statistics(b):
    """mean_*, sigma_*, epsn_* of Particles (particles.py:311-366)."""
    out = {}
    for k in ('x', 'xp', 'y', 'yp', 'z', 'dp'):
        out['mean_' + k] = np.mean(getattr(b, k))
        out['sigma_' + k] = std(getattr(b, k))
    out['epsn_x'] = emittance(b.x, b.xp, b.dp) * b.betagamma
    out['epsn_y'] = emittance(b.y, b.yp, b.dp) * b.betagamma
    out['epsn_z'] = 4 * np.pi * emittance(b.z, b.dp, None) * b.p0 / e
    return out


# ---------------------------------------------------------------------------
# slicing
# ---------------------------------------------------------------------------
class SliceData(object):
    """What a SliceSet holds (particles/slicing.py:58-126, 384-393)."""

    def __init__(self, z_bins, idx, beam):
        self.z_bins = z_bins
        self.slice_index_of_particle = idx
        self.n_slices = len(z_bins) - 1
        inside = (idx < self.n_slices) & (idx >= 0)
        self.particles_within_cuts = np.where(inside)[0].astype(np.int32)
        self.n_macroparticles_per_slice = np.bincount(
            idx[self.particles_within_cuts], minlength=self.n_slices).astype(np.int32)
        self.beta, self.gamma, self.p0 = beam.beta, beam.gamma, beam.p0
        self.age = 0

    @property
    def z_centers(self):
        return self.z_bins[:-1] + 0.5 * (self.z_bins[1:] - self.z_bins[:-1])

    def convert_to_time(self, z):
        return z / (self.beta * c)


def long_cuts(beam, n_sigma_z=None, z_cuts=None):
    """Slicer.get_long_cuts (slicing.py:395-413)."""
    if z_cuts is not None:
        return z_cuts
    if n_sigma_z:
        return (np.mean(beam.z) - n_sigma_z * std(beam.z),
                np.mean(beam.z) + n_sigma_z * std(beam.z))
    tail, head = np.min(beam.z), np.max(beam.z)
    head += abs(head) * 1e-15
    return tail, head


def slice_from_cuts(beam, n_slices, tail, head):
    """compute_sliceset_kwargs with given cuts (slicing.py:604-615)."""
    width = (head - tail) / float(n_slices)
    z_bins = np.arange(tail, head + 1e-7 * width, width, dtype=np.float64)
    with np.errstate(invalid='ignore'):
        idx = np.floor((beam.z - tail) / width).astype(np.int32)
    return SliceData(z_bins, idx, beam)


def slice_uniform(beam, n_slices, n_sigma_z=None, z_cuts=None):
    tail, head = long_cuts(beam, n_sigma_z, z_cuts)
    return slice_from_cuts(beam, n_slices, tail, head)


def _per_slice_sum(sd, values):
    p = sd.particles_within_cuts
    return np.bincount(sd.slice_index_of_particle[p], weights=values[p],
                       minlength=sd.n_slices)


def mean_per_slice(sd, u):
    """stats.pyx:235-253."""
    s = _per_slice_sum(sd, u)
    n = sd.n_macroparticles_per_slice
    out = s.copy()
    nz = n != 0
    out[nz] = s[nz] / n[nz]
    return out


def cov_per_slice(sd, a, b):
    """Two-pass per-slice covariance, /(n-1) for n > 1 (stats.pyx:290-319)."""
    ma, mb = mean_per_slice(sd, a), mean_per_slice(sd, b)
    idx = sd.slice_index_of_particle
    p = sd.particles_within_cuts
    da = np.zeros(len(a))
    db = np.zeros(len(b))
    da[p] = a[p] - ma[idx[p]]
    db[p] = b[p] - mb[idx[p]]
    s = _per_slice_sum(sd, da * db)
    n = sd.n_macroparticles_per_slice
    big = n > 1
    s[big] = s[big] / (n[big] - 1)
    return s


def std_per_slice(sd, u):
    """stats.pyx:258-286."""
    return np.sqrt(cov_per_slice(sd, u, u))


def emittance_per_slice(sd, u, up, dp=None):
    """stats.pyx:374-425."""
    cu2, cuup, cup2 = cov_per_slice(sd, u, u), cov_per_slice(sd, u, up), cov_per_slice(sd, up, up)
    n = sd.n_macroparticles_per_slice
    out = np.zeros(sd.n_slices)
    if dp is not None:
        cudp, cupdp, cdp2 = (cov_per_slice(sd, u, dp), cov_per_slice(sd, up, dp),
                             cov_per_slice(sd, dp, dp))
    for i in range(sd.n_slices):
        if n[i] > 1:
            s11, s12, s22 = cu2[i], cuup[i], cup2[i]
            if dp is not None and cdp2[i] != 0.:
                s11 -= cudp[i] * cudp[i] / cdp2[i]
                s12 -= cudp[i] * cupdp[i] / cdp2[i]
                s22 -= cupdp[i] * cupdp[i] / cdp2[i]
            out[i] = np.sqrt(s11 * s22 - s12 * s12)
    return out


# ---------------------------------------------------------------------------
# transverse
# ---------------------------------------------------------------------------
def segment_matrices(ax0, bx0, ax1, bx1, ay0, by0, ay1, by1):
    """I (cosine) and J (sine) parts (transverse_tracking.py:71-103)."""
    I = np.zeros((4, 4))
    J = np.zeros((4, 4))
    I[0, 0] = np.sqrt(bx1 / bx0)
    I[1, 0] = np.sqrt(1. / (bx0 * bx1)) * (ax0 - ax1)
    I[1, 1] = np.sqrt(bx0 / bx1)
    I[2, 2] = np.sqrt(by1 / by0)
    I[3, 2] = np.sqrt(1. / (by0 * by1)) * (ay0 - ay1)
    I[3, 3] = np.sqrt(by0 / by1)
    J[0, 0] = np.sqrt(bx1 / bx0) * ax0
    J[0, 1] = np.sqrt(bx0 * bx1)
    J[1, 0] = -(np.sqrt(1. / (bx0 * bx1)) * (1. + ax0 * ax1))
    J[1, 1] = -np.sqrt(bx0 / bx1) * ax1
    J[2, 2] = np.sqrt(by1 / by0) * ay0
    J[2, 3] = np.sqrt(by0 * by1)
    J[3, 2] = -(np.sqrt(1. / (by0 * by1)) * (1. + ay0 * ay1))
    J[3, 3] = -np.sqrt(by0 / by1) * ay1
    return I, J


class Chroma(object):
    """ChromaticitySegment (detuners.py:53-98), Qp already scaled by dmu."""

    def __init__(self, dQp_x, dQp_y):
        self.cx = [q / float(factorial(i + 1)) for i, q in enumerate(np.atleast_1d(dQp_x))]
        self.cy = [q / float(factorial(i + 1)) for i, q in enumerate(np.atleast_1d(dQp_y))]

    @staticmethod
    def _poly(cf, dp):
        if len(cf) == 1:
            return cf[0] * dp
        if len(cf) == 2:
            return (cf[0] + cf[1] * dp) * dp
        if len(cf) == 3:
            return (cf[0] + (cf[1] + cf[2] * dp) * dp) * dp
        from numpy.polynomial.polynomial import polyval
        return polyval(dp, [0] + list(cf))

    def detune(self, b):
        return self._poly(self.cx, b.dp), self._poly(self.cy, b.dp)


class Amplitude(object):
    """AmplitudeDetuningSegment (detuners.py:106-151)."""

    def __init__(self, dapp_x, dapp_y, dapp_xy, dapp_yx, ax, bx, ay, by):
        self.d = (dapp_x, dapp_y, dapp_xy, dapp_yx)
        self.o = (ax, bx, ay, by)

    def detune(self, b):
        ax, bx, ay, by = self.o
        dx, dy, dxy, dyx = self.d
        Jx = 1 / 2 * ((1 + ax**2) / bx * b.x**2 + 2 * ax * b.x * b.xp + bx * b.xp**2)
        Jy = 1 / 2 * ((1 + ay**2) / by * b.y**2 + 2 * ay * b.y * b.yp + by * b.yp**2)
        return (dx * Jx + dxy * Jy) / b.p0, (dy * Jy + dyx * Jx) / b.p0


class TransverseSegment(object):
    """TransverseSegmentMap.track (transverse_tracking.py:138-196)."""

    def __init__(self, ax0, bx0, Dx0, ax1, bx1, Dx1, ay0, by0, Dy0, ay1, by1, Dy1,
                 dQ_x, dQ_y, detuners=()):
        self.I, self.J = segment_matrices(ax0, bx0, ax1, bx1, ay0, by0, ay1, by1)
        self.D = (Dx0, Dx1, Dy0, Dy1)
        self.dQ_x, self.dQ_y = dQ_x, dQ_y
        self.detuners = list(detuners)
        self.dispersion = not np.allclose(self.D, np.zeros(4), atol=1e-3)

    def track(self, b):
        dphi_x, dphi_y = self.dQ_x, self.dQ_y
        for d in self.detuners:
            qx, qy = d.detune(b)
            dphi_x = dphi_x + qx
            dphi_y = dphi_y + qy
        dphi_x = dphi_x * (2. * np.pi)
        dphi_y = dphi_y * (2. * np.pi)
        if self.detuners:
            s_x, c_x = np.sin(np.atleast_1d(dphi_x)), np.cos(np.atleast_1d(dphi_x))
            s_y, c_y = np.sin(np.atleast_1d(dphi_y)), np.cos(np.atleast_1d(dphi_y))
        else:
            c_x, c_y, s_x, s_y = np.cos(dphi_x), np.cos(dphi_y), np.sin(dphi_x), np.sin(dphi_y)
        I, J = self.I, self.J
        M00, M01 = I[0, 0] * c_x + J[0, 0] * s_x, I[0, 1] * c_x + J[0, 1] * s_x
        M10, M11 = I[1, 0] * c_x + J[1, 0] * s_x, I[1, 1] * c_x + J[1, 1] * s_x
        M22, M23 = I[2, 2] * c_y + J[2, 2] * s_y, I[2, 3] * c_y + J[2, 3] * s_y
        M32, M33 = I[3, 2] * c_y + J[3, 2] * s_y, I[3, 3] * c_y + J[3, 3] * s_y
        Dx0, Dx1, Dy0, Dy1 = self.D
        if self.dispersion:
            b.x = b.x - Dx0 * b.dp
            b.y = b.y - Dy0 * b.dp
        b.x, b.xp = M00 * b.x + M01 * b.xp, M10 * b.x + M11 * b.xp
        b.y, b.yp = M22 * b.y + M23 * b.yp, M32 * b.y + M33 * b.yp
        if self.dispersion:
            b.x = b.x + Dx1 * b.dp
            b.y = b.y + Dy1 * b.dp


def smooth_segment(Q_x, Q_y, beta_x, beta_y, Qp_x=0, Qp_y=0, app_x=0, app_y=0, app_xy=0):
    """One-segment smooth-optics ring (alpha = 0, D = 0) with the detuner
    order of Synchrotron (machines/synchrotron.py:474-479): chromaticity,
    then amplitude detuning; dmu = 1 for one segment."""
    dets = []
    if np.any(np.atleast_1d(Qp_x) != 0) or np.any(np.atleast_1d(Qp_y) != 0):
        dets.append(Chroma([q * 1.0 for q in np.atleast_1d(Qp_x)],
                           [q * 1.0 for q in np.atleast_1d(Qp_y)]))
    if app_x != 0 or app_y != 0 or app_xy != 0:
        dets.append(Amplitude(app_x * 1.0, app_y * 1.0, app_xy * 1.0, app_xy * 1.0,
                              0., beta_x, 0., beta_y))
    dQ_x = Q_x * (np.diff([0., 1.])[0] / 1.0)
    dQ_y = Q_y * (np.diff([0., 1.])[0] / 1.0)
    return TransverseSegment(0., beta_x, 0., 0., beta_x, 0., 0., beta_y, 0., 0., beta_y, 0.,
                             dQ_x, dQ_y, dets)


# ---------------------------------------------------------------------------
# longitudinal
# ---------------------------------------------------------------------------
def eta0(alpha_array, gamma):
    """LongitudinalMap._eta0 (longitudinal_tracking.py:79-81)."""
    return alpha_array[0] - gamma**-2


def drift(b, alpha_array, length, shrinkage_p_increment=0):
    """Drift.track (:107-111) with eta(dp) = 0 + eta0*dp**0."""
    beta_ratio = 1 - shrinkage_p_increment / (b.gamma**3 * b.p0)
    eta = 0 + eta0(alpha_array, b.gamma) * (b.dp**0)
    b.z = beta_ratio * b.z - eta * b.dp * length


def rf_kick(b, circumference, harmonic, voltage, phi_offset=0, phi_lock=0, p_increment=0,
            D_x=0, D_y=0):
    """Kick.track_with/without_dispersion (:185-203)."""
    if D_x or D_y:
        b.x = b.x - D_x * b.dp
        b.y = b.y - D_y * b.dp
    amplitude = np.abs(b.charge) * voltage / (b.beta * c)
    phi = harmonic * (2 * np.pi * b.z / circumference) + phi_offset + phi_lock
    delta_p = b.dp * b.p0
    delta_p += amplitude * np.sin(phi) - p_increment
    b.p0 += p_increment
    b.dp = delta_p / b.p0
    if D_x or D_y:
        b.x = b.x + D_x * b.dp
        b.y = b.y + D_y * b.dp


class RFSystemsOracle(object):
    """RFSystems (:280-665): drift(C/2), kicks, drift(C/2) (:362-369); the
    accelerating kick is the first; phase lock of the other kicks by
    -h_k/h_0 * phi_s with phi_s = arcsin(p_increment beta c / (|q| V_0))
    (:556-575, 530-542); with p_increment the transverse coordinates shrink
    by sqrt(betagamma_before / betagamma_after) (:516-530, 545-548); the
    last drift gets shrinkage_p_increment = p_increment only with
    shrink_longitudinal (:451-456)."""

    def __init__(self, circumference, harmonic_list, voltage_list, phi_offset_list,
                 alpha_array, gamma_reference, charge, p_increment=0, D_x=0, D_y=0,
                 phase_lock=True, shrink_transverse=True, shrink_longitudinal=False):
        self.C = circumference
        self.h, self.V, self.dphi = list(harmonic_list), list(voltage_list), list(phi_offset_list)
        self.alpha_array = alpha_array
        self.p_increment = p_increment
        self.D = (D_x, D_y)
        self.shrink_transverse = shrink_transverse
        self.shrink_longitudinal = shrink_longitudinal
        self.phi_lock = [0] * len(self.h)
        if phase_lock:
            phi_s = self.phi_s(gamma_reference, charge)
            for k in range(1, len(self.h)):
                self.phi_lock[k] -= self.h[k] / self.h[0] * phi_s

    def phi_s(self, gamma, charge):
        beta = np.sqrt(1 - gamma**-2)
        if self.p_increment == 0 and self.V[0] == 0:
            return 0
        deltaE = self.p_increment * beta * c
        return np.arcsin(deltaE / (np.abs(charge) * self.V[0]))

    def track(self, b):
        betagamma_old = b.betagamma
        drift(b, self.alpha_array, 0.5 * self.C)
        for k, (h, V, dphi, lock) in enumerate(zip(self.h, self.V, self.dphi, self.phi_lock)):
            rf_kick(b, self.C, h, V, dphi, lock, self.p_increment if k == 0 else 0, *self.D)
        drift(b, self.alpha_array, 0.5 * self.C,
              self.p_increment if self.shrink_longitudinal else 0)
        if self.p_increment and self.shrink_transverse:
            f = np.sqrt(betagamma_old / b.betagamma)
            b.x = b.x * f
            b.xp = b.xp * f
            b.y = b.y * f
            b.yp = b.yp * f


class LinearMapOracle(object):
    """LinearMap.track_with_dispersion (:731-759)."""

    def __init__(self, alpha_array, circumference, Q_s, D_x=0, D_y=0):
        self.alpha_array, self.C, self.Q_s, self.D_x, self.D_y = (
            alpha_array, circumference, Q_s, D_x, D_y)

    def track(self, b):
        b.x = b.x - self.D_x * b.dp
        b.y = b.y - self.D_y * b.dp
        omega_0 = 2 * np.pi * b.beta * c / self.C
        omega_s = self.Q_s * omega_0
        dQ_s = 2 * np.pi * self.Q_s
        cosdQ_s, sindQ_s = np.cos(dQ_s), np.sin(dQ_s)
        z0, dp0 = b.z, b.dp
        eta = 0 + eta0(self.alpha_array, b.gamma) * 1
        longfac = eta * b.beta * c / omega_s
        b.z = z0 * cosdQ_s - longfac * dp0 * sindQ_s
        b.dp = dp0 * cosdQ_s + z0 / longfac * sindQ_s
        b.x = b.x + self.D_x * b.dp
        b.y = b.y + self.D_y * b.dp


# ---------------------------------------------------------------------------
# wakes
# ---------------------------------------------------------------------------
def resonator_transverse(yokoya, R_shunt, frequency, Q):
    """Resonator.function_transverse (wakes.py:450-470)."""
    omega = 2 * np.pi * frequency
    alpha = omega / (2 * Q)
    omegabar = np.sqrt(np.abs(omega**2 - alpha**2))

    def wake(dt, *args, **kwargs):
        dt = dt.clip(max=0)
        if Q > 0.5:
            return (yokoya * R_shunt * omega**2 / (Q * omegabar) * np.exp(alpha * dt) *
                    np.sin(omegabar * dt))
        if Q == 0.5:
            return yokoya * R_shunt * omega**2 / Q * np.exp(alpha * dt) * dt
        return (yokoya * R_shunt * omega**2 / (Q * omegabar) * np.exp(alpha * dt) *
                np.sinh(omegabar * dt))
    return wake


def resonator_longitudinal(R_shunt, frequency, Q):
    """Resonator.function_longitudinal (wakes.py:472-493)."""
    omega = 2 * np.pi * frequency
    alpha = omega / (2 * Q)
    omegabar = np.sqrt(np.abs(omega**2 - alpha**2))

    def wake(dt, *args, **kwargs):
        t = -(np.sign(dt) - 1) * R_shunt * alpha * np.exp(alpha * dt)
        if Q > 0.5:
            return t * (np.cos(omegabar * dt) + alpha / omegabar * np.sin(omegabar * dt))
        if Q == 0.5:
            return t * (1. + alpha * dt)
        return t * (np.cosh(omegabar * dt) + alpha / omegabar * np.sinh(omegabar * dt))
    return wake


def resistive_wall_transverse(yokoya, pipe_radius, length, conductivity, dt_min):
    """ResistiveWall.function_transverse (wakes.py:611-624)."""
    mu_r = 1
    Z_0 = 119.9169832 * np.pi
    b = np.array([pipe_radius]).flatten()

    def wake(dt, *args, **kwargs):
        return (yokoya * (np.sign(dt + np.abs(dt_min)) - 1) / 2. * np.sqrt(kwargs['beta']) *
                length / np.pi / b**3 *
                np.sqrt(-mu_r / np.pi / conductivity / dt.clip(max=-abs(dt_min)))) * \
            np.sqrt(Z_0 * c)
    return wake


def wake_table_transverse(time_ns, strength):
    """WakeTable.function_transverse (wakes.py:290-339): ns -> s, V/pC/mm ->
    V/C/m with the sign flip, scipy's interp1d (the reference's own
    interpolator, bounds errors included) at -dt, dt clipped at 0 for an
    ultrarelativistic table."""
    from scipy.interpolate import interp1d
    time = 1e-9 * np.asarray(time_ns, dtype=np.float64)
    strength = -1e15 * np.asarray(strength, dtype=np.float64)
    f = interp1d(time, strength)
    if time[0] == 0 and strength[0] == 0:
        return lambda dt, *a, **k: f(-np.asarray(dt, dtype=np.float64).clip(max=0))
    if time[0] < 0:
        return lambda dt, *a, **k: f(-np.asarray(dt, dtype=np.float64))
    raise ValueError('does not meet requirements.')


def wake_table_longitudinal(time_ns, strength):
    """WakeTable.function_longitudinal (wakes.py:341-376): V/pC -> V/C, zero
    outside the table, half the value at dt = 0 for a table from 0."""
    from scipy.interpolate import interp1d
    time = 1e-9 * np.asarray(time_ns, dtype=np.float64)
    strength = 1e12 * np.asarray(strength, dtype=np.float64)
    f = interp1d(time, strength, bounds_error=False, fill_value=0.)

    def wake(dt, *a, **k):
        dt = np.asarray(dt, dtype=np.float64)
        w = f(-dt)
        if time[0] == 0:
            return (np.sign(-dt) + 1.) / 2. * w
        if time[0] < 0:
            return w
        raise ValueError('Longitudinal wake component does not meet requirements.')
    return wake


# ---------------------------------------------------------------------------
# 2.5-D transverse Gaussian space charge (spacecharge/spacecharge.py:159-417)
# ---------------------------------------------------------------------------
SC_RATIO_THRESHOLD = 1e-3      # :168-172
SC_ABSOLUTE_THRESHOLD = 1e-10  # :174-178


def sc_efieldn_mit(x, y, sig_x, sig_y):
    """_efieldn_mit (:275-300) with scipy.special.wofz (pmath.wofz)."""
    from scipy.special import wofz
    from scipy.constants import epsilon_0
    sig_sqrt = np.sqrt(2 * (sig_x**2 - sig_y**2))
    w1 = wofz(x / sig_sqrt + 1j * (y / sig_sqrt))
    ex = np.exp(-x * x / (2 * sig_x * sig_x) + -y * y / (2 * sig_y * sig_y))
    w2 = wofz(x * sig_y / (sig_x * sig_sqrt) + 1j * (y * sig_x / (sig_y * sig_sqrt)))
    denom = 2. * epsilon_0 * np.sqrt(np.pi) * sig_sqrt
    return (w1.imag - ex * w2.imag) / denom, (w1.real - ex * w2.real) / denom


def sc_efieldn_round(x, y, sig_r):
    """_efieldn_round (:392-399)."""
    from scipy.constants import epsilon_0
    r2 = x * x + y * y
    amplitude = (1 - np.exp(-r2 / (2 * sig_r * sig_r))) / (2 * np.pi * epsilon_0 * r2)
    return x * amplitude, y * amplitude


def sc_efieldn_checked(x, y, sig_x, sig_y):
    """add_sigma_check (:401-417)."""
    tol = dict(rtol=SC_RATIO_THRESHOLD, atol=SC_ABSOLUTE_THRESHOLD)
    if np.allclose(sig_y, sig_x, **tol):
        if np.allclose(sig_y, 0, **tol):
            return np.zeros(x.shape), np.zeros(x.shape)
        return sc_efieldn_round(x, y, sig_x)
    if np.all(sig_x < sig_y):
        en_y, en_x = sc_efieldn_mit(y, x, sig_y, sig_x)
        return en_x, en_y
    return sc_efieldn_mit(x, y, sig_x, sig_y)


def transverse_gaussian_space_charge(b, n_slices, length, n_sigma_z=None, z_cuts=None,
                                     sig_check=True):
    """TransverseGaussianSpaceCharge.track (:202-224) with get_efieldn
    (:226-246): the kick of every particle inside the cuts, slice by slice."""
    from scipy.constants import c
    sd = slice_uniform(b, n_slices, n_sigma_z, z_cuts)
    mx, my = mean_per_slice(sd, b.x), mean_per_slice(sd, b.y)
    sx, sy = std_per_slice(sd, b.x), std_per_slice(sd, b.y)
    prefactor = (b.charge * length / (b.p0 * b.betagamma * b.gamma * c))
    charge_per_mp = b.particlenumber_per_mp * b.charge
    lam = sd.n_macroparticles_per_slice * charge_per_mp / np.diff(sd.z_bins)
    idx = sd.slice_index_of_particle
    efield = sc_efieldn_checked if sig_check else sc_efieldn_mit
    for s in range(n_slices):
        p_id = np.where(idx == s)[0]
        if len(p_id) == 0:
            continue
        x = b.x[p_id] - mx[s]
        y = b.y[p_id] - my[s]
        en_x, en_y = efield(np.abs(x), np.abs(y), sx[s], sy[s])
        en_x = np.abs(en_x) * np.sign(x)
        en_y = np.abs(en_y) * np.sign(y)
        b.xp[p_id] = b.xp[p_id] + (en_x * lam[s]) * prefactor
        b.yp[p_id] = b.yp[p_id] + (en_y * lam[s]) * prefactor


# kick kinds: (moment, kicked coordinate, multiplying coordinate or None)
KICK_KINDS = {
    'ConstantX': ('n', 'xp', None), 'ConstantY': ('n', 'yp', None),
    'ConstantZ': ('n', 'dp', None),
    'DipoleX': ('x', 'xp', None), 'DipoleXY': ('y', 'xp', None),
    'DipoleY': ('y', 'yp', None), 'DipoleYX': ('x', 'yp', None),
    'QuadX': ('n', 'xp', 'x'), 'QuadXY': ('n', 'xp', 'y'),
    'QuadY': ('n', 'yp', 'y'), 'QuadYX': ('n', 'yp', 'x'),
}


class Kick(object):
    """One WakeKick (wake_kicks.py:33-139 and the apply methods)."""

    def __init__(self, kind, wake_function, n_turns_wake, uniform_bin=True, z_cuts=None):
        self.kind = kind
        self.moment, self.target, self.factor = KICK_KINDS[kind]
        self.wake = wake_function
        self.n_turns_wake = n_turns_wake
        self.convolve = uniform_bin and (n_turns_wake == 1 or bool(z_cuts))

    @staticmethod
    def wake_factor(b):
        return (-(b.charge)**2 / (b.mass * b.gamma * (b.beta * c)**2) *
                b.particlenumber_per_mp)

    def _conv(self, t_tgt, s_src, moments, beta):
        if self.convolve:
            dt = np.concatenate((t_tgt - s_src[-1], (t_tgt - s_src[0])[1:]))
            return np.convolve(moments, self.wake(dt, beta=beta), 'valid')
        dt = [t_tgt] - np.transpose([s_src])
        return np.dot(moments, self.wake(dt, beta=beta))

    def signal(self, b, slices, ages):
        times = [s.convert_to_time(s.z_centers) for s in slices]
        acc = 0
        for i in range(min(len(ages), self.n_turns_wake)):
            s = slices[i]
            if self.moment == 'n':
                m = s.n_macroparticles_per_slice
            elif self.moment == 'x':
                m = s.n_macroparticles_per_slice * s.mean_x
            else:
                m = s.n_macroparticles_per_slice * s.mean_y
            acc += self._conv(times[0], times[i] + ages[i], m, s.beta)
        return self.wake_factor(b) * acc

    def apply(self, b, slices, ages):
        k = self.signal(b, slices, ages)
        s0 = slices[0]
        p = s0.particles_within_cuts
        kick = np.take(k, s0.slice_index_of_particle[p])
        target = getattr(b, self.target)
        if self.factor is None:
            target[p] += kick
        else:
            target[p] += kick * getattr(b, self.factor)[p]


class WakeFieldOracle(object):
    """WakeField.track (wakes.py:121-145)."""

    def __init__(self, n_slices, kicks, n_sigma_z=None, z_cuts=None, uniform_charge=False):
        self.n_slices, self.n_sigma_z, self.z_cuts = n_slices, n_sigma_z, z_cuts
        self.uniform_charge = uniform_charge
        self.kicks = kicks
        T = max(k.n_turns_wake for k in kicks)
        self.slices = deque([], maxlen=T)
        self.ages = deque([], maxlen=T)

    def track(self, b):
        for i in range(len(self.ages)):
            self.ages[i] += b.circumference / (b.beta * c)
        if self.uniform_charge:
            sd = slice_uniform_charge(b, self.n_slices, self.n_sigma_z, self.z_cuts)
        else:
            sd = slice_uniform(b, self.n_slices, self.n_sigma_z, self.z_cuts)
        sd.mean_x = mean_per_slice(sd, b.x)
        sd.mean_y = mean_per_slice(sd, b.y)
        self.slices.appendleft(sd)
        self.ages.appendleft(0.)
        for k in self.kicks:
            k.apply(b, self.slices, self.ages)
        return sd


def slice_uniform_charge(beam, n_slices, n_sigma_z=None, z_cuts=None, perm=None):
    """UniformChargeSlicer.compute_sliceset_kwargs (slicing.py:636-686),
    with random.sample drawing from Python's global generator.  perm: the
    sort permutation to use (np.argsort(z) when None)."""
    from random import sample
    z_cut_tail, z_cut_head = long_cuts(beam, n_sigma_z, z_cuts)
    n_part = len(beam.z)
    id_new = np.argsort(beam.z) if perm is None else np.asarray(perm)
    id_old = np.arange(n_part).take(id_new)
    z_sorted = beam.z.take(id_new)
    n_cut_tail = np.searchsorted(z_sorted, z_cut_tail)
    n_cut_head = n_part - np.searchsorted(z_sorted, z_cut_head)
    n_part_within_cuts = n_part - n_cut_tail - n_cut_head
    rand_slice_i = sample(list(range(n_slices)), n_part_within_cuts % n_slices)
    n_part_per_slice = (n_part_within_cuts // n_slices) * np.ones(n_slices, dtype=np.int32)
    n_part_per_slice[rand_slice_i] += 1
    n_all = np.hstack((n_cut_tail, n_part_per_slice, n_cut_head))
    first_indices = np.cumsum(n_all, dtype=np.int32)[:-1]
    z_bins = np.empty(n_slices + 1)
    z_bins[1:-1] = (z_sorted[first_indices[1:-1] - 1] + z_sorted[first_indices[1:-1]]) / 2
    z_bins[0], z_bins[-1] = z_cut_tail, z_cut_head
    idx_sorted = -np.ones(n_part, dtype=np.int32)
    for i in range(n_slices):
        idx_sorted[first_indices[i]:first_indices[i + 1]] = i
    idx = np.empty(n_part, dtype=np.int32)
    np.put(idx, id_old, idx_sorted)
    sd = SliceData(z_bins, idx, beam)
    sd.n_macroparticles_per_slice = n_part_per_slice
    return sd
