"""Particles: the macroparticle ensemble, with its six FP64 coordinate
columns resident in device memory.

Same contract as PyHEADTAIL particles/particles.py:16-390:
  - constructor (macroparticlenumber, particlenumber_per_mp, charge, mass,
    circumference, gamma, coords_n_momenta_dict); update() copies;
  - gamma / beta / betagamma / p0 setters re-derive each other with the same
    arithmetic (:74-103), so host scalars stay bit-identical to the
    reference's;
  - get_slices(slicer, statistics=...) caches one SliceSet per slicer
    configuration and re-slices whenever statistics are requested (:118-147);
  - mean_* / sigma_* / epsn_* statistics (:311-366), dispersion_* and the
    statistical Twiss parameters alpha/beta/gamma_Twiss_* (:368-390).

Coordinates are torch float64 tensors on the current HIP device; numpy
arrays given to the constructor or update() are copied there.  Elements of
the fused engine may leave device work pending on a bunch (the wake kick of
its last turn); reading any coordinate applies it first, so every observer
sees the same state as after the reference's element-by-element turn.
"""
import numpy as np
import torch
from scipy.constants import c, e

from .. import _phx
from ..general.element import Printing

COORDS = ("x", "xp", "y", "yp", "z", "dp")


def _kinematics_of_gamma(gamma, mass):
    """(beta, betagamma, p0) derived from gamma as the reference's gamma
    setter does (particles/particles.py:74-103)."""
    betagamma = np.sqrt(gamma**2 - 1)
    return np.sqrt(1 - gamma**-2), betagamma, betagamma * mass * c


class Kinematics(object):
    """gamma, beta, betagamma and p0, each settable; setting one re-derives
    the others through gamma with the reference's expressions, so host
    scalars stay bit-identical to the reference's.  Needs self.mass."""

    def _set_gamma(self, gamma):
        self._gamma = gamma
        self._beta, self._betagamma, self._p0 = _kinematics_of_gamma(gamma, self.mass)

    gamma = property(lambda self: self._gamma, _set_gamma)
    beta = property(lambda self: self._beta,
                    lambda self, v: self._set_gamma(1. / np.sqrt(1 - v**2)))
    betagamma = property(lambda self: self._betagamma,
                         lambda self, v: self._set_gamma(np.sqrt(v**2 + 1)))
    p0 = property(lambda self: self._p0,
                  lambda self, v: self._set_gamma(np.sqrt(1 + (v / (self.mass * c))**2)))


class ScalarState(Kinematics):
    """The host scalars of a bunch without coordinates: what the fused
    engine advances turn by turn to precompute per-turn parameters."""

    def __init__(self, beam):
        self.mass = beam.mass
        self.charge = beam.charge
        self.circumference = beam.circumference
        self.macroparticlenumber = beam.macroparticlenumber
        self.particlenumber_per_mp = beam.particlenumber_per_mp
        self._gamma, self._beta = beam._gamma, beam._beta
        self._betagamma, self._p0 = beam._betagamma, beam._p0

    @property
    def intensity(self):
        return self.particlenumber_per_mp * self.macroparticlenumber

    @property
    def charge_per_mp(self):
        return self.particlenumber_per_mp * self.charge

    def write_back(self, beam):
        beam._gamma, beam._beta = self._gamma, self._beta
        beam._betagamma, beam._p0 = self._betagamma, self._p0


def as_device_f64(value, n=None):
    """A contiguous float64 tensor on the current device (copy if needed)."""
    if isinstance(value, torch.Tensor):
        t = value
        if t.device.type != "cuda" or t.dtype != torch.float64 or not t.is_contiguous():
            t = t.to(device=_phx.device(), dtype=torch.float64).contiguous()
    else:
        t = torch.as_tensor(np.ascontiguousarray(value, dtype=np.float64)).to(_phx.device())
    if n is not None and t.numel() != n:
        raise ValueError("coordinate array has %d entries, expected %d" % (t.numel(), n))
    return t


def _coordinate(name):
    def getter(self):
        if self._pending is not None:
            self._run_pending()
        try:
            return self._coords[name]
        except KeyError:
            # a bunch without this plane (e.g. transverse only), as a missing
            # attribute of the reference's Particles
            raise AttributeError("this Particles object has no coordinate %r" % name)

    def setter(self, value):
        if self._pending is not None:
            self._run_pending()
        self._coords[name] = as_device_f64(value, self.macroparticlenumber)

    return property(getter, setter, doc="device float64 column %r" % name)


class Particles(Kinematics, Printing):
    """Macroparticle ensemble with device-resident coordinates."""

    x = _coordinate("x")
    xp = _coordinate("xp")
    y = _coordinate("y")
    yp = _coordinate("yp")
    z = _coordinate("z")
    dp = _coordinate("dp")

    def __init__(self, macroparticlenumber, particlenumber_per_mp, charge, mass,
                 circumference, gamma, coords_n_momenta_dict={}, *args, **kwargs):
        self._coords = {}
        self._pending = None
        self.macroparticlenumber = macroparticlenumber
        self.particlenumber_per_mp = particlenumber_per_mp
        self.charge = charge
        self.mass = mass
        self.circumference = circumference
        self.gamma = gamma
        self._slice_sets = {}
        self.coords_n_momenta = set()
        self.id = torch.arange(1, macroparticlenumber + 1, dtype=torch.int32,
                               device=_phx.device())
        self.update(coords_n_momenta_dict)

    # -- deferred device work ------------------------------------------------
    def _set_pending(self, owner):
        """Register an object whose flush(beam) must run before anyone reads
        the coordinates (the fused engine's pending wake kick)."""
        if self._pending is not None and self._pending is not owner:
            self._run_pending()
        self._pending = owner

    def _run_pending(self):
        owner, self._pending = self._pending, None
        owner.flush(self)

    def _raw(self, name):
        """Coordinate tensor without running pending work (engine use)."""
        return self._coords[name]

    # -- reference API -------------------------------------------------------
    @property
    def intensity(self):
        return self.particlenumber_per_mp * self.macroparticlenumber

    @intensity.setter
    def intensity(self, value):
        self.particlenumber_per_mp = value / float(self.macroparticlenumber)

    @property
    def charge_per_mp(self):
        return self.particlenumber_per_mp * self.charge

    @charge_per_mp.setter
    def charge_per_mp(self, value):
        self.particlenumber_per_mp = value / self.charge

    @property
    def z_beamframe(self):
        return self.z * self.gamma

    @z_beamframe.setter
    def z_beamframe(self, value):
        self.z = value / self.gamma

    def get_coords_n_momenta_dict(self):
        return {coord: getattr(self, coord) for coord in self.coords_n_momenta}

    def get_slices(self, slicer, *args, **kwargs):
        if slicer not in self._slice_sets or kwargs.get('statistics'):
            self._slice_sets[slicer] = slicer.slice(self, *args, **kwargs)
        return self._slice_sets[slicer]

    def clean_slices(self):
        self._slice_sets = {}

    def update(self, coords_n_momenta_dict):
        if any(len(v) != self.macroparticlenumber for v in coords_n_momenta_dict.values()):
            raise ValueError("lengths of given phase space coordinate arrays"
                             " do not coincide with self.macroparticlenumber.")
        for coord, array in list(coords_n_momenta_dict.items()):
            if isinstance(array, torch.Tensor):
                t = as_device_f64(array).clone()
            else:
                t = as_device_f64(np.array(array, dtype=np.float64, copy=True))
            if coord in COORDS:
                setattr(self, coord, t)
            else:
                setattr(self, coord, t)
        self.coords_n_momenta.update(coords_n_momenta_dict.keys())

    def add(self, coords_n_momenta_dict):
        if any(s in self.coords_n_momenta for s in coords_n_momenta_dict):
            raise ValueError("One or more of the specified coordinates or"
                             " momenta already exist and cannot be added."
                             " Use self.update(...) for this purpose.")
        self.update(coords_n_momenta_dict)

    def sort_for(self, attr):
        permutation = torch.argsort(getattr(self, attr), stable=True)
        self.reorder(permutation)

    def reorder(self, permutation, except_for_attrs=[]):
        for attr in ['id'] + list(self.coords_n_momenta):
            if attr in except_for_attrs:
                continue
            setattr(self, attr, getattr(self, attr)[permutation])
        self.clean_slices()

    def to_numpy(self):
        """Host copy {coord: np.ndarray} of all coordinates (synchronises)."""
        return {k: getattr(self, k).cpu().numpy() for k in self.coords_n_momenta}

    def __add__(self, other):
        result = Particles(self.macroparticlenumber + other.macroparticlenumber,
                           self.particlenumber_per_mp, self.charge, self.mass,
                           self.circumference, self.gamma, {})
        for coord in self.coords_n_momenta:
            result.update({coord: torch.cat((getattr(self, coord), getattr(other, coord)))})
        result.id = torch.cat((self.id, other.id))
        return result

    def __radd__(self, other):
        if other == 0:
            result = Particles(self.macroparticlenumber, self.particlenumber_per_mp,
                               self.charge, self.mass, self.circumference, self.gamma, {})
            for coord in self.coords_n_momenta:
                result.update({coord: getattr(self, coord)})
            result.id = self.id.clone()
            return result
        return self.__add__(other)

    # -- statistics ----------------------------------------------------------
    def _sums(self, arrays):
        """Device reduction of up to six coordinate arrays (see
        phx_bunch_sums); returns the sums as a numpy array (synchronises)."""
        L = _phx.lib()
        k = len(arrays)
        n = self.macroparticlenumber
        out = torch.empty(L.phx_bunch_sums_len(k), dtype=torch.float64, device=_phx.device())
        ptrs = (_phx.P * k)(*[a.data_ptr() for a in arrays])
        ws = _phx.workspace(n)
        _phx.call(L.phx_bunch_sums, ptrs, k, n, out.data_ptr(), ws.data_ptr(), _phx.stream())
        return out.cpu().numpy()

    def _mean(self, u):
        s = self._sums([u])
        return np.float64(s[0] / self.macroparticlenumber)

    def _cov_from_sums(self, s, k, i, j):
        """cobra_functions/stats.pyx:cov with the shifted single-pass sums."""
        n = self.macroparticlenumber
        if n < 2:
            return 0.
        a_sum, b_sum = s[k + i], s[k + j]
        lo, hi = min(i, j), max(i, j)
        p = lo * k - lo * (lo - 1) // 2 + (hi - lo)
        ab_sum = s[2 * k + p]
        return (ab_sum - a_sum * b_sum / n) / (n - 1)

    def _std(self, u):
        s = self._sums([u])
        return np.sqrt(self._cov_from_sums(s, 1, 0, 0))

    def _emittance(self, u, up, dp):
        """cobra_functions/stats.pyx:emittance (:84-122)."""
        arrays = [u, up] + ([dp] if dp is not None else [])
        k = len(arrays)
        s = self._sums(arrays)
        cov_u2 = self._cov_from_sums(s, k, 0, 0)
        cov_up2 = self._cov_from_sums(s, k, 1, 1)
        cov_u_up = self._cov_from_sums(s, k, 1, 0)
        term_u2_dp = term_u_up_dp = term_up2_dp = 0.
        if dp is not None:
            cov_dp2 = self._cov_from_sums(s, k, 2, 2)
            if cov_dp2 != 0:
                cov_u_dp = self._cov_from_sums(s, k, 0, 2)
                cov_up_dp = self._cov_from_sums(s, k, 1, 2)
                term_u2_dp = cov_u_dp * cov_u_dp / cov_dp2
                term_u_up_dp = cov_u_dp * cov_up_dp / cov_dp2
                term_up2_dp = cov_up_dp * cov_up_dp / cov_dp2
        sigma11 = cov_u2 - term_u2_dp
        sigma12 = cov_u_up - term_u_up_dp
        sigma22 = cov_up2 - term_up2_dp
        return np.sqrt(sigma11 * sigma22 - sigma12 * sigma12)

    def effective_normalized_emittance_x(self, **kwargs):
        return self._emittance(self.x, self.xp, None) * self.betagamma

    def effective_normalized_emittance_y(self, **kwargs):
        return self._emittance(self.y, self.yp, None) * self.betagamma

    def effective_normalized_emittance_z(self, **kwargs):
        return 4 * np.pi * self._emittance(self.z, self.dp, None) * self.p0 / e

    def epsn_x(self, **kwargs):
        return self._emittance(self.x, self.xp, getattr(self, 'dp', None)) * self.betagamma

    def epsn_y(self, **kwargs):
        return self._emittance(self.y, self.yp, getattr(self, 'dp', None)) * self.betagamma

    def epsn_z(self, **kwargs):
        return self.effective_normalized_emittance_z(**kwargs)

    # -- dispersion and statistical Twiss parameters (stats.pyx:57-179) ------
    def _dispersion(self, u):
        """<u dp> / <dp^2>, or 0 when <dp^2> is 0."""
        dp = self.dp
        mean_u_dp, mean_dp2 = self._mean(u * dp), self._mean(dp * dp)
        return mean_u_dp / mean_dp2 if mean_dp2 > 0 else 0

    def dispersion_x(self):
        return self._dispersion(self.x)

    def dispersion_y(self):
        return self._dispersion(self.y)

    def _twiss(self, u, up, dp, entry):
        """Entry (0, 0) beta, (0, 1) alpha, (1, 1) gamma of the beam matrix
        with the dispersion terms removed, over the emittance.  Without dp the
        dispersion terms are 0 (cov_dp2 stands at 1); with dp, cov_dp2 is
        used as it is (stats.pyx:get_alpha / get_beta / get_gamma)."""
        arrays = [u, up] + ([dp] if dp is not None else [])
        k = len(arrays)
        s = self._sums(arrays)
        cov = lambda i, j: self._cov_from_sums(s, k, i, j)   # noqa: E731
        if dp is None:
            cov_dp2, disp = 1., (0., 0.)
        else:
            cov_dp2, disp = cov(2, 2), (cov(0, 2), cov(1, 2))
        i, j = entry
        sigma = cov(i, j) - disp[i] * disp[j] / cov_dp2
        return (-sigma if entry == (0, 1) else sigma) / self._emittance(u, up, dp)

    def alpha_Twiss_x(self):
        return self._twiss(self.x, self.xp, getattr(self, 'dp', None), (0, 1))

    def alpha_Twiss_y(self):
        return self._twiss(self.y, self.yp, getattr(self, 'dp', None), (0, 1))

    def beta_Twiss_x(self):
        return self._twiss(self.x, self.xp, getattr(self, 'dp', None), (0, 0))

    def beta_Twiss_y(self):
        return self._twiss(self.y, self.yp, getattr(self, 'dp', None), (0, 0))

    def gamma_Twiss_x(self):
        return self._twiss(self.x, self.xp, getattr(self, 'dp', None), (1, 1))

    def gamma_Twiss_y(self):
        return self._twiss(self.y, self.yp, getattr(self, 'dp', None), (1, 1))


def _add_moment_methods(cls):
    """mean_<u>() and sigma_<u>() for the six coordinates."""
    for name in COORDS:
        def mean(self, _name=name, **kwargs):
            return self._mean(getattr(self, _name))

        def sigma(self, _name=name, **kwargs):
            return self._std(getattr(self, _name))
        mean.__name__, sigma.__name__ = 'mean_' + name, 'sigma_' + name
        setattr(cls, mean.__name__, mean)
        setattr(cls, sigma.__name__, sigma)


_add_moment_methods(Particles)
