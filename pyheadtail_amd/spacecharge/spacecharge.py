"""2.5-D transverse Gaussian space charge (PyHEADTAIL spacecharge/
spacecharge.py:159-417): the Bassetti-Erskine field of each slice's
Gaussian, applied slice by slice to the particles inside the cuts, in one
device kernel per kick (phx_spacecharge_kick).  The Faddeeva function is
evaluated on the device (Weideman's rational approximation, N = 40) in
place of scipy.special.wofz.
"""
import ctypes

import numpy as np
from scipy.constants import c, epsilon_0

from .. import _phx
from ..general.element import Element


class TransverseGaussianSpaceCharge(Element):
    """Transverse space charge of a Gaussian distribution, slice-wise."""

    # abs(1 - sig_y / sig_x) < ratio_threshold ==> round beam
    ratio_threshold = 1e-3
    # abs(sig_y - sig_x) < absolute_threshold ==> round beam
    absolute_threshold = 1e-10

    def __init__(self, slicer, length, sig_check=True, other_efieldn=None):
        """slicer: the slices over which the field is applied (n_slices = 1:
        one longitudinally averaged kick); length: the s interval over which
        the force is integrated; sig_check: exchange x and y for
        sigma_x < sigma_y and use the round-beam field for sigma_x ~ sigma_y.
        other_efieldn selects one of the reference's alternative host
        implementations of the same field; only the device field is
        provided here."""
        if other_efieldn is not None:
            raise NotImplementedError("TransverseGaussianSpaceCharge: only the device "
                                      "Bassetti-Erskine field is provided")
        self.slicer = slicer
        self.length = length
        self.sig_check = bool(sig_check)

    def _params(self, beam):
        p = _phx.SpaceCharge()
        p.prefactor = (beam.charge * self.length /
                       (beam.p0 * beam.betagamma * beam.gamma * c))
        p.charge_per_mp = beam.particlenumber_per_mp * beam.charge
        p.denom_pre = 2. * epsilon_0 * np.sqrt(np.pi)
        p.round_pre = 2 * np.pi * epsilon_0
        p.rtol = type(self).ratio_threshold
        p.atol = type(self).absolute_threshold
        p.sig_check = 1 if self.sig_check else 0
        return p

    def track(self, beam):
        slices = beam.get_slices(self.slicer,
                                 statistics=["mean_x", "mean_y", "sigma_x", "sigma_y"])
        sl = _phx.SCSlices()
        sl.n_slices = slices.n_slices
        sl.counts = slices.n_macroparticles_per_slice.data_ptr()
        sl.z_bins = slices.z_bins.data_ptr()
        sl.mean_x, sl.mean_y = slices.mean_x.data_ptr(), slices.mean_y.data_ptr()
        sl.sigma_x, sl.sigma_y = slices.sigma_x.data_ptr(), slices.sigma_y.data_ptr()
        p = self._params(beam)
        coords = _phx.ptr6([beam.x, beam.xp, beam.y, beam.yp, beam.z, beam.dp])
        _phx.call(_phx.lib().phx_spacecharge_kick, coords, beam.macroparticlenumber,
                  slices.slice_index_of_particle.data_ptr(), slices._state.data_ptr(),
                  ctypes.byref(sl), ctypes.byref(p), _phx.stream())
