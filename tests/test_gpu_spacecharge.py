"""2.5-D transverse Gaussian space charge on the device against the oracle
(spacecharge.py:159-417 restated with scipy.special.wofz), and the device
Faddeeva function against scipy.special.wofz.

Cases: flat beam (sigma_x > sigma_y), tall beam (the planes exchanged), a
round beam (identical x and y: the round-beam formula), sig_check off, one
slice over the whole bunch, slices with a single particle (sigma 0: no
kick).  The kicks are compared, relative to their largest value."""
import ctypes

import numpy as np
import pytest
import torch

from oracle import refpath as R
from oracle import workloads as W
from tests.machines import COORDS, to_numpy, to_product

pytestmark = pytest.mark.gpu

KICK_TOL = 1e-11


def test_device_wofz_accuracy():
    print("\n[case] device Faddeeva function vs scipy.special.wofz", flush=True)
    from scipy.special import wofz
    from pyheadtail_amd import _phx
    rng = np.random.default_rng(0)
    xs = np.concatenate([rng.uniform(0, 6, 20000), rng.uniform(0, 60, 20000),
                         10**rng.uniform(-6, 3, 20000), [0., 0., 1e-300, 5.]])
    ys = np.concatenate([rng.uniform(0, 6, 20000), rng.uniform(0, 60, 20000),
                         10**rng.uniform(-8, 3, 20000), [0., 1e-3, 0., 0.]])
    dev = _phx.device()
    x = torch.from_numpy(xs).to(dev)
    y = torch.from_numpy(ys).to(dev)
    wr, wi = torch.empty_like(x), torch.empty_like(x)
    _phx.call(_phx.lib().phx_math_wofz, x.data_ptr(), y.data_ptr(), len(xs), wr.data_ptr(),
              wi.data_ptr(), _phx.stream())
    got = wr.cpu().numpy() + 1j * wi.cpu().numpy()
    ref = wofz(xs + 1j * ys)
    err = np.abs(got - ref) / np.abs(ref)
    print("[err] max |w - wofz| / |wofz| = %.3e" % err.max(), flush=True)
    assert err.max() <= 5e-14


def _bunch(kind, n=40000, seed=3):
    ob = W.linear_bunch(n, seed=seed)
    if kind == 'flat':
        ob.y *= 0.6
        ob.yp *= 0.6
    elif kind == 'tall':
        ob.x *= 0.5
        ob.xp *= 0.5
    elif kind == 'round':
        ob.y = ob.x.copy()
        ob.yp = ob.xp.copy()
    return ob


@pytest.mark.parametrize("kind,n_slices,sig_check", [
    ('flat', 64, True), ('tall', 64, True), ('round', 64, True), ('flat', 64, False),
    ('flat', 1, True), ('tall', 400, True)])
def test_spacecharge_kick_vs_oracle(kind, n_slices, sig_check):
    print("\n[case] space charge %s beam, %d slices, sig_check=%s" % (kind, n_slices, sig_check),
          flush=True)
    from pyheadtail_amd.particles.slicing import UniformBinSlicer
    from pyheadtail_amd.spacecharge.spacecharge import TransverseGaussianSpaceCharge
    ob = _bunch(kind)
    before = {k: getattr(ob, k).copy() for k in COORDS}
    pb = to_product(ob)
    sc = TransverseGaussianSpaceCharge(UniformBinSlicer(n_slices, n_sigma_z=3),
                                       W.CIRCUMFERENCE / 8, sig_check=sig_check)
    sc.track(pb)
    R.transverse_gaussian_space_charge(ob, n_slices, W.CIRCUMFERENCE / 8, n_sigma_z=3,
                                       sig_check=sig_check)
    got = to_numpy(pb)
    for k in ('xp', 'yp'):
        dk_ref = getattr(ob, k) - before[k]
        dk = got[k] - before[k]
        scale = np.max(np.abs(dk_ref))
        err = np.max(np.abs(dk - dk_ref)) / scale
        print("[err] %s kick %.3e (max kick %.3e)" % (k, err, scale), flush=True)
        assert scale > 0 and err <= KICK_TOL, (k, err)
    for k in ('x', 'y', 'z', 'dp'):
        np.testing.assert_array_equal(got[k], before[k])
