"""The oracle's space-charge field (a restatement of spacecharge.py:226-417
with scipy.special.wofz, the reference's own Faddeeva function) pinned by
what electrostatics fixes: the far field of any Gaussian line charge is that
of a line charge, lambda / (2 pi eps0 r); the elliptic field tends to the
round one as sigma_y -> sigma_x; exchanging the planes exchanges the
components; the round field is the closed form (1 - exp(-r^2/2s^2)) /
(2 pi eps0 r^2) (x, y) and is linear near the axis with slope 1/(4 pi eps0
s^2)."""
import numpy as np
from scipy.constants import epsilon_0

from oracle import refpath as R


def test_far_field_is_a_line_charge():
    rng = np.random.default_rng(1)
    r = 400e-3
    phi = rng.uniform(0, np.pi / 2, 50)
    x, y = r * np.cos(phi), r * np.sin(phi)
    for sx, sy in ((2e-3, 1e-3), (1e-3, 2.5e-3), (1.5e-3, 1.5e-3)):
        ex, ey = R.sc_efieldn_checked(x, y, sx, sy)
        want = 1 / (2 * np.pi * epsilon_0 * r)
        np.testing.assert_allclose(np.hypot(ex, ey), want, rtol=1e-4)
        np.testing.assert_allclose(np.arctan2(ey, ex), phi, atol=1e-4)


def test_elliptic_tends_to_round_and_planes_exchange():
    rng = np.random.default_rng(2)
    x, y = rng.uniform(0, 5e-3, 200), rng.uniform(0, 5e-3, 200)
    s = 1.7e-3
    ex_r, ey_r = R.sc_efieldn_round(x, y, s)
    ex_e, ey_e = R.sc_efieldn_mit(x, y, s * (1 + 2e-3), s * (1 - 2e-3))
    np.testing.assert_allclose(ex_e, ex_r, rtol=1e-2)
    np.testing.assert_allclose(ey_e, ey_r, rtol=1e-2)
    ex1, ey1 = R.sc_efieldn_checked(x, y, 2e-3, 1e-3)
    ex2, ey2 = R.sc_efieldn_checked(y, x, 1e-3, 2e-3)
    np.testing.assert_allclose(ex1, ey2, rtol=1e-13)
    np.testing.assert_allclose(ey1, ex2, rtol=1e-13)


def test_round_field_closed_form_and_axis_slope():
    s = 1e-3
    x = np.array([1e-5, 1e-3, 3e-3])
    ex, ey = R.sc_efieldn_round(x, 0 * x, s)
    want = (1 - np.exp(-x**2 / (2 * s**2))) / (2 * np.pi * epsilon_0 * x)
    np.testing.assert_allclose(ex, want, rtol=1e-12)
    assert np.all(ey == 0)
    # (1 - e^-a)/a = 1 - a/2 + ..., a = x^2/2s^2 = 5e-5
    np.testing.assert_allclose(ex[0] / x[0], 1 / (4 * np.pi * epsilon_0 * s**2), rtol=1e-4)
    # zero sigma in both planes: no field
    ex0, ey0 = R.sc_efieldn_checked(x, x, 0., 0.)
    assert np.all(ex0 == 0) and np.all(ey0 == 0)
