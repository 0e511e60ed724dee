"""The oracle's space-charge field (a restatement of spacecharge.py:226-417
with scipy.special.wofz, the reference's own Faddeeva function) pinned by
what electrostatics fixes: the far field of any Gaussian line charge is that
of a line charge, lambda / (2 pi eps0 r); the elliptic field tends to the
round one as sigma_y -> sigma_x; exchanging the planes exchanges the
components; the round field is the closed form (1 - exp(-r^2/2s^2)) /
(2 pi eps0 r^2) (x, y) and is linear near the axis with slope 1/(4 pi eps0
s^2); the assembled kick (prefactor q L / (p0 beta gamma gamma c), line
density N q_mp / width) is the textbook small-amplitude space-charge kick."""
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


def test_kick_assembly_small_amplitude_closed_form():
    """The whole kick of transverse_gaussian_space_charge (prefactor, line
    density, field) against the textbook small-amplitude kick of a Gaussian
    beam: inside an elliptic Gaussian slice of line charge density lambda,
    E_x = lambda x / (2 pi eps0 sig_x (sig_x + sig_y)) (E_y likewise) near
    the axis, and a particle of charge q and momentum p0 = gamma m beta c
    crossing the length L receives dx' = q E_x L (1 - beta^2) / (p0 beta c)
    (the beam's magnetic field cancels the fraction beta^2 of the electric
    force).  lambda = N q_mp / slice width, here one slice of width
    6 sigma_z.  Probe particles at 1e-3 sigma from the centre: the cubic
    correction is 1e-6 of the kick."""
    from scipy.constants import c, e, m_p
    rng = np.random.default_rng(7)
    n, sx_b, sy_b = 400000, 2.0e-3, 1.3e-3
    gamma, ppm, L = 27.7, 1e11 / 400000, 1727.875
    x = rng.normal(0, sx_b, n)
    y = rng.normal(0, sy_b, n)
    z = rng.uniform(-0.2, 0.2, n)
    # probes near the centre (they belong to the slice like any particle)
    k = 8
    x[:k] = np.mean(x[k:]) + 1e-3 * sx_b * np.linspace(-1, 1, k)
    y[:k] = np.mean(y[k:]) + 1e-3 * sy_b * np.linspace(1, -1, k)
    coords = dict(x=x, xp=np.zeros(n), y=y, yp=np.zeros(n), z=z, dp=np.zeros(n))
    b = R.Bunch(n, ppm, e, m_p, 6911.5, gamma, coords)
    R.transverse_gaussian_space_charge(b, 1, L, n_sigma_z=3)
    # the slice's statistics and line density from first principles
    mx, my = np.mean(x), np.mean(y)
    sx, sy = np.std(x, ddof=1), np.std(y, ddof=1)
    width = 6 * np.std(z, ddof=1)            # one slice over mean_z +- 3 sigma_z
    lam = n * ppm * e / width
    beta = np.sqrt(1 - gamma**-2)
    p0 = gamma * m_p * beta * c
    ex = lam * (x[:k] - mx) / (2 * np.pi * epsilon_0 * sx * (sx + sy))
    ey = lam * (y[:k] - my) / (2 * np.pi * epsilon_0 * sy * (sx + sy))
    want_xp = e * ex * L * (1 - beta**2) / (p0 * beta * c)
    want_yp = e * ey * L * (1 - beta**2) / (p0 * beta * c)
    np.testing.assert_allclose(b.xp[:k], want_xp, rtol=1e-5, atol=0)
    np.testing.assert_allclose(b.yp[:k], want_yp, rtol=1e-5, atol=0)
    # the plausible slips (beta gamma^2 -> beta gamma, a factor 2 in the
    # field, the slice width counted in bins) fall far outside that bound
    for slip in (gamma, 2.0, 0.5, 1.0 / 6):
        assert not np.allclose(b.xp[:k], want_xp * slip, rtol=1e-5, atol=0)
