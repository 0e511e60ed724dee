"""Seeded bunches and one-turn maps of the benchmark workloads, for the
oracle (test infrastructure only).

Bunch generation restates ParticleGenerator (particles/generators.py:
299-405): np.random.seed(seed), then normal draws for x, xp, y, yp, z, dp
in that order, then linear matching x = sqrt(beta) u,
xp = -alpha/sqrt(beta) u + sqrt(1/beta) u' (:124-146).  The Gaussian of the
non-linear-RF workload is redrawn outside the separatrix with margin 0.05
(cut_distribution, :220-259; Synchrotron.generate_6D_Gaussian_bunch,
machines/synchrotron.py:309-364) using the closed form of a stationary
single-harmonic bucket above transition.

Workload parameters:
  protons, p0 = sqrt(27.7^2 - 1) m_p c (gamma from the reference's p0
  setter), C = 6911.5 m, beta_x = beta_y = 54.6 m, alpha = 0,
  D = 0, eps_n = 2.5e-6 m, Q_x = 20.13, Q_y = 20.18, alpha_c = 1.9e-3;
  linear:    LinearMap with Q_s = 0.017, sigma_z = 0.23 m, sigma_dp = 1e-3
             drawn as stated (no longitudinal matching), 50 slices over
             +-3 sigma_z;
  nonlinear: RFSystems h = 4620, V = 4.5 MV, dphi = 0; Q'_x = Q'_y = 5;
             amplitude detuning app_x = app_y = 1e-10*p0, app_xy = -5e-11*p0
             (the reference divides by p0 in the detuner); intensity 1.5e11;
             500 slices over +-3 sigma_z; CircularResonator(R_s = 1e6 Ohm/m,
             f_r = 1e9 Hz, Q = 1, one turn of wake memory); sigma_dp from
             the matched beta_z of the bucket's Q_s.
"""
import numpy as np
from scipy.constants import c, e, m_p

from . import refpath as R

# The reference's machines take the momentum: Synchrotron(p0=...) sets gamma
# through its p0 setter, 1/(c m) sqrt(p0^2 + m^2 c^2) (machines/synchrotron.py:
# 278-280), which gives 27.700000000000003 for the nominal gamma 27.7.
P0 = np.sqrt(27.7**2 - 1) * m_p * c


def gamma_from_p0(p0, mass=m_p):
    return 1 / (c * mass) * np.sqrt(p0**2 + mass**2 * c**2)


GAMMA = gamma_from_p0(P0)
CIRCUMFERENCE = 6911.5
BETA_X = BETA_Y = 54.6
EPSN = 2.5e-6
Q_X, Q_Y = 20.13, 20.18
ALPHA_C = 1.9e-3
SIGMA_Z = 0.23
H_RF, V_RF = 4620, 4.5e6


def p0_of(gamma, mass=m_p):
    return np.sqrt(gamma**2 - 1) * mass * c


def q_s_single_harmonic(gamma, alpha_c, h, V, charge=e, mass=m_p):
    """RFBucket.Q_s (trackers/rf_bucket.py:268-279) for one harmonic."""
    beta = np.sqrt(1 - gamma**-2)
    p0 = np.sqrt(gamma**2 - 1) * mass * c
    eta = alpha_c - gamma**-2
    return np.sqrt(np.abs(charge) * np.abs(eta) * h * V / (2 * np.pi * p0 * beta * c))


def stationary_acceptance(gamma, alpha_c, h, V, C, margin, charge=e, mass=m_p):
    """is_in_separatrix(margin) of a stationary single-harmonic bucket above
    transition: H(z, dp) = -eta beta c dp^2/2 + V(z)/p0 with
    V(z) = |q|/C * R/h * V (cos(h z/R) - cos(pi)) (potential zero at the
    unstable fixed points z = +-C/(2h)); sign(eta) > 0 makes H(convex) =
    H; accepted if |z| < C/(2h) and H > margin*H(0, 0)."""
    beta = np.sqrt(1 - gamma**-2)
    p0 = np.sqrt(gamma**2 - 1) * mass * c
    eta = alpha_c - gamma**-2
    Rr = C / (2 * np.pi)
    zmax = C / (2. * h)

    def pot(z):
        return np.abs(charge) / C * (Rr / h * V) * (np.cos(h * z / Rr) + 1.)

    def ham(z, dp):
        return -0.5 * eta * beta * c * dp**2 + pot(z) / p0

    h0 = ham(0., 0.)

    def accepted(z, dp):
        return np.logical_and(np.abs(z) < zmax, np.sign(eta) * ham(z, dp) >
                              margin * np.sign(eta) * h0)
    return accepted


def gaussian_bunch(n, seed, sigma_dp, intensity=1.5e11, gamma=GAMMA, accept=None,
                   sigma_z=SIGMA_Z):
    np.random.seed(seed)
    betagamma = np.sqrt(gamma**2 - 1)
    eps_geo = EPSN / betagamma
    std = np.sqrt(eps_geo)
    x, xp = np.random.normal(0., std, n), np.random.normal(0., std, n)
    y, yp = np.random.normal(0., std, n), np.random.normal(0., std, n)
    z, dp = np.random.normal(0., sigma_z, n), np.random.normal(0., sigma_dp, n)
    if accept is not None:
        out = ~accept(z, dp)
        while out.any():
            k = int(np.sum(out))
            z[out] = np.random.normal(0., sigma_z, k)
            dp[out] = np.random.normal(0., sigma_dp, k)
            out = ~accept(z, dp)
    coords = {}
    for u, up, name in ((x, xp, 'x'), (y, yp, 'y')):
        beta = BETA_X if name == 'x' else BETA_Y
        coords[name] = np.sqrt(beta) * u + 0 * up
        coords[name + 'p'] = -0. / np.sqrt(beta) * u + np.sqrt(1. / beta) * up
    coords['z'], coords['dp'] = z, dp
    return R.Bunch(n, intensity / n, e, m_p, CIRCUMFERENCE, gamma, coords)


def linear_bunch(n, seed=12345):
    """Longitudinally unmatched Gaussian as stated: sigma_dp = 1e-3."""
    return gaussian_bunch(n, seed, 1e-3)


def nonlinear_sigma_dp(gamma=GAMMA):
    Qs = q_s_single_harmonic(gamma, ALPHA_C, H_RF, V_RF)
    eta = ALPHA_C - gamma**-2
    beta_z = np.abs(eta) * CIRCUMFERENCE / 2.0 / np.pi / Qs
    return SIGMA_Z / beta_z


def nonlinear_bunch(n, seed=12345):
    acc = stationary_acceptance(GAMMA, ALPHA_C, H_RF, V_RF, CIRCUMFERENCE, 0.05)
    return gaussian_bunch(n, seed, nonlinear_sigma_dp(), accept=acc)


def linear_map_oracle(n_slices=50):
    """[LinearMap, TransverseSegment] (RF_at='middle', one segment) and the
    50-slice slicing recorded every turn."""
    seg = R.smooth_segment(Q_X, Q_Y, BETA_X, BETA_Y)
    lin = R.LinearMapOracle([ALPHA_C], CIRCUMFERENCE, 0.017)
    return [lin, seg]


def nonlinear_map_oracle(n_slices=500, n_turns_wake=1, p0=None):
    p0 = P0 if p0 is None else p0
    seg = R.smooth_segment(Q_X, Q_Y, BETA_X, BETA_Y, Qp_x=5., Qp_y=5.,
                           app_x=1e-10 * p0, app_y=1e-10 * p0, app_xy=-5e-11 * p0)
    rf = R.RFSystemsOracle(CIRCUMFERENCE, [H_RF], [V_RF], [0.], [ALPHA_C], GAMMA, e)
    wx = R.resonator_transverse(1., 1e6, 1e9, 1.)
    wy = R.resonator_transverse(1., 1e6, 1e9, 1.)
    kicks = [R.Kick('DipoleX', wx, n_turns_wake), R.Kick('DipoleY', wy, n_turns_wake)]
    wake = R.WakeFieldOracle(n_slices, kicks, n_sigma_z=3)
    return [rf, seg, wake]
