"""Longitudinal resonator physics from the impedance, by quadrature.

Nothing here evaluates the time-domain wake function that the library and
the oracle use: every number comes from the resonator impedance in the
circuit convention (Z(w_r) = Rs),

    Z(w) = Rs / (1 + i Q (w/w_r - w_r/w)),

through textbook relations:
* point-charge loss factor  k = (1/pi) int_0^inf Re Z(w) dw    (V/C);
* wake of a point charge    W(tau) = (2/pi) int_0^inf Re Z(w) cos(w tau) dw
  for tau > 0 (causality: the real part fixes the wake);
* loss factor of a Gaussian bunch of rms duration st
                            k(st) = (1/pi) int_0^inf Re Z(w) exp(-w^2 st^2) dw.
The beam-loading theorem ties them: a charge q alone loses q^2 k of energy,
so it sees the potential q k = q W(0+)/2; a bunch of charge Q loses Q^2 k(st).
In the element's units (charge in C, energy in MeV) one electron of a
macroparticle of charge q_m that sees the potential V changes its relative
momentum by -V / (E [MeV] * 1e6).

Leading particles have the larger z: the reference's ATransWake sums the
wake of the particles j with z_j >= z_i onto particle i
(accel.cpp:1634-1646), and its drift lowers z for a longer path
(atpass.cpp:416-427).
"""
import math

import numpy as np
from scipy.integrate import quad

C0 = 2.99792458e8


def _re_z(rs, q):
    # Re Z at x = w/w_r, written without division by x (Re Z(0) = 0)
    return lambda x: rs * x * x / (x * x + q * q * (x * x - 1.0) ** 2)


def loss_factor_point(rs, q, fr):
    wr = 2 * math.pi * fr
    return wr / math.pi * quad(_re_z(rs, q), 0, np.inf, limit=500, epsabs=0, epsrel=1e-13)[0]


def wake(rs, q, fr, tau):
    assert tau > 0
    wr = 2 * math.pi * fr
    return 2 / math.pi * wr * quad(_re_z(rs, q), 0, np.inf, weight="cos", wvar=wr * tau, limlst=200)[0]


def loss_factor_gaussian(rs, q, fr, sigma_z):
    wr = 2 * math.pi * fr
    st = sigma_z / C0
    f = _re_z(rs, q)
    upper = 12.0 / (wr * st)  # exp(-144): the rest of the integral is below round-off
    return wr / math.pi * quad(lambda x: f(x) * math.exp(-(x * wr * st) ** 2), 0, upper, limit=500, epsabs=0,
                               epsrel=1e-12)[0]


def quasi_gaussian_z(n, sigma_z):
    """n longitudinal positions at the normal quantiles (i + 1/2)/n: a
    smooth sample with no statistical noise."""
    from scipy.special import ndtri
    beam = np.zeros((n, 6))
    beam[:, 4] = sigma_z * ndtri((np.arange(n) + 0.5) / n)
    return beam
