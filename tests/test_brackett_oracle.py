"""Pins of the Brackett-series restatement in oracle/hydrogen_oracle.py
(src/hydrogen_line_absorption.jl:152-470).  CPU only.

* bracket_line_interpolator: the reference's own test (test/line_absorption.jl:50-72) -- the
  convolved profile of Br gamma peaks within window/500 of the line centre.
* brackett_oscillator_strength: against exact hydrogen oscillator strengths 4 -> m, summed over l
  from radial integrals of the analytic wavefunctions (numerical quadrature), within what the fit
  achieves: 1e-4 for Br alpha (as the reference states against Goldwire 1968), 4e-3 above.
* exponential_integral_1: against scipy's E1 within 1e-4 (the accuracy of its approximations).
* holtsmark_profile: the fit to the quasistatic line profile (SOFBET) has unit two-sided norm, the
  1.5 beta^(-5/2) asymptote, and its three pieces (inner tables, medium asymptote with C7/D7,
  far asymptote) meet continuously at beta = 25.12 and 500 for every shielding knot -- which ties
  the last row of PROB7 to C7 and D7.
* The line's convolved profile integrates (over a window wide enough to hold it) to the product of
  its components' integrals, as a convolution must.
Greim's K_nm beyond the tabulated members and the profile's digits beyond these properties are
unpinned (no reference data).
"""
import math

import numpy as np
import pytest
from scipy import integrate, special

from oracle import hydrogen_oracle as ho


def test_convolved_profile_peaks_at_line_centre():
    m = 7
    E = ho.RydbergH_eV * (1 / 16 - 1 / m ** 2)
    lam0 = ho.hplanck_eV * ho.c_cgs / E
    start, step, prof, window = ho.bracket_line_interpolator(m, lam0, 8000.0, 1e14, 1e5)
    test = np.linspace(lam0 - window / 10, lam0 + window / 10, 10001)
    vals = ho._interp_range(start, step, prof, test)
    assert abs(test[np.argmax(vals)] - lam0) < window / 500


def _radial(n, l, r):
    """Normalized hydrogen radial function R_nl(r) (atomic units)."""
    rho = 2 * r / n
    norm = math.sqrt((2 / n) ** 3 * math.factorial(n - l - 1) / (2 * n * math.factorial(n + l)))
    return norm * np.exp(-rho / 2) * rho ** l * special.eval_genlaguerre(n - l - 1, 2 * l + 1, rho)


def _exact_f(n, m):
    """Absorption oscillator strength n -> m of hydrogen, summed over the l sublevels."""
    dE = 0.5 * (1 / n ** 2 - 1 / m ** 2)  # Hartree
    total = 0.0
    rmax = 4 * m * m + 60
    for l in range(n):
        for lp in (l - 1, l + 1):
            if lp < 0 or lp >= m:
                continue
            R, _ = integrate.quad(lambda r: _radial(n, l, r) * _radial(m, lp, r) * r ** 3, 0, rmax,
                                  limit=400, epsabs=0, epsrel=1e-11)
            total += (2 * l + 1) * (2 / 3) * dE * max(l, lp) / (2 * l + 1) * R * R
    return total / n ** 2


def test_brackett_oscillator_strengths_match_exact_values():
    assert _exact_f(2, 3) == pytest.approx(0.6407, rel=1e-3)  # H alpha: the quadrature itself
    # the fit's own accuracy: 4e-5 for Br alpha, a smooth 3e-3 at most from m = 8 on
    assert ho.brackett_oscillator_strength(4, 5) == pytest.approx(_exact_f(4, 5), rel=1e-4)
    for m in (6, 7, 8, 10, 15, 20, 30):
        assert ho.brackett_oscillator_strength(4, m) == pytest.approx(_exact_f(4, m), rel=4e-3), m


@pytest.mark.parametrize("x", [1e-4, 3e-3, 0.01, 0.05, 0.3, 0.99, 1.0, 1.5, 4.0, 12.0, 29.0])
def test_exponential_integral_1(x):
    assert ho.exponential_integral_1(x) == pytest.approx(special.exp1(x), rel=1e-4)
    assert ho.exponential_integral_1(31.0) == 0.0 and ho.exponential_integral_1(-1.0) == 0.0


@pytest.mark.parametrize("P", [0.0, 0.2, 0.4, 0.6, 0.8])
def test_holtsmark_pieces_meet(P):
    e = 1e-9
    # at 25.12 the table's last row and C7/D7 agree to 5e-4 (a slipped digit in either shows as a
    # jump of ~1e-2); at 500 the medium piece's correction DD/(CC + beta^1.5) is left, <= 2.2e-3
    for b, tol in ((25.12, 1e-3), (500.0, 3e-3)):
        lo, hi = ho.holtsmark_profile(b - e, P), ho.holtsmark_profile(b + e, P)
        assert abs(lo - hi) <= tol * hi, (b, lo, hi)
    far = 1e4
    assert ho.holtsmark_profile(far, P) == pytest.approx(1.5 * far ** -2.5, rel=1e-3)


def test_holtsmark_norm():
    pts = [1.0, 8.0, 10.0, 25.12, 500.0]
    tot, _ = integrate.quad(lambda b: ho.holtsmark_profile(b, 0.0), 0, 1e6, points=pts, limit=800)
    tot += (1.5 / 1.5) * 1e6 ** -1.5  # the tail beyond 1e6 (leading term)
    assert tot == pytest.approx(0.5, rel=5e-3)


def test_convolution_preserves_integrals():
    m = 5
    E = ho.RydbergH_eV * (1 / 16 - 1 / m ** 2)
    lam0 = ho.hplanck_eV * ho.c_cgs / E
    T, ne = 5000.0, 1e11  # a Doppler width well above the Stark width: all three convolved
    start, step, prof, window = ho.bracket_line_interpolator(m, lam0, T, ne, 1e5, window_size=40,
                                                             n_points=2001)
    lams = start + step * np.arange(prof.size)
    impact, quasi = ho.brackett_line_stark_profiles(m, lams, lam0, T, ne)
    sdop = ho.doppler_width(lam0, T, ho.H_MASS, 1e5)
    assert sdop / (1.6678e-18 * ho.greim_1960_Knm(4, m) * 1.25e-9 * ne ** (2 / 3) * ho.c_cgs) > 0.25
    # the Gaussian integrates to 1 over the window, so the profile's integral is that of the
    # product of the two Stark components' integrals
    assert np.sum(prof) * step == pytest.approx(np.sum(impact) * step * np.sum(quasi) * step,
                                                rel=2e-2)
