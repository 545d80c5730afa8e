"""The far field's expansion, checked on the CPU against the exact powers it replaces.

A far-wing series term c_k (d + t)^(-2k) (d: line centre to sub-tile centre, t: point offset from
the sub-tile centre, |t| <= h) is summed on the device as sum_{m < 12} W_km tau^m with
tau = t/h and W_km = binom(-2k, m) (h/d)^m d^(-2k).  The setup puts a sub-tile in a line's far
field only when its centre is at least 8 * 31.5 + 1 = 253 grid steps from the line centre, with
h = 31.5 steps (64-point sub-tiles on an evenly spaced grid), so |h/d| <= 31.5/253.  Here the
truncation of that expansion is measured over the whole range of (h/d, tau) and each k, relative
to each point's exact value, against the bounds stated in the source (2e-10 for the leading term,
1.5e-6 for the R^4 term), and the coefficients are checked against the binomial series of
(1 + x)^(-2k) term by term.
"""
import math

import numpy as np

TERMS = 12
U_MAX = 31.5 / 253.0


def w_coeffs(k, m_max=TERMS):
    """binom(-2k, m) for m < m_max by the recurrence the far-field kernel uses."""
    out = []
    b = 1.0
    for m in range(m_max):
        if m > 0:
            b *= -(2 * k + m - 1) / m
        out.append(b)
    return np.array(out)


def test_coefficients_are_the_binomial_series():
    for k in range(1, 5):
        c = w_coeffs(k)
        exact = [(-1) ** m * math.comb(2 * k + m - 1, m) for m in range(TERMS)]
        assert np.allclose(c, np.array(exact, dtype=float), rtol=1e-15, atol=0), k


def test_truncation_within_stated_bound():
    u = np.linspace(-U_MAX, U_MAX, 2001)[:, None]
    tau = np.linspace(-1.0, 1.0, 401)[None, :]
    x = u * tau
    for k, bound in ((1, 2.1e-10), (2, 1e-8), (3, 1.6e-7), (4, 1.5e-6)):
        c = w_coeffs(k)
        approx = sum(c[m] * x ** m for m in range(TERMS))
        exact = (1.0 + x) ** (-2 * k)
        err = float(np.max(np.abs(approx - exact) / exact))
        assert err <= bound, (k, err, bound)


def test_one_step_closer_breaks_the_k4_bound():
    """The margin matters: at ratio 4 instead of 8 the 12-term expansion of the R^4 term is off
    by more than 1e-4, which a dropped margin or a halved distance in the setup would cause."""
    u = np.linspace(-1 / 4, 1 / 4, 2001)[:, None]
    tau = np.linspace(-1.0, 1.0, 401)[None, :]
    x = u * tau
    c = w_coeffs(4)
    approx = sum(c[m] * x ** m for m in range(TERMS))
    exact = (1.0 + x) ** (-8)
    assert float(np.max(np.abs(approx - exact) / exact)) > 1e-4
