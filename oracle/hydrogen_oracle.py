"""ORACLE -- test infrastructure only: a Float64 numpy restatement of Korg's
``hydrogen_line_absorption!`` (``src/hydrogen_line_absorption.jl:75-150``) for one layer, with the
Stehle & Hutcheon (1999) tables as Korg interpolates them (``src/hydrogen_line_absorption.jl:6-39``:
trilinear in (T, n_e, log(dnu / F0)) with flat extrapolation, the first log-detuning knot at
-floatmax, log 0 clipped to -700) and the MHD occupation probability ``hummer_mihalas_w``
(``src/statmech.jl:463-497``).  The Brackett-series part (``:152-169``) is not restated: its lines
lie beyond 1.44 um and reach no grid this checks.

Only the tests use this module, as the checker of the device library; the product never imports
it.  Tables are passed in as plain arrays (``transitions``: a list of dicts with ``lower``,
``upper``, ``log_gf``, ``temps``, ``nes``, ``dnu`` (delta_nu_over_F0), ``profile`` [dnu, ne, T] and
``lambda0`` [ne, T] in Angstrom, in the file's order).
"""
import math

import numpy as np

# src/constants.jl:3-19
kboltz_cgs = 1.380649e-16
kboltz_eV = 8.617333262145e-5
hplanck_eV = 4.135667696e-15
c_cgs = 2.99792458e10
electron_mass_cgs = 9.1093897e-28
electron_charge_cgs = 4.80320425e-10
amu_cgs = 1.6605402e-24
bohr_radius_cgs = 5.29177210903e-9
RydbergH_eV = 13.598287264
eV_to_cgs = 1.602e-12
H_MASS = 1.008 * amu_cgs  # src/atomic_data.jl (atomic mass of H) x amu
FLOATMAX = np.finfo(np.float64).max
FLOATMIN = np.finfo(np.float64).tiny


def hummer_mihalas_w(T, n_eff, nH, nHe, ne):
    """``src/statmech.jl:463-497`` without the Hubeny generalization."""
    r_level = math.sqrt(5 / 2 * n_eff ** 4 + 1 / 2 * n_eff ** 2) * bohr_radius_cgs
    neutral = nH * (r_level + math.sqrt(3) * bohr_radius_cgs) ** 3 + \
        nHe * (r_level + 1.02 * bohr_radius_cgs) ** 3
    if n_eff > 3:
        K = 16 / 3 * (n_eff / (n_eff + 1)) ** 2 * ((n_eff + 7 / 6) / (n_eff ** 2 + n_eff + 1 / 2))
    else:
        K = 1.0
    chi = RydbergH_eV / n_eff ** 2 * eV_to_cgs
    e = electron_charge_cgs
    charged = 16 * ((e ** 2) / (chi * math.sqrt(K))) ** 3 * ne
    return math.exp(-4 * math.pi / 3 * (neutral + charged))


def sigma_line(lam):
    """``src/line_absorption.jl:229-237``."""
    e, m, c = electron_charge_cgs, electron_mass_cgs, c_cgs
    return (math.pi * e ** 2 / m / c) * (lam ** 2 / c)


def doppler_width(lam0, T, m, xi):
    """``src/line_absorption.jl:189``."""
    return lam0 * math.sqrt(kboltz_cgs * T / m + (xi ** 2) / 2) / c_cgs


def scaled_vdW_abo(sigma, alpha, m, T):
    """``src/line_absorption.jl:208-221`` (ABO branch)."""
    v0 = 1e6
    inv_mu = 1 / (1.008 * amu_cgs) + 1 / m
    vbar = math.sqrt(8 * kboltz_cgs * T / math.pi * inv_mu)
    return 2 * (4 / math.pi) ** (alpha / 2) * math.gamma((4 - alpha) / 2) * v0 * sigma * \
        (vbar / v0) ** (1 - alpha)


def voigt_hjerting(a, v):
    """``src/line_absorption.jl:251-310`` (scalar)."""
    sqrt_pi = math.sqrt(math.pi)
    v2 = v * v

    def harris(v):
        H0 = math.exp(-v2)
        if v < 1.3:
            H1 = -1.12470432 + (-0.15516677 + (3.288675912 + (-2.34357915 + 0.42139162 * v) * v) * v) * v
        elif v < 2.4:
            H1 = -4.48480194 + (9.39456063 + (-6.61487486 + (1.98919585 - 0.22041650 * v) * v) * v) * v
        else:
            H1 = (0.554153432 + (0.278711796 + (-0.1883256872 + (0.042991293 - 0.003278278 * v) * v)
                                 * v) * v) / (v2 - 3 / 2)
        return H0, H1, (1 - 2 * v2) * H0

    if a <= 0.2 and v >= 5:
        invv2 = 1 / v2
        return (a / sqrt_pi * invv2) * (1 + 1.5 * invv2 + 3.75 * invv2 ** 2)
    if a <= 0.2:
        H0, H1, H2 = harris(v)
        return H0 + (H1 + H2 * a) * a
    if a <= 1.4 and a + v < 3.2:
        H0, H1, H2 = harris(v)
        M0 = H0
        M1 = H1 + 2 / sqrt_pi * M0
        M2 = H2 - M0 + 2 / sqrt_pi * M1
        M3 = 2 / (3 * sqrt_pi) * (1 - H2) - (2 / 3) * v2 * M1 + (2 / sqrt_pi) * M2
        M4 = 2 / 3 * v2 * v2 * M0 - 2 / (3 * sqrt_pi) * M1 + 2 / sqrt_pi * M3
        psi = 0.979895023 + (-0.962846325 + (0.532770573 - 0.122727278 * a) * a) * a
        return psi * (M0 + (M1 + (M2 + (M3 + M4 * a) * a) * a) * a)
    r2 = v2 / (a * a)
    a_invu = 1 / math.sqrt(2) / ((r2 + 1) * a)
    a2 = a_invu * a_invu
    return math.sqrt(2 / math.pi) * a_invu * (1 + (3 * r2 - 1 + ((r2 - 2) * 15 * r2 + 2) * a2) * a2)


def line_profile(lam0, sigma, gamma, amplitude, lam):
    """``src/line_absorption.jl:245-249``."""
    inv = 1 / (sigma * math.sqrt(2))
    scaling = inv / math.sqrt(math.pi) * amplitude
    return voigt_hjerting(gamma * inv, abs(lam - lam0) * inv) * scaling


def _interval(knots, x):
    """Gridded linear interpolation's cell and weight, flat beyond the ends."""
    n = len(knots)
    if x <= knots[0]:
        return 0, 0.0
    if x >= knots[-1]:
        return n - 2, 1.0
    i = int(np.searchsorted(knots, x, side="right")) - 1
    return i, (x - knots[i]) / (knots[i + 1] - knots[i])


def _bilinear(table, knots0, knots1, x0, x1):
    """table[i1][i0] (C order [ne][T]) at (x0 = T, x1 = n_e)."""
    i, s = _interval(knots0, x0)
    j, t = _interval(knots1, x1)
    lo = table[j][i] * (1 - s) + table[j][i + 1] * s
    hi = table[j + 1][i] * (1 - s) + table[j + 1][i + 1] * s
    return lo * (1 - t) + hi * t


def searchsortedfirst(wls, x):
    return int(np.searchsorted(wls, x, side="left")) + 1


def searchsortedlast(wls, x):
    return int(np.searchsorted(wls, x, side="right"))


def hydrogen_line_absorption(alpha_row, wls, T, ne, nH_I, nHe_I, UH_I, xi, window_size,
                             transitions, use_MHD=True):
    """Adds hydrogen-line opacity to ``alpha_row`` (the grid ``wls`` in cm), one layer."""
    nus = c_cgs / wls
    dnudl = c_cgs / wls ** 2
    n_max = max(t["upper"] for t in transitions)
    ws = [hummer_mihalas_w(T, n, nH_I, nHe_I, ne) if use_MHD else 1.0 for n in range(1, n_max + 1)]
    beta = 1 / (kboltz_eV * T)
    F0 = 1.25e-9 * ne ** (2 / 3)
    for tr in transitions:
        temps, nes = tr["temps"], tr["nes"]
        if not (temps[0] < T < temps[-1] and nes[0] < ne < nes[-1]):
            continue
        lam0 = _bilinear(tr["lambda0"] * 1e-8, temps, nes, T, ne)
        Elo = RydbergH_eV * (1 - 1 / tr["lower"] ** 2)
        Eup = RydbergH_eV * (1 - 1 / tr["upper"] ** 2)
        levels = ws[tr["upper"] - 1] * (math.exp(-beta * Elo) - math.exp(-beta * Eup)) / UH_I
        amplitude = 10.0 ** tr["log_gf"] * nH_I * sigma_line(lam0) * levels
        lb = searchsortedfirst(wls, lam0 - window_size)
        ub = searchsortedlast(wls, lam0 + window_size)
        if lb >= ub:
            continue
        if tr["lower"] == 2 and tr["upper"] in (3, 4, 5):
            lam0, sABO, aABO = {3: (6.56460998e-5, 1180.0, 0.677),
                                4: (4.8626810200000004e-5, 2320.0, 0.455),
                                5: (4.34168232e-5, 4208.0, 0.380)}[tr["upper"]]
            Gamma = scaled_vdW_abo(sABO * bohr_radius_cgs ** 2, aABO, H_MASS, T) * nH_I
            gamma = Gamma * lam0 ** 2 / (c_cgs * 4 * math.pi)
            sigma = doppler_width(lam0, T, H_MASS, xi)
            for j in range(lb - 1, ub):
                alpha_row[j] += line_profile(lam0, sigma, gamma, amplitude, wls[j])
        nu0 = c_cgs / lam0
        with np.errstate(divide="ignore"):
            logP = np.log(tr["profile"])
        logP[logP == -np.inf] = -700.0
        knots = np.concatenate([[-FLOATMAX], np.log(tr["dnu"][1:])])
        for j in range(lb - 1, ub):
            x = abs(nus[j] - nu0) / F0
            x = x + (x == 0) * FLOATMIN
            y = math.log(x)
            k, w = _interval(knots, y)
            a = _bilinear(logP[k], temps, nes, T, ne)
            b = _bilinear(logP[k + 1], temps, nes, T, ne)
            alpha_row[j] += math.exp(a * (1 - w) + b * w) * dnudl[j] * amplitude
    return alpha_row
