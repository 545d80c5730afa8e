"""ORACLE -- test infrastructure only: a Float64 numpy restatement of Korg's
``hydrogen_line_absorption!`` (``src/hydrogen_line_absorption.jl:75-150``) for one layer, with the
Stehle & Hutcheon (1999) tables as Korg interpolates them (``src/hydrogen_line_absorption.jl:6-39``:
trilinear in (T, n_e, log(dnu / F0)) with flat extrapolation, the first log-detuning knot at
-floatmax, log 0 clipped to -700) and the MHD occupation probability ``hummer_mihalas_w``
(``src/statmech.jl:463-497``), and the Brackett series (``:152-169``): its Stark profiles
(``brackett_line_stark_profiles``, ``:259-351``: electron impact + quasistatic Holtsmark, ``:391-445``)
convolved with each other and, when the Doppler width matters, with a Gaussian on a 201-point grid
(``bracket_line_interpolator``, ``:211-249``), then interpolated linearly onto the grid.  The
convolutions are direct sums (the reference uses DSP.conv; the two agree to rounding).

Only the tests use this module, as the checker of the device library; the product never imports
it.  Tables are passed in as plain arrays (``transitions``: a list of dicts with ``lower``,
``upper``, ``log_gf``, ``temps``, ``nes``, ``dnu`` (delta_nu_over_F0), ``profile`` [dnu, ne, T] and
``lambda0`` [ne, T] in Angstrom, in the file's order).
"""
import math

import numpy as np

# src/constants.jl:3-19
kboltz_cgs = 1.380649e-16
hplanck_cgs = 6.62607015e-27
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


def exponential_integral_1(x):
    """``src/line_absorption.jl:169-182`` (piecewise approximation of E1)."""
    if x < 0:
        return 0.0
    if x <= 0.01:
        return -math.log(x) - 0.577215 + x
    if x <= 1.0:
        return -math.log(x) - 0.57721566 + \
            x * (0.99999193 + x * (-0.24991055 + x * (0.05519968 + x * (-0.00976004 + x * 0.00107857))))
    if x <= 30.0:
        return (x * (x + 2.334733) + 0.25062) / (x * (x + 3.330657) + 1.681534) / x * math.exp(-x)
    return 0.0


def brackett_oscillator_strength(n, m):
    """``src/hydrogen_line_absorption.jl:180-191``."""
    GINF = 0.2027 / n ** 0.71
    GCA = 0.124 / n
    FKN = 1.9603 * n
    WTC = 0.45 - 2.4 / n ** 3 * (n - 1)
    FK = FKN * (m / ((m - n) * (m + n))) ** 3
    XMN12 = (m - n) ** 1.2
    WT = (XMN12 - 1) / (XMN12 + WTC)
    return FK * (1 - WT * GINF - (0.222 + GCA / m) * (1 - WT))


_GRIEM_KMN = [[0.0001716, 0.0090190, 0.1001000, 0.5820000],
              [0.0005235, 0.0177200, 0.1710000, 0.8660000],
              [0.0008912, 0.0250700, 0.2230000, 1.0200000]]


def greim_1960_Knm(n, m):
    """``src/hydrogen_line_absorption.jl:370-389``."""
    if m - n <= 3 and n <= 4:
        return _GRIEM_KMN[m - n - 1][n - 1]
    return 5.5e-5 * n ** 4 * m ** 4 / (m ** 2 - n ** 2) / (1 + 0.13 / (m - n))


_H_PROB7 = np.array([[0.005, 0.128, 0.260, 0.389, 0.504], [0.004, 0.109, 0.220, 0.318, 0.389],
                     [-0.007, 0.079, 0.162, 0.222, 0.244], [-0.018, 0.041, 0.089, 0.106, 0.080],
                     [-0.026, -0.003, 0.003, -0.023, -0.086], [-0.025, -0.048, -0.087, -0.148, -0.234],
                     [-0.008, -0.085, -0.165, -0.251, -0.343], [0.018, -0.111, -0.223, -0.321, -0.407],
                     [0.032, -0.130, -0.255, -0.354, -0.431], [0.014, -0.148, -0.269, -0.359, -0.427],
                     [-0.005, -0.140, -0.243, -0.323, -0.386], [0.005, -0.095, -0.178, -0.248, -0.307],
                     [-0.002, -0.068, -0.129, -0.187, -0.241], [-0.007, -0.049, -0.094, -0.139, -0.186],
                     [-0.010, -0.036, -0.067, -0.103, -0.143]])
_H_C7 = [511.318, 1.532, 4.044, 19.266, 41.812]
_H_D7 = [-6.070, -4.528, -8.759, -14.984, -23.956]
_H_PP = [0.0, 0.2, 0.4, 0.6, 0.8]
_H_BETA = [1.0, 1.259, 1.585, 1.995, 2.512, 3.162, 3.981, 5.012, 6.310, 7.943, 10.0, 12.59,
           15.85, 19.95, 25.12]


def holtsmark_profile(beta, P):
    """``src/hydrogen_line_absorption.jl:397-445`` (0-based table indices)."""
    if beta > 500:
        return (1.5 / math.sqrt(beta) + 27 / beta ** 2) / beta ** 2
    im = min(int(math.floor(5 * P + 1)), 4) - 1
    ip = im + 1
    WTPP = 5 * (P - _H_PP[im])
    WTPM = 1 - WTPP
    if beta <= 25.12:
        jp = max(1, next(i for i, k in enumerate(_H_BETA) if beta <= k))
        jm = jp - 1
        WTBP = (beta - _H_BETA[jm]) / (_H_BETA[jp] - _H_BETA[jm])
        WTBM = 1 - WTBP
        CBP = _H_PROB7[jp, ip] * WTPP + _H_PROB7[jp, im] * WTPM
        CBM = _H_PROB7[jm, ip] * WTPP + _H_PROB7[jm, im] * WTPM
        CORR = 1 + CBP * WTBP + CBM * WTBM
        WT = max(min(0.5 * (10 - beta), 1), 0)
        PR1 = 8 / (83 + (2 + 0.95 * beta ** 2) * beta) if beta <= 10 else 0.0
        PR2 = (1.5 / math.sqrt(beta) + 27 / beta ** 2) / beta ** 2 if beta >= 8 else 0.0
        return (PR1 * WT + PR2 * (1 - WT)) * CORR
    CC = _H_C7[ip] * WTPP + _H_C7[im] * WTPM
    DD = _H_D7[ip] * WTPP + _H_D7[im] * WTPM
    CORR = 1 + DD / (CC + beta * math.sqrt(beta))
    return (1.5 / math.sqrt(beta) + 27 / beta ** 2) / beta ** 2 * CORR


def brackett_line_stark_profiles(m, lams, lam0, T, ne):
    """``src/hydrogen_line_absorption.jl:272-351``: the electron-impact and quasistatic profiles
    of the Brackett line to level m at the wavelengths ``lams`` (cm, ascending)."""
    n = 4
    nus = c_cgs / lams
    nu0 = c_cgs / lam0
    ne_1_6 = ne ** (1 / 6)
    F0 = 1.25e-9 * ne ** (2 / 3)
    GCON1 = 0.2 + 0.09 * math.sqrt(T / 10_000) / (1 + ne / 1.0e13)
    GCON2 = 0.2 / (1 + ne / 1.0e15)
    Knm = greim_1960_Knm(n, m)
    Y1WHT = 1e14 if m - n <= 3 else 1e13
    WTY1 = 1 / (1 + ne / Y1WHT)
    Y1B = 2 / (1 + 0.012 / T * math.sqrt(ne / T))
    C1CON = Knm / lam0 * (m ** 2 - n ** 2) ** 2 / (n ** 2 * m ** 2) * 1e-8
    Y1NUM = 320
    Y1SCAL = Y1NUM * ((T / 10_000) ** 0.3 / ne_1_6) * WTY1 + Y1B * (1 - WTY1)
    C1 = F0 * 78940 / T * C1CON * Y1SCAL
    C2 = F0 ** 2 / (5.96e-23 * ne) * (Knm / lam0) ** 2 * 1e-16
    betas = np.abs(lams - lam0) / F0 / Knm * 1e8
    G1 = 6.77 * math.sqrt(C1)
    impact = np.zeros(lams.size)
    quasi = np.zeros(lams.size)
    PP = ne_1_6 * 0.08989 / math.sqrt(T)
    for k, b in enumerate(betas):
        y1 = C1 * b
        y2 = C2 * b ** 2
        if y2 <= 1e-4 and y1 <= 1e-5:
            width = G1 * max(0.0, 0.2114 + math.log(math.sqrt(C2) / C1)) * (1 - GCON1 - GCON2)
        else:
            GAM = G1 * (0.5 * math.exp(-min(80.0, y1)) + exponential_integral_1(y1) -
                        0.5 * exponential_integral_1(y2)) * \
                (1 - GCON1 / (1 + (90 * y1) ** 3) - GCON2 / (1 + 2000 * y1))
            width = 0.0 if GAM <= 1e-20 else GAM
        impact[k] = width / (math.pi * (width ** 2 + b ** 2)) if width > 0 else 0.0
        ps = (0.9 * y1) ** 2
        qe = (ps + 0.03 * math.sqrt(y1)) / (ps + 1.0)
        quasi[k] = holtsmark_profile(b, PP) * (1 + qe)
    dbeta_dlam = 1e8 / (Knm * F0)
    red = bool(np.any(nus < nu0))  # findlast(nus .< nu0) exists: every point is rescaled
    for prof in (impact, quasi):
        prof *= np.sqrt(lams / lam0)
        if red:
            prof *= np.exp((hplanck_cgs * (nus - nu0)) / kboltz_cgs / T)
        prof *= dbeta_dlam
    return impact, quasi


def _conv_center(a, b):
    """The central len(a) values of the full convolution (``conv(a, b)[start:start+n-1]``,
    start = (n + 1) / 2, 1-based).  Direct sums: every term is non-negative, so any summation order
    agrees to rounding."""
    n = a.size
    return np.convolve(a, b)[(n - 1) // 2:(n - 1) // 2 + n]


def bracket_line_interpolator(m, lam0, T, ne, xi, lmin=0.0, lmax=math.inf, n_points=201,
                              window_size=5, include_doppler_threshold=0.25):
    """``src/hydrogen_line_absorption.jl:211-249``: (first wavelength, step, convolved profile,
    window); the profile is interpolated linearly between its wavelengths start + k step.
    (None, None, None, 0.0) when the window misses [lmin, lmax]."""
    n = 4
    F0 = 1.25e-9 * ne ** (2 / 3)
    Knm = greim_1960_Knm(n, m)
    stark_width = 1.6678e-18 * Knm * F0 * c_cgs
    sdop = doppler_width(lam0, T, H_MASS, xi)
    window = window_size * max(sdop, stark_width)
    start = max(lmin, lam0 - window)
    end = min(lam0 + window, lmax)
    if start > lmax or end < lmin or start == end:
        return None, None, None, 0.0
    step = (end - start) / (n_points - 1)
    lams = start + step * np.arange(n_points)
    impact, quasi = brackett_line_stark_profiles(m, lams, lam0, T, ne)
    if sdop / stark_width > include_doppler_threshold:
        dop = np.exp(-0.5 * (lams - lam0) ** 2 / sdop ** 2) / math.sqrt(2 * math.pi) / sdop
        quasi = _conv_center(quasi, dop) * step
    return start, step, _conv_center(impact, quasi) * step, window


def _interp_range(start, step, vals, x):
    """Linear interpolation of vals at start + k step (``linear_interpolation(range, vals)``)."""
    t = (x - start) / step
    i = np.clip(np.floor(t).astype(np.int64), 0, vals.size - 2)
    w = t - i
    return vals[i] * (1 - w) + vals[i + 1] * w


def brackett_lines(alpha_row, wls, T, ne, nH_I, UH_I, xi, ws, n_max):
    """``src/hydrogen_line_absorption.jl:152-169``: adds the Brackett lines 4 -> m, m = 5 ..
    n_max, to one layer; ``ws[k]`` is the occupation probability of level k + 1."""
    n = 4
    beta = 1 / (kboltz_eV * T)
    E_low = RydbergH_eV * (1 - 1 / n ** 2)
    for m in range(5, n_max + 1):
        E = RydbergH_eV * (1 / n ** 2 - 1 / m ** 2)
        lam0 = hplanck_eV * c_cgs / E
        levels = ws[m - 1] * math.exp(-beta * E_low) * (1 - math.exp(-beta * E)) / UH_I
        gf = 2 * n ** 2 * brackett_oscillator_strength(n, m)
        amplitude = gf * nH_I * sigma_line(lam0) * levels
        start, step, prof, window = bracket_line_interpolator(m, lam0, T, ne, xi, wls[0], wls[-1])
        lb = searchsortedfirst(wls, lam0 - window)
        ub = searchsortedlast(wls, lam0 + window)
        if prof is None or lb > ub:
            continue
        sl = slice(lb - 1, ub)
        alpha_row[sl] += _interp_range(start, step, prof, wls[sl]) * amplitude
    return alpha_row


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
    brackett_lines(alpha_row, wls, T, ne, nH_I, UH_I, xi, ws, n_max)
    return alpha_row
