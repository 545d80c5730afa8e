"""Pins of the Float64 CPU oracle (oracle/line_oracle.c) against the reference's own test
assertions and against what the mathematics and physics fix.  CPU only.

Reference tests restated here:
  test/line_absorption.jl:4-17   profile normalisation and near-monotonicity
  test/wavelengths.jl:60-86      searchsortedfirst / searchsortedlast known indices
Closed forms / special cases (no golden alpha exists in the reference for metal lines):
  Voigt: a = 0 gives exp(-v^2) (v < 5); large a gives the Lorentzian; bounded distance from the
  exact Voigt function Re w(v + ia) (scipy.special.wofz); H1 close to the Dawson-function form.
  Inverse densities: pdf(inverse(rho)) == rho.  Doppler width, Stark/vdW temperature laws,
  ABO alpha = 0 hard-sphere limit, the classical line cross-section pi e^2/(m_e c).
  Whole path: integrated line opacity == amplitude, Doppler FWHM == 2 sqrt(2 ln 2) sigma,
  Lorentz FWHM == 2 gamma, and both thread schedules agree.
"""
import math

import numpy as np
import pytest
from scipy.special import dawsn, gamma as gamma_fn, wofz

from oracle import oracle
from korg_line import Line, LineList, Problem, Species, Wavelengths
from korg_line.constants import (amu_cgs, c_cgs, electron_charge_cgs, electron_mass_cgs,
                                 hplanck_eV, kboltz_cgs, kboltz_eV)
from korg_line.species import atomic_masses


# ----------------------------------------------------------------------------- reference tests
@pytest.mark.parametrize("sigma", [1e-7, 1e-8, 1e-9])
@pytest.mark.parametrize("gam", [3e-8, 3e-9, 3e-10])
def test_line_profile_reference_properties(sigma, gam):
    """test/line_absorption.jl:4-17, same grid (4750:0.01:5250 A), centre 5e-5 cm, amplitude 7."""
    wls = Wavelengths((4750, 5250)).all_wls
    delta = 0.01 * 1e-8
    amplitude = 7.0
    phi = np.array([oracle.line_profile(5e-5, sigma, gam, amplitude, w) for w in wls])
    half = int(math.ceil(phi.size / 2))
    assert np.all(np.diff(phi[:half]) > -1e-3 * phi.max())
    assert np.all(np.diff(phi[half - 1:]) < 1e-3 * phi.max())
    assert 0.98 < np.sum(phi * delta) / amplitude < 1


@pytest.mark.parametrize("conv", [1, 1e-8])
def test_searchsorted_reference_indices(conv):
    """test/wavelengths.jl:60-86 (targets in A and in cm)."""
    w = Wavelengths(0.5 + np.arange(2, 11)).all_wls
    assert oracle.searchsortedfirst(w, 4.0 * conv) == 3
    assert oracle.searchsortedfirst(w, 1.5 * conv) == 1
    assert oracle.searchsortedfirst(w, 3 * conv) == 2
    assert oracle.searchsortedlast(w, 8 * conv) == 6
    assert oracle.searchsortedlast(w, 4 * conv) == 2
    assert oracle.searchsortedlast(w, 11 * conv) == 9
    w = Wavelengths((1, 10, 1.0)).all_wls
    assert oracle.searchsortedfirst(w, 5.4 * conv) == 6
    assert oracle.searchsortedlast(w, 5.6 * conv) == 5
    w = Wavelengths([(3, 5, 1.0), (11, 12.5, 0.5), (16, 20, 1.0)]).all_wls
    assert oracle.searchsortedfirst(w, 2 * conv) == 1
    assert oracle.searchsortedfirst(w, 11.9 * conv) == 6
    assert oracle.searchsortedfirst(w, 45 * conv) == 13
    assert oracle.searchsortedlast(w, 2 * conv) == 0
    assert oracle.searchsortedlast(w, 4 * conv) == 2
    assert oracle.searchsortedlast(w, 11 * conv) == 4
    assert oracle.searchsortedlast(w, 13.1 * conv) == 7
    assert oracle.searchsortedlast(w, 55 * conv) == 12


# ----------------------------------------------------------------------------- Voigt function
def test_voigt_doppler_limit():
    """a = 0: H = exp(-v^2) for v < 5 (branch B reduces to H0), 0 in branch A."""
    for v in np.linspace(0, 4.99, 200):
        assert oracle.voigt_hjerting(0.0, v) == pytest.approx(math.exp(-v * v), rel=1e-15)
    for v in (5.0, 7.0, 100.0):
        assert oracle.voigt_hjerting(0.0, v) == 0.0


def test_voigt_lorentz_limit():
    """a >> 1: H -> a / (sqrt(pi) (a^2 + v^2)), the Lorentzian."""
    for a in (20.0, 100.0):
        for v in (0.0, 3.0, 30.0, 300.0):
            ref = a / (math.sqrt(math.pi) * (a * a + v * v))
            assert oracle.voigt_hjerting(a, v) == pytest.approx(ref, rel=2e-3)


@pytest.mark.parametrize("a,bound", [(0.0, 1e-10), (0.01, 1e-4), (0.1, 2e-3), (0.2, 1e-2),
                                     (0.3, 5e-3), (0.5, 6e-3), (1.0, 4e-2), (1.4, 0.13),
                                     (1.5, 0.08), (3.0, 4e-3), (10.0, 1e-4)])
def test_voigt_close_to_exact_faddeeva(a, bound):
    """Distance of Hunger's approximation from the exact H(a, v) = Re w(v + ia), relative to
    H(a, 0).  The bounds are the approximation's own accuracy (worst near a = 1.4, ~12%), so a
    dropped term or wrong sign in any branch exceeds them."""
    peak = wofz(1j * a).real
    v = np.linspace(0, 12, 1201)
    err = np.array([abs(oracle.voigt_hjerting(a, x) - wofz(x + 1j * a).real) for x in v]) / peak
    assert err.max() < bound


def test_harris_series_against_dawson():
    """H1(v) approximates -(2/sqrt(pi)) (1 - 2 v F(v)) (F = Dawson); H2 = (1 - 2v^2) e^{-v^2}."""
    for v in np.linspace(0, 4.9, 99):
        h0, h1, h2 = oracle.harris_series(v)
        assert h0 == pytest.approx(math.exp(-v * v), rel=1e-15)
        assert h2 == pytest.approx((1 - 2 * v * v) * math.exp(-v * v), rel=1e-14, abs=1e-300)
        exact = -2 / math.sqrt(math.pi) * (1 - 2 * v * dawsn(v))
        assert abs(h1 - exact) < 6e-3


# ----------------------------------------------------------------------------- closed forms
@pytest.mark.parametrize("rho_rel", [1e-6, 1e-3, 0.3, 0.99])
def test_inverse_densities_round_trip(rho_rel):
    sigma, gam = 2e-10, 3e-11
    peak_g = 1 / (math.sqrt(2 * math.pi) * sigma)
    x = oracle.inverse_gaussian_density(rho_rel * peak_g, sigma)
    assert math.exp(-0.5 * x * x / sigma ** 2) * peak_g == pytest.approx(rho_rel * peak_g, rel=1e-9)
    peak_l = 1 / (math.pi * gam)
    x = oracle.inverse_lorentz_density(rho_rel * peak_l, gam)
    assert 1 / (math.pi * gam * (1 + x * x / gam ** 2)) == pytest.approx(rho_rel * peak_l, rel=1e-9)
    assert oracle.inverse_gaussian_density(1.01 * peak_g, sigma) == 0.0
    assert oracle.inverse_lorentz_density(1.01 * peak_l, gam) == 0.0


def test_doppler_width_known_value():
    """Thermal Doppler width of Fe at 5000 A, 5000 K: the most probable speed sqrt(2kT/m) is
    1.2201 km/s, so lambda v/c = 0.020349 A, which is sigma * sqrt(2)."""
    m = 55.85 * amu_cgs
    s = oracle.doppler_width(5000e-8, 5000.0, m, 0.0)
    assert s * math.sqrt(2) * 1e8 == pytest.approx(0.020349, rel=1e-4)
    # microturbulence adds in quadrature to the most-probable speed: sqrt(2kT/m + xi^2)
    s2 = oracle.doppler_width(5000e-8, 5000.0, m, 1e5)
    v_mp = math.sqrt(2 * kboltz_cgs * 5000 / m)
    assert s2 * math.sqrt(2) == pytest.approx(5000e-8 / c_cgs * math.hypot(v_mp, 1e5), rel=1e-12)


def test_broadening_temperature_laws():
    assert oracle.scaled_stark(2.5e-5, 1e4) == 2.5e-5
    assert oracle.scaled_stark(1.0, 2e4) == pytest.approx(2 ** (1 / 6), rel=1e-15)
    assert oracle.scaled_vdW((3e-8, -1.0), 1.0, 1e4) == 3e-8
    assert oracle.scaled_vdW((1.0, -1.0), 1.0, 5e3) == pytest.approx(0.5 ** 0.3, rel=1e-15)


def test_abo_limits():
    """ABO (Anstee & O'Mara): alpha = 0 is the hard-sphere rate 2 sigma vbar; the temperature law
    is gamma ~ T^((1 - alpha)/2)."""
    m = 55.85 * amu_cgs
    sig = 1.3917417470792187e-14
    T = 6000.0
    vbar = math.sqrt(8 * kboltz_cgs * T / math.pi * (1 / (1.008 * amu_cgs) + 1 / m))
    assert oracle.scaled_vdW((sig, 0.0), m, T) == pytest.approx(2 * sig * vbar, rel=1e-13)
    alpha = 0.227
    r = oracle.scaled_vdW((sig, alpha), m, 2 * T) / oracle.scaled_vdW((sig, alpha), m, T)
    assert r == pytest.approx(2 ** ((1 - alpha) / 2), rel=1e-13)



@pytest.mark.parametrize("alpha", [0.227, 0.3, 0.45])
@pytest.mark.parametrize("T", [3500.0, 6000.0, 12000.0])
def test_abo_rate_is_maxwell_average(alpha, T):
    """The ABO broadening rate per perturber is twice the Maxwell-averaged rate <sigma(v) v> of
    the cross-section sigma(v) = sigma0 (v / v0)^-alpha (v0 = 1e6 cm/s) over the distribution of
    relative speeds of the radiator and a hydrogen atom (reduced mass mu): the average is done
    here by numerical quadrature of the physical definition, not by the closed form."""
    from scipy.integrate import quad
    m = 55.85 * amu_cgs
    sig = 1.3917417470792187e-14
    mu = 1.0 / (1.0 / (1.008 * amu_cgs) + 1.0 / m)
    s = math.sqrt(kboltz_cgs * T / mu)  # speed scale of the Maxwell distribution

    def integrand(x):  # x = v / s; f(v) dv = sqrt(2/pi) x^2 exp(-x^2/2) dx
        v = x * s
        return sig * (v / 1e6) ** (-alpha) * v * math.sqrt(2 / math.pi) * x * x * math.exp(-x * x / 2)

    rate, _ = quad(integrand, 0, 40, epsabs=0, epsrel=1e-13, limit=200)
    assert oracle.scaled_vdW((sig, alpha), m, T) == pytest.approx(2 * rate, rel=1e-10)


def test_continuum_interpolation_on_knots():
    """The gridded continuum is interpolated linearly between the knots that bracket a line
    centre: for a quadratic f(x) = x^2 that interpolant is exactly f(x) + (x - x_i)(x_{i+1} - x)
    on [x_i, x_{i+1}].  Passing those values at the line centres (KORG_CNTM_AT_LINES) must give
    the same windows as passing the knots (KORG_CNTM_GRID); the continuum only sets the cutoff."""
    from korg_line import workloads
    P, _, _ = workloads.solar_small(n_lines=400)
    knots = np.sort(np.concatenate([np.linspace(P.all_wls[0] - 3e-8, P.all_wls[-1] + 3e-8, 23),
                                    [5030.55e-8, 5030.9e-8]]))
    x = (knots - knots[0]) / (knots[-1] - knots[0])  # 0..1
    scale = np.geomspace(1e-9, 1e-6, P.n_layers)     # a different level per layer
    grid_vals = scale[:, None] * (1e-3 + x[None, :] ** 2)
    wl = P.lines.wl * (1e-8 if P.lines.wl[0] >= 1 else 1.0)
    i = np.clip(np.searchsorted(knots, wl, side="right") - 1, 0, len(knots) - 2)
    xl = (wl - knots[0]) / (knots[-1] - knots[0])
    exact = 1e-3 + xl ** 2 + (xl - x[i]) * (x[i + 1] - xl)
    at_lines = exact[:, None] * scale[None, :]       # (lines, layers)
    mk = lambda mode, vals, kn: Problem(P.all_wls, P.lines, P.species_mass, P.species_is_mol,
                                       P.n_div_U, P.temps, P.n_e, P.n_HI, P.n_HeI, P.n_H2, P.xi,
                                       mode, vals, kn, P.cutoff_threshold)
    a = np.zeros((P.n_layers, P.n_wl), order="F")
    lb1, ub1, _ = oracle.line_absorption(a, mk(1, grid_vals, knots), return_windows=True)
    lb2, ub2, _ = oracle.line_absorption(a, mk(0, at_lines, None), return_windows=True)
    assert np.array_equal(lb1, lb2) and np.array_equal(ub1, ub2)
    assert np.any(lb1 != oracle.line_absorption(a, mk(0, at_lines * 10.0, None),
                                                 return_windows=True)[0])


def test_sigma_line_classical_cross_section():
    """sigma_line * c / lambda^2 is pi e^2/(m_e c) = 0.026540 cm^2 s^-1."""
    lam = 5000e-8
    assert oracle.sigma_line(lam) * c_cgs / lam ** 2 == pytest.approx(0.0265400, rel=1e-5)


# ----------------------------------------------------------------------------- whole path
def _single_line_problem(line, wls, T=5000.0, n_div_U=1e10, ne=1e14, nH=1e16, xi=0.0,
                         cntm=1e-20, cutoff=3e-4):
    ll = LineList.from_lines([line])
    L = 1
    return Problem(wls.all_wls, ll, [ll.species[0].get_mass()], [ll.species[0].ismolecule()],
                   np.full((1, L), n_div_U), [T], [ne], [nH], [0.0], [0.0], [xi], 0,
                   np.full((1, L), cntm), None, cutoff)


def _run(P):
    a = np.zeros((P.n_layers, P.n_wl), order="F")
    lb, ub, nev = oracle.line_absorption(a, P, nthreads=1, return_windows=True)
    return a, lb, ub


def _fwhm(x, y):
    h = y.max() / 2
    above = np.nonzero(y >= h)[0]
    i0, i1 = above[0], above[-1]
    xl = np.interp(h, [y[i0 - 1], y[i0]], [x[i0 - 1], x[i0]])
    xr = np.interp(h, [y[i1 + 1], y[i1]], [x[i1 + 1], x[i1]])
    return xr - xl


def test_pure_doppler_line_shape_and_strength():
    """gamma = 0: Gaussian of standard deviation sigma = doppler_width; the line integrates to
    gf (pi e^2/m_e c)(lambda^2/c) n/U (e^{-E_l/kT} - e^{-E_u/kT})."""
    wls = Wavelengths((4999, 5001, 0.0005))
    T, nU = 5000.0, 1e10
    line = Line(5000.0, -1.0, "Fe I", 2.0, 0.0, 0.0, 0.0)
    a, lb, ub = _run(_single_line_problem(line, wls, T=T, n_div_U=nU, cntm=1e-30, cutoff=1e-12))
    y = a[0]
    sigma = math.sqrt(kboltz_cgs * T / (55.85 * amu_cgs)) * 5000e-8 / c_cgs
    assert _fwhm(wls.all_wls, y) == pytest.approx(2 * math.sqrt(2 * math.log(2)) * sigma, rel=2e-3)
    lam = 5000e-8
    E_u = 2.0 + hplanck_eV * c_cgs / lam
    strength = 10 ** -1.0 * 0.0265400 * lam ** 2 / c_cgs * nU * \
        (math.exp(-2.0 / (kboltz_eV * T)) - math.exp(-E_u / (kboltz_eV * T)))
    assert np.sum(y) * 0.0005e-8 == pytest.approx(strength, rel=1e-4)
    assert wls.all_wls[np.argmax(y)] == pytest.approx(lam, abs=0.0005e-8)


def test_pure_lorentz_line_width():
    """Natural broadening only, cold gas: FWHM in wavelength = 2 gamma, gamma = Gamma lambda^2/(4 pi c)."""
    wls = Wavelengths((4999.9, 5000.1, 0.00002))
    grad = 1e11
    line = Line(5000.0, 0.0, "Fe I", 0.0, grad, 0.0, 0.0)
    a, _, _ = _run(_single_line_problem(line, wls, T=1.0, cntm=1e-30, cutoff=1e-12))
    gam = grad * (5000e-8) ** 2 / (4 * math.pi * c_cgs)
    assert _fwhm(wls.all_wls, a[0]) == pytest.approx(2 * gam, rel=5e-3)


def test_window_is_where_line_falls_below_cutoff():
    """The window half-width w = sqrt(D^2 + L^2) satisfies D, L <= w, with the Gaussian and
    Lorentzian densities equal to cutoff * alpha_cntm / amplitude at D and L."""
    wls = Wavelengths((4990, 5010))
    line = Line(5000.0, -1.0, "Fe I", 1.0, 1e9, 0.0, 0.0)
    cntm, cutoff = 1e-8, 3e-4
    P = _single_line_problem(line, wls, T=5000.0, cntm=cntm, cutoff=cutoff)
    a, lb, ub = _run(P)
    lam = 5000e-8
    sigma = oracle.doppler_width(lam, 5000.0, 55.85 * amu_cgs, 0.0)
    gam = 1e9 * lam ** 2 / (4 * math.pi * c_cgs)
    amp = np.sum(a[0]) * 0.01e-8
    rho = cntm * cutoff / amp
    D = sigma * math.sqrt(-2 * math.log(math.sqrt(2 * math.pi) * sigma * rho))
    Lw = math.sqrt(gam / (math.pi * rho) - gam * gam)
    w = math.hypot(D, Lw)
    x = wls.all_wls
    assert lb[0] == np.searchsorted(x, lam - w * (1 - 1e-3)) + 1 or \
        lb[0] == np.searchsorted(x, lam - w * (1 + 1e-3)) + 1
    assert x[ub[0] - 1] <= lam + w * (1 + 1e-3) and x[lb[0] - 1] >= lam - w * (1 + 1e-3)


def test_schedules_agree():
    """The reference's per-chunk schedule and the deterministic one agree to rounding."""
    from korg_line import workloads
    P, _, _ = workloads.solar_small(n_lines=300)
    a = np.zeros((P.n_layers, P.n_wl), order="F")
    b = np.zeros((P.n_layers, P.n_wl), order="F")
    oracle.line_absorption(a, P, nthreads=4, mode=oracle.DETERMINISTIC)
    oracle.line_absorption(b, P, nthreads=4, mode=oracle.KORG_CHUNKS)
    np.testing.assert_allclose(a, b, rtol=1e-13, atol=1e-300)
    c = np.zeros((P.n_layers, P.n_wl), order="F")
    oracle.line_absorption(c, P, nthreads=1, mode=oracle.DETERMINISTIC)
    assert np.array_equal(a, c)


def test_accumulates_and_empty_list_noop():
    from korg_line import workloads
    P, _, _ = workloads.solar_small(n_lines=50)
    base = np.random.default_rng(0).random((P.n_layers, P.n_wl))
    a = np.asfortranarray(base.copy())
    z = np.zeros((P.n_layers, P.n_wl), order="F")
    oracle.line_absorption(a, P)
    oracle.line_absorption(z, P)
    assert np.array_equal(a, base + z)  # lines summed from zero, then added once
    E = Problem(P.all_wls, P.lines.take(np.zeros(0, dtype=int)), P.species_mass,
                P.species_is_mol, P.n_div_U, P.temps, P.n_e, P.n_HI, P.n_HeI, P.n_H2, P.xi,
                P.cntm_mode, P.cntm_values, P.cntm_knots, P.cutoff_threshold)
    b = np.asfortranarray(base.copy())
    oracle.line_absorption(b, E)
    assert np.array_equal(b, base)


def test_molecules_skip_stark_and_vdw():
    """src/line_absorption.jl:91-94: only gamma_rad broadens a molecular line."""
    wls = Wavelengths((4999, 5001, 0.001))
    mol = Line(5000.0, 0.0, Species("CO"), 0.5, 1e8, 1e-3, -7.0)
    bare = Line(5000.0, 0.0, Species("CO"), 0.5, 1e8, 0.0, 0.0)
    a1, _, _ = _run(_single_line_problem(mol, wls, ne=1e16, nH=1e17, cntm=1e-30, cutoff=1e-9))
    a2, _, _ = _run(_single_line_problem(bare, wls, ne=1e16, nH=1e17, cntm=1e-30, cutoff=1e-9))
    assert np.array_equal(a1, a2)
    assert Species("CO").get_mass() == atomic_masses[5] + atomic_masses[7]
