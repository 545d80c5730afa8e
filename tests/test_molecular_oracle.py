"""Pins of the molecular cross-section oracle (oracle/molecular_oracle.py) and of the host side of
the precompute (korg_line.molecular).  CPU only.

The interpolation: against SciPy's RegularGridInterpolator (linear, fill 0 outside), a textbook
routine independent of the restatement; exact at the knots; zero outside the table.
The precompute's line_absorption! calls carry the reference's settings
(src/molecular_cross_sections.jl:56-77): n_e = 0, the molecule's density 1/cutoff_alpha over its
partition function, NaN perturbers, xi = vmic * 1e5, continuum 1, cutoff 1.0.
"""
import numpy as np
import pytest
from scipy.interpolate import RegularGridInterpolator

from oracle import molecular_oracle as mo
from korg_line import Line, LineList, Species, Wavelengths
from korg_line.molecular import DEFAULT_LOG_T, DEFAULT_VMIC, MolecularCrossSection


def _table(seed=1):
    rng = np.random.default_rng(seed)
    wls = Wavelengths((15020, 15021))
    vm = np.array([0.0, 0.5, 1.5, 3.0])
    lt = np.array([3.2, 3.5, 3.6, 4.1])
    sigma = rng.uniform(1e-24, 1e-20, (vm.size, wls.all_wls.size, lt.size))
    return MolecularCrossSection(wls, vm, lt, sigma, Species("H2O"))


def test_interpolation_matches_scipy_and_knots():
    t = _table()
    itp = RegularGridInterpolator((t.vmic_vals, t.wls.all_wls, t.log_temp_vals), t.sigma,
                                  method="linear", bounds_error=False, fill_value=0.0)
    rng = np.random.default_rng(2)
    wl = np.concatenate([rng.uniform(15019.9e-8, 15021.1e-8, 200), t.wls.all_wls[[0, 5, -1]]])
    for vm, lt in [(0.7, 3.55), (0.0, 3.2), (3.0, 4.1), (2.2, 3.9), (3.1, 3.5), (1.0, 4.2)]:
        got = mo.cross_section(t, vm, lt, wl)
        want = itp(np.stack([np.full(wl.size, vm), wl, np.full(wl.size, lt)], axis=1))
        np.testing.assert_allclose(got, want, rtol=1e-13, atol=0)
    # at the knots: the table itself
    np.testing.assert_array_equal(mo.cross_section(t, 1.5, 3.6, t.wls.all_wls), t.sigma[2, :, 2])
    # outside any dimension: zero
    assert not np.any(mo.cross_section(t, 3.01, 3.6, t.wls.all_wls))
    assert not np.any(mo.cross_section(t, 1.0, 3.19, t.wls.all_wls))


def test_interpolation_adds_density_weighted_cross_sections():
    t = _table()
    temps = np.array([10 ** 3.3, 10 ** 3.95, 10 ** 4.5])
    n = {Species("H2O"): np.array([1e10, 2e12, 3e14])}
    a = np.full((3, 40), 7.0, order="F")
    wl = t.wls.all_wls[10:50]
    mo.interpolate_molecular_cross_sections(a, [t], wl, temps, 1.2, n)
    for l in range(3):
        want = 7.0 + mo.cross_section(t, 1.2, np.log10(temps[l]), wl) * n[Species("H2O")][l]
        np.testing.assert_allclose(a[l], want, rtol=1e-15)
    assert np.all(a[2] == 7.0)  # log T 4.5 is beyond the table


def test_precompute_problems_carry_the_reference_settings():
    lines = [Line(15020.5 + 0.1 * k, -3.0 + 0.1 * k, "H2O", 0.5) for k in range(5)]
    ll = LineList.from_lines(lines)
    U = lambda lnT: 1.0 + np.exp(lnT) / 100.0
    probs, sp, wls, vm, lt = MolecularCrossSection.precompute_problems(
        ll, Wavelengths((15020, 15022)), U, cutoff_alpha=1e-30)
    assert sp == Species("H2O") and len(probs) == DEFAULT_VMIC.size
    np.testing.assert_array_equal(lt, DEFAULT_LOG_T)
    assert lt[0] == 3.0 and abs(lt[-1] - 5.0) < 1e-12 and lt.size == 51
    for P, v in zip(probs, vm):
        T = 10.0 ** DEFAULT_LOG_T
        np.testing.assert_array_equal(P.temps, T)
        assert not P.n_e.any() and np.all(np.isnan(P.n_HI)) and np.all(np.isnan(P.n_H2))
        np.testing.assert_array_equal(P.xi, np.full(T.size, v * 1e5))
        np.testing.assert_allclose(P.n_div_U[0], 1e30 / U(np.log(T)), rtol=1e-15)
        assert P.cutoff_threshold == 1.0 and np.all(P.cntm_values == 1.0)
        assert P.species_is_mol[0] == 1
    with pytest.raises(ValueError):
        mixed = LineList.from_lines(lines + [Line(15021.5, -3.0, "CO", 0.5)])
        MolecularCrossSection.precompute_problems(mixed, Wavelengths((15020, 15022)), U)
