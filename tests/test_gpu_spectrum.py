"""The continuum on the grid (korg_continuum_batched) and the emergent flux
(korg_flux_batched): synthesize's alpha = continuum + lines and its flux and continuum flux
(src/synthesize.jl:225-226, 248-250, 286-290, 310-318).  Needs a GPU.

Tolerances: the continuum against the linear-interpolation formula at 1e-15 relative (Float64) and
within one Float32 ulp of its rounding; the flux against the Float64 oracle transfer (oracle/rt_oracle.c) of
the same total absorption at 1e-10; the normalized flux against the fit step's at 1e-13.
"""
import numpy as np
import pytest

from oracle import oracle
from korg_line import DeviceProblem, FitTarget, FluxTarget, workloads
from korg_line.atmosphere import Atmosphere

pytestmark = pytest.mark.gpu


def _star(seed, Teff, wl, n_lines=250):
    atm = Atmosphere.scaled(Teff, 4.3, -0.2)
    P = workloads.make_problem(atm, workloads.synthetic_linelist(
        seed, n_lines, wl[0] - 2, wl[1] + 2, (-4.0, 0.5), (0.0, 5.0)), workloads.Wavelengths(wl))
    tau_ref = 10.0 ** atm.lgTau5
    alpha_ref = atm.continuum(np.array([5000e-8]))[:, 0]
    return P, atm, (tau_ref, alpha_ref)


def _cntm_formula(P):
    """alpha_cntm on the grid: per point the knot interval (last knot <= lambda, clamped to the
    end intervals), then (1 - t) v_i + t v_(i+1)."""
    k, x = P.cntm_knots, P.all_wls
    i = np.clip(np.searchsorted(k, x, side="right") - 1, 0, k.size - 2)
    t = (x - k[i]) / (k[i + 1] - k[i])
    return (1.0 - t) * P.cntm_values[:, i] + t * P.cntm_values[:, i + 1]


STARS = ((1, 5200.0, (5000.0, 5010.0)), (2, 6100.0, (5003.0, 5006.0)), (3, 4500.0, (4990.0, 5002.0)))


def test_continuum_on_grid_matches_interpolation(ctx):
    stars = [_star(*s) for s in STARS]
    Ds = [DeviceProblem(P) for P, _, _ in stars]
    for dt, rtol in (("float64", 1e-15), ("float32", 0.0)):
        As = [d.new_alpha(dt) for d in Ds]
        for a in As:  # the fill overwrites whatever alpha holds
            a.copy_from(np.full(a.shape, 7.0, dtype=dt))
        ctx.continuum_batched(As, Ds)
        for (P, _, _), a in zip(stars, As):
            ref = _cntm_formula(P)
            got = a.numpy().T
            if dt == "float64":
                np.testing.assert_allclose(got, ref, rtol=rtol, atol=0)
            else:  # the Float32 rounding of the Float64 value (within one ulp)
                np.testing.assert_allclose(got, ref.astype(np.float32), rtol=6e-8, atol=0)


def test_flux_of_continuum_plus_lines_matches_oracle(ctx):
    stars = [_star(*s) for s in STARS]
    Ds = [DeviceProblem(P) for P, _, _ in stars]
    As = [d.new_alpha("float64") for d in Ds]
    ctx.continuum_batched(As, Ds)
    ctx.line_absorption_batched(As, Ds, "fp64")  # accumulates onto the continuum
    tg = [FluxTarget(P, *r) for P, _, r in stars]
    ctx.flux_batched(As, Ds, tg)
    for (P, atm, (tr, ar)), a, t in zip(stars, As, tg):
        F, Fc = t.flux(), t.cntm()
        cg = _cntm_formula(P)
        lines = np.zeros((P.n_layers, P.n_wl), order="F")
        oracle.line_absorption(lines, P)
        np.testing.assert_allclose(a.numpy().T, cg + lines, rtol=1e-12, atol=0)
        Fo = oracle.rt_flux(cg + lines, P.all_wls, P.temps, tr, ar) * 1e-8
        Fco = oracle.rt_flux(cg, P.all_wls, P.temps, tr, ar) * 1e-8
        np.testing.assert_allclose(F, Fo, rtol=1e-10)
        np.testing.assert_allclose(Fc, Fco, rtol=1e-10)
        assert np.all(F <= Fc * (1 + 1e-12))  # lines only absorb


def test_normalized_flux_equals_fit_step(ctx):
    stars = [_star(*s) for s in STARS[:2]]
    Ds = [DeviceProblem(P) for P, _, _ in stars]
    As = [d.new_alpha("float64") for d in Ds]
    ctx.continuum_batched(As, Ds)
    ctx.line_absorption_batched(As, Ds, "fp64")
    tg = [FluxTarget(P, *r) for P, _, r in stars]
    ctx.flux_batched(As, Ds, tg)
    Ds2 = [DeviceProblem(P) for P, _, _ in stars]
    As2 = [d.new_alpha("float64") for d in Ds2]
    Ts = [FitTarget(P, tr, ar, np.ones(P.n_wl), 0.01) for P, _, (tr, ar) in stars]
    ctx.fit_gradient_batched(As2, Ds2, Ts, "fp64")
    for f, t in zip(tg, Ts):
        np.testing.assert_allclose(f.flux() / f.cntm(), t.flux(), rtol=1e-13)


def test_continuum_requires_knots(ctx):
    """A star whose continuum is given at line centres (not on knots) is rejected."""
    from korg_line import _lib
    atm = Atmosphere.scaled(5200.0, 4.3, -0.2)
    P = workloads.make_problem(atm, workloads.synthetic_linelist(1, 50, 4998.0, 5012.0, (-4.0, 0.5),
                                                                 (0.0, 5.0)),
                               workloads.Wavelengths((5000.0, 5010.0)), cntm="at_lines")
    D = DeviceProblem(P)
    a = D.new_alpha("float64")
    with pytest.raises(_lib.KorgError):
        ctx.continuum_batched([a], [D])
