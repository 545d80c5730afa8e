"""Pins of the Float64 adjoint (oracle/line_oracle.c: oracle_voigt_hjerting_grad,
oracle_line_absorption_vjp) against central finite differences of the oracle's own forward.
CPU only.

The windows of the reference are integer functions of the inputs and are held fixed (as
ForwardDiff does through searchsorted); every check here uses steps small enough that no window
changes, and asserts that.
"""
import numpy as np
import pytest

from oracle import oracle
from korg_line import workloads
from korg_line.problem import Problem

GROUPS = ("T", "n_e", "xi", "n_HI", "n_HeI", "n_H2", "n_div_U")


@pytest.mark.parametrize("a,v", [
    (0.05, 7.0), (0.15, 12.0),             # branch A: a <= 0.2, v >= 5
    (0.0, 0.7), (0.1, 1.9), (0.18, 3.3),   # branch B: the three Harris pieces
    (0.5, 0.6), (1.0, 1.5), (1.3, 1.2),    # branch C: a <= 1.4, a + v < 3.2
    (0.3, 4.0), (2.0, 0.5), (5.0, 9.0),    # branch D
])
def test_voigt_gradient_matches_central_differences(a, v):
    h, da, dv = oracle.voigt_hjerting_grad(a, v)
    assert h == oracle.voigt_hjerting(a, v)
    ea, ev = 1e-6 * max(a, 1e-3), 1e-6 * v
    fa = (oracle.voigt_hjerting(a + ea, v) - oracle.voigt_hjerting(a - ea, v)) / (2 * ea)
    fv = (oracle.voigt_hjerting(a, v + ev) - oracle.voigt_hjerting(a, v - ev)) / (2 * ev)
    assert da == pytest.approx(fa, rel=1e-6, abs=1e-12 * abs(h))
    assert dv == pytest.approx(fv, rel=1e-6, abs=1e-12 * abs(h))


def _perturbed(P, name, d):
    """A copy of problem P with the per-layer input `name` moved by d (same shape)."""
    kw = dict(temps=P.temps, n_e=P.n_e, xi=P.xi, n_HI=P.n_HI, n_HeI=P.n_HeI, n_H2=P.n_H2,
              n_div_U=P.n_div_U)
    key = {"T": "temps"}.get(name, name)
    kw[key] = kw[key] + d
    return Problem(P.all_wls, P.lines, P.species_mass, P.species_is_mol, kw["n_div_U"],
                   kw["temps"], kw["n_e"], kw["n_HI"], kw["n_HeI"], kw["n_H2"], kw["xi"],
                   P.cntm_mode, P.cntm_values, P.cntm_knots, P.cutoff_threshold)


def _objective(P, G):
    a = np.zeros((P.n_layers, P.n_wl), order="F")
    lb, ub, _ = oracle.line_absorption(a, P, return_windows=True)
    return float(np.sum(G * a)), lb, ub


@pytest.fixture(scope="module")
def small_problem():
    P, _, _ = workloads.solar_small(n_lines=150)
    rng = np.random.default_rng(7)
    G = np.asfortranarray(rng.standard_normal((P.n_layers, P.n_wl)))
    return P, G


@pytest.mark.parametrize("name", GROUPS)
def test_vjp_matches_central_differences(small_problem, name):
    """<dL/dx, d> == (L(x + h d) - L(x - h d)) / 2h for L = <G, alpha>, random G and d."""
    P, G = small_problem
    grads = oracle.line_absorption_vjp(G, P)
    rng = np.random.default_rng(abs(hash(name)) % 2**32)
    base = {"T": P.temps, "n_e": P.n_e, "xi": P.xi, "n_HI": P.n_HI, "n_HeI": P.n_HeI,
            "n_H2": P.n_H2, "n_div_U": P.n_div_U}[name]
    d = rng.uniform(0.5, 1.5, base.shape) * np.abs(base)  # a relative direction per element
    g = grads[name]
    predicted = float(np.sum(g * d))
    # the largest step that moves no window; n_HeI and n_H2 enter only through the small share
    # they have of n_eff, so a larger relative step still moves n_eff by < 1e-6
    steps = (1e-3, 1e-4, 1e-5, 1e-6) if name in ("n_HeI", "n_H2") else (1e-6, 1e-7, 1e-8, 1e-9)
    for h in steps:
        fp, lbp, ubp = _objective(_perturbed(P, name, h * d), G)
        fm, lbm, ubm = _objective(_perturbed(P, name, -h * d), G)
        if np.array_equal(lbp, lbm) and np.array_equal(ubp, ubm):
            break
    else:
        pytest.fail("every step moved a window")
    fd = (fp - fm) / (2 * h)
    scale = float(np.sum(np.abs(g * d))) + 1e-300
    assert abs(predicted - fd) <= 1e-6 * scale, (predicted, fd, scale)


def test_vjp_is_linear_in_G_and_zero_for_zero_G(small_problem):
    P, G = small_problem
    g1 = oracle.line_absorption_vjp(G, P)
    g2 = oracle.line_absorption_vjp(4.0 * G, P)  # a power of two: exact in floating point
    for k in g1:
        assert np.array_equal(g2[k], 4.0 * g1[k]), k
    g0 = oracle.line_absorption_vjp(np.zeros_like(G), P)
    assert all(np.all(v == 0) for v in g0.values())


def test_vjp_n_div_U_is_the_line_opacity_per_unit_density(small_problem):
    """alpha is linear in n_div_U, so dL/d(n_div_U[s, l]) * n_div_U[s, l] summed over species
    equals <G, alpha> exactly (Euler's theorem for a degree-1 homogeneous function)."""
    P, G = small_problem
    g = oracle.line_absorption_vjp(G, P)
    L_val, _, _ = _objective(P, G)
    assert float(np.sum(g["n_div_U"] * P.n_div_U)) == pytest.approx(L_val, rel=1e-12)
