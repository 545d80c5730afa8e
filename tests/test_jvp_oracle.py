"""Pins of the oracle's frozen-window mode and of its Jacobian-vector product
(oracle_line_absorption_jvp, the forward-mode derivative a ForwardDiff.Dual alpha receives).
CPU only.

Frozen windows: freezing the computed windows changes nothing (bitwise); freezing the windows of
a smaller cutoff gives that smaller cutoff's alpha exactly (the amplitudes do not depend on the
cutoff, only the windows do).

JVP: against central differences of the frozen-window forward, per input group; exactly linear in
n/U (alpha is linear in n/U, so the JVP along d is the forward with n/U := d); and the adjoint
identity <G, J x> = <J^T G, x> with the oracle's reverse mode, to 1e-12.
"""
import numpy as np
import pytest

from oracle import oracle
from korg_line import workloads
from korg_line.problem import Problem

GROUPS = ("T", "n_e", "xi", "n_HI", "n_HeI", "n_H2", "n_div_U")
BASE = {"T": "temps"}


def _forward(P):
    a = np.zeros((P.n_layers, P.n_wl), order="F")
    lb, ub, _ = oracle.line_absorption(a, P, return_windows=True)
    return a, lb, ub


def _with(P, **kw):
    """A copy of P with some per-layer inputs replaced (frozen windows kept)."""
    args = dict(temps=P.temps, n_e=P.n_e, xi=P.xi, n_HI=P.n_HI, n_HeI=P.n_HeI, n_H2=P.n_H2,
                n_div_U=P.n_div_U)
    args.update(kw)
    Q = Problem(P.all_wls, P.lines, P.species_mass, P.species_is_mol, args["n_div_U"],
                args["temps"], args["n_e"], args["n_HI"], args["n_HeI"], args["n_H2"],
                args["xi"], P.cntm_mode, P.cntm_values, P.cntm_knots, P.cutoff_threshold)
    if P.window_lb is not None:
        Q = Q.with_windows(P.window_lb, P.window_ub)
    return Q


@pytest.fixture(scope="module")
def problem():
    P, _, _ = workloads.solar_small(n_lines=120)
    _, lb, ub = _forward(P)
    return P, lb, ub


def test_frozen_computed_windows_change_nothing(problem):
    P, lb, ub = problem
    a, _, _ = _forward(P)
    b, lb2, ub2 = _forward(P.with_windows(lb, ub))
    assert np.array_equal(a, b)
    assert np.array_equal(lb, lb2) and np.array_equal(ub, ub2)


def test_frozen_windows_of_a_smaller_cutoff_give_its_alpha(problem):
    P, _, _ = problem
    Q = Problem(P.all_wls, P.lines, P.species_mass, P.species_is_mol, P.n_div_U, P.temps, P.n_e,
                P.n_HI, P.n_HeI, P.n_H2, P.xi, P.cntm_mode, P.cntm_values, P.cntm_knots, 3e-5)
    aq, lbq, ubq = _forward(Q)
    wide = (ubq - lbq) > 0
    assert np.sum(wide) > 50
    ap, lbp, ubp = _forward(P.with_windows(lbq, ubq))
    assert np.array_equal(lbp, lbq) and np.array_equal(ubp, ubq)
    assert np.array_equal(ap, aq)
    # and a frozen window of one point adds exactly one column per line
    one = np.where(lbq <= ubq, (lbq + ubq) // 2, 1)
    a1, lb1, ub1 = _forward(P.with_windows(one, np.where(lbq <= ubq, one, 0)))
    cols = np.unique(one[lbq <= ubq] - 1)
    assert np.all(a1[:, np.setdiff1d(np.arange(P.n_wl), cols)] == 0.0)
    assert np.all(a1[:, cols] > 0.0)


def _direction(P, name, seed):
    rng = np.random.default_rng(seed)
    base = getattr(P, BASE.get(name, name))
    return rng.uniform(0.5, 1.5, base.shape) * np.abs(base)


@pytest.mark.parametrize("name", GROUPS)
def test_jvp_matches_central_differences(problem, name):
    P0, lb, ub = problem
    P = P0.with_windows(lb, ub)  # the windows a ForwardDiff evaluation holds fixed
    d = _direction(P, name, 100 + GROUPS.index(name))
    J = oracle.line_absorption_jvp(P, [{name: d}])[0]
    base = getattr(P, BASE.get(name, name))
    key = BASE.get(name, name)
    # n_HeI and n_H2 enter only through their small share of n_eff: larger steps; n_e only
    # through Stark broadening
    steps = {"n_HeI": (1e-4, 1e-5), "n_H2": (1e-4, 1e-5), "n_e": (1e-5, 1e-6)}.get(name, (1e-6, 1e-7))
    best = np.inf
    for h in steps:
        ap, _, _ = _forward(_with(P, **{key: base + h * d}))
        am, _, _ = _forward(_with(P, **{key: base - h * d}))
        fd = (ap - am) / (2 * h)
        best = min(best, float(np.linalg.norm(J - fd) / np.linalg.norm(J)))
    assert np.linalg.norm(J) > 0
    assert best <= 1e-6, best


def test_jvp_is_exactly_linear_in_n_div_U(problem):
    P0, lb, ub = problem
    P = P0.with_windows(lb, ub)
    d = _direction(P, "n_div_U", 7)
    J = oracle.line_absorption_jvp(P, [{"n_div_U": d}])[0]
    a, _, _ = _forward(_with(P, n_div_U=d))
    np.testing.assert_allclose(J, a, rtol=1e-13, atol=1e-300)


def test_jvp_adjoint_identity_and_directions(problem):
    """<G, J x> == <J^T G, x> for a tangent moving every input at once; several directions in one
    call equal one call each."""
    P, _, _ = problem
    rng = np.random.default_rng(5)
    G = np.asfortranarray(rng.standard_normal((P.n_layers, P.n_wl)))
    xs = [{n: _direction(P, n, 10 * k + i) for i, n in enumerate(GROUPS)} for k in range(3)]
    Js = oracle.line_absorption_jvp(P, xs)
    g = oracle.line_absorption_vjp(G, P)
    for k, x in enumerate(xs):
        lhs = float(np.sum(G * Js[k]))
        rhs = sum(float(np.sum(g[n] * x[n])) for n in GROUPS)
        scale = sum(float(np.sum(np.abs(g[n] * x[n]))) for n in GROUPS)
        assert abs(lhs - rhs) <= 1e-12 * scale, (lhs, rhs)
        one = oracle.line_absorption_jvp(P, [x])[0]
        assert np.array_equal(one, Js[k])
