"""Forward mode (korg_line_absorption_jvp, _jvp_device) and frozen windows on the device, against
the oracle, through the C ABI.  Needs a GPU.

Tolerances: Float64 throughout, so the device equals the oracle to 1e-11 relative per element
(atol 1e-12 of the direction's largest element, since derivatives change sign); the adjoint
identity <G, J x> = <J^T G, x> between the device's forward and reverse modes to 1e-12 of
sum |J^T G * x|; frozen windows: index work bit-exact, Float64 alpha to 1e-12, Float32 to 1e-4.
"""
import numpy as np
import pytest

from oracle import oracle
from korg_line import Context, DeviceProblem, Problem, workloads
from korg_line.hipmem import DeviceArray

pytestmark = pytest.mark.gpu

GROUPS = ("T", "n_e", "xi", "n_HI", "n_HeI", "n_H2", "n_div_U")
BASE = {"T": "temps"}


def _tangent(P, seed, names=GROUPS):
    rng = np.random.default_rng(seed)
    return {n: rng.uniform(-1.0, 1.0, getattr(P, BASE.get(n, n)).shape) *
            np.abs(getattr(P, BASE.get(n, n))) for n in names}


def _smaller_cutoff_windows(P, cutoff=3e-5):
    Q = Problem(P.all_wls, P.lines, P.species_mass, P.species_is_mol, P.n_div_U, P.temps, P.n_e,
                P.n_HI, P.n_HeI, P.n_H2, P.xi, P.cntm_mode, P.cntm_values, P.cntm_knots, cutoff)
    lb, ub, _ = oracle.line_absorption(np.zeros((P.n_layers, P.n_wl), order="F"), Q,
                                       mode=oracle.WINDOWS_ONLY, return_windows=True)
    return lb, ub


def _close(got, ref, rtol=1e-11):
    np.testing.assert_allclose(got, ref, rtol=rtol, atol=1e-12 * np.max(np.abs(ref)))


@pytest.fixture(scope="module")
def prob():
    P, _, _ = workloads.solar_small(n_lines=400)
    return P


def test_frozen_windows_match_oracle(ctx, prob):
    P = prob
    lb, ub = _smaller_cutoff_windows(P)
    F = P.with_windows(lb, ub)
    ref = np.zeros((P.n_layers, P.n_wl), order="F")
    rlb, rub, rn = oracle.line_absorption(ref, F, return_windows=True)
    got = np.zeros_like(ref)
    ctx.line_absorption(got, F, "fp64")
    glb, gub = ctx.windows(len(P.lines))
    assert np.array_equal(glb, rlb) and np.array_equal(gub, rub)
    assert ctx.stats()["n_evals"] == rn
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-300)
    g32 = np.zeros_like(ref)
    ctx.line_absorption(g32, F, "fp32")
    m = np.abs(ref) >= 1e-8 * np.abs(ref).max()
    assert np.max(np.abs(g32[m] - ref[m]) / np.abs(ref[m])) <= 1e-4
    # freezing the computed windows changes nothing, in either precision and through the device
    # entry point
    a = np.zeros_like(ref)
    ctx.line_absorption(a, P, "fp64")
    clb, cub = ctx.windows(len(P.lines))
    D = DeviceProblem(P.with_windows(clb, cub))
    for prec in ("fp64", "fp32"):
        x = np.zeros_like(ref)
        ctx.line_absorption(x, P, prec)
        y = D.new_alpha("float64")
        ctx.line_absorption_device(y, D, prec)
        assert np.array_equal(y.numpy().T, x), prec


def test_jvp_matches_oracle(ctx, prob):
    P = prob
    xs = [_tangent(P, 10 + k) for k in range(3)] + [_tangent(P, 20, ("T",)),
                                                      _tangent(P, 21, ("n_div_U", "xi"))]
    ref = oracle.line_absorption_jvp(P, xs)
    got = ctx.line_absorption_jvp(P, xs)  # 5 directions: two passes of the profile kernel
    for k in range(len(xs)):
        assert np.abs(ref[k]).max() > 0
        _close(got[k], ref[k])
    assert ctx.stats()["n_evals"] > 0


def test_jvp_with_frozen_windows_matches_oracle(ctx, prob):
    P = prob
    lb, ub = _smaller_cutoff_windows(P)
    F = P.with_windows(lb, ub)
    xs = [_tangent(F, 30), _tangent(F, 31, ("n_e", "n_HI"))]
    ref = oracle.line_absorption_jvp(F, xs)
    got = ctx.line_absorption_jvp(F, xs)
    for k in range(2):
        _close(got[k], ref[k])


def test_jvp_adjoint_identity_with_device_vjp(ctx, prob):
    P = prob
    rng = np.random.default_rng(3)
    G = np.asfortranarray(rng.standard_normal((P.n_layers, P.n_wl)))
    g = ctx.line_absorption_vjp(G, P)
    xs = [_tangent(P, 40 + k) for k in range(2)]
    Js = ctx.line_absorption_jvp(P, xs)
    for x, J in zip(xs, Js):
        lhs = float(np.sum(G * J))
        rhs = sum(float(np.sum(g[n] * x[n])) for n in GROUPS)
        scale = sum(float(np.sum(np.abs(g[n] * x[n]))) for n in GROUPS)
        assert abs(lhs - rhs) <= 1e-12 * scale, (lhs, rhs, scale)


def test_jvp_device_entry_accumulates_in_both_dtypes(ctx, prob):
    P = prob
    D = DeviceProblem(P)
    xs = [_tangent(P, 50), _tangent(P, 51, ("T", "n_div_U"))]
    ref = oracle.line_absorption_jvp(P, xs)
    rng = np.random.default_rng(9)
    base = [rng.standard_normal((P.n_layers, P.n_wl)) * np.abs(r).max() for r in ref]
    td = [{k: DeviceArray.from_numpy(np.ascontiguousarray(v)) for k, v in x.items()} for x in xs]
    for dt, rtol in (("float64", 1e-11), ("float32", 1e-5)):
        outs = [D.new_alpha(dt, fill=b) for b in base]
        ctx.line_absorption_jvp_device(D, td, outs)
        for o, b, r in zip(outs, base, ref):
            got = o.numpy().T.astype(np.float64)
            want = b.astype(dt).astype(np.float64) + r
            np.testing.assert_allclose(got, want, rtol=rtol, atol=rtol * np.abs(want).max())


def test_jvp_empty_list_and_no_directions(ctx, prob):
    P = prob
    E = Problem(P.all_wls, P.lines.take(np.arange(0)), P.species_mass, P.species_is_mol,
                P.n_div_U, P.temps, P.n_e, P.n_HI, P.n_HeI, P.n_H2, P.xi, P.cntm_mode,
                P.cntm_values, P.cntm_knots, P.cutoff_threshold)
    d = [np.full((P.n_layers, P.n_wl), 3.0, order="F")]
    ctx.line_absorption_jvp(E, [_tangent(P, 1)], d)
    assert np.all(d[0] == 3.0)
    assert ctx.line_absorption_jvp(P, []) == []


def test_jvp_mirror_with_partition_functions(ctx):
    """line_absorption_jvp_ (the Dual path of the mirror: n/U moves with the densities and, through
    U(T), with T) against central differences of the frozen-window device forward, and against
    line_absorption_vjp_ by the adjoint identity."""
    from korg_line import Species, Wavelengths, line_absorption_jvp_, line_absorption_vjp_
    from korg_line.atmosphere import Atmosphere, continuum_grid
    from korg_line.problem import ContinuumGrid
    atm = Atmosphere.from_model("sun")
    wls = Wavelengths((5000, 5030))
    ll = workloads.synthetic_linelist(14, 150, 4995.0, 5035.0, (-4.0, 0.5), (0.0, 5.0))
    cg = ContinuumGrid(*continuum_grid(atm, float(wls.all_wls[0]), float(wls.all_wls[-1])))
    L, n = atm.n_layers, wls.all_wls.size
    xi = 1e5 * np.ones(L)
    rng = np.random.default_rng(15)
    fe = Species("Fe I")
    x = {"temps": rng.uniform(0.5, 1.5, L) * atm.temps, "n_e": rng.uniform(0.5, 1.5, L) * atm.n_e,
         "xi": rng.uniform(0.5, 1.5, L) * xi,
         "n_densities": {fe: rng.uniform(0.5, 1.5, L) * atm.n_densities[fe],
                         Species("H I"): rng.uniform(0.5, 1.5, L) * atm.n_densities[Species("H I")]}}
    J = [np.zeros((L, n), order="F")]
    line_absorption_jvp_(J, ll, wls, atm.temps, atm.n_e, atm.n_densities, atm.partition_fns, xi,
                         cg, [x], ctx=ctx)
    J = J[0]
    # central differences of the device forward with the windows of the unperturbed call
    P0 = Problem.from_korg_args(ll, wls, atm.temps, atm.n_e, atm.n_densities, atm.partition_fns,
                                xi, cg)
    a0 = np.zeros((L, n), order="F")
    ctx.line_absorption(a0, P0, "fp64")
    lb, ub = ctx.windows(len(ll))

    def forward(h):
        nd = dict(atm.n_densities)
        for sp, d in x["n_densities"].items():
            nd[sp] = atm.n_densities[sp] + h * d
        P = Problem.from_korg_args(ll, wls, atm.temps + h * x["temps"], atm.n_e + h * x["n_e"],
                                   nd, atm.partition_fns, xi + h * x["xi"], cg).with_windows(lb, ub)
        a = np.zeros((L, n), order="F")
        ctx.line_absorption(a, P, "fp64")
        return a
    h = 1e-6
    fd = (forward(h) - forward(-h)) / (2 * h)
    assert np.linalg.norm(J - fd) <= 1e-6 * np.linalg.norm(J), np.linalg.norm(J - fd) / np.linalg.norm(J)
    G = np.asfortranarray(rng.standard_normal((L, n)))
    g = line_absorption_vjp_(G, ll, wls, atm.temps, atm.n_e, atm.n_densities, atm.partition_fns,
                             xi, cg, ctx=ctx)
    rhs = (float(np.sum(g["temps"] * x["temps"])) + float(np.sum(g["n_e"] * x["n_e"])) +
           float(np.sum(g["xi"] * x["xi"])) +
           sum(float(np.sum(g["n_densities"][sp] * d)) for sp, d in x["n_densities"].items()))
    lhs = float(np.sum(G * J))
    assert abs(lhs - rhs) <= 1e-10 * abs(rhs), (lhs, rhs)
