"""The fit step on the device (korg_fit_gradient_batched, korg_plan_fit): line opacity, radiative
transfer, chi^2 and its gradient with respect to the per-layer inputs, against the Float64 oracle
chain -- oracle line opacity, oracle radiative transfer (rt_oracle.c), its reverse mode, and the
oracle line adjoint.  Needs a GPU.

Tolerances: the Float64 device step against the oracle chain at 1e-9 relative (normalized flux,
chi^2) and 1e-8 of the largest gradient of each quantity (the device recomputes the optical depth
walking up from the bottom layer, so the adjoint's operands differ from the oracle's in the last
bits); the batched and captured forms bitwise against the direct single-star step.
"""
import numpy as np
import pytest

from oracle import oracle
from korg_line import Context, DeviceProblem, FitTarget, Problem, workloads

pytestmark = pytest.mark.gpu

NAMES = ("T", "n_e", "xi", "n_HI", "n_HeI", "n_H2", "n_div_U")


def _cntm_on_grid(P):
    return np.array([np.interp(P.all_wls, P.cntm_knots, P.cntm_values[l])
                     for l in range(P.n_layers)])


def _reference(atm):
    tau_ref = 10.0 ** atm.lgTau5
    alpha_ref = atm.continuum(np.array([5000e-8]))[:, 0]
    return tau_ref, alpha_ref


def oracle_step(P, atm, data, sigma):
    """chi^2, normalized flux and the per-layer gradient by the oracle chain (Float64)."""
    tau_ref, alpha_ref = _reference(atm)
    al = np.zeros((P.n_layers, P.n_wl), order="F")
    oracle.line_absorption(al, P)
    ac = _cntm_on_grid(P)
    at = ac + al
    F = oracle.rt_flux(at, P.all_wls, P.temps, tau_ref, alpha_ref)
    Fc = oracle.rt_flux(ac, P.all_wls, P.temps, tau_ref, alpha_ref)
    Fn = F / Fc
    r = (Fn - data) / sigma
    gF = 2 * r / sigma / Fc
    G = oracle.rt_flux_vjp(at, P.all_wls, P.temps, tau_ref, alpha_ref, gF)
    grads = oracle.line_absorption_vjp(np.asfortranarray(G), P)
    return float(np.sum(r * r)), Fn, grads


def make_star(seed=1, n_lines=300, wl=(5000.0, 5015.0), Teff=None):
    if Teff is None:
        P, atm, wls = workloads.solar_small(n_lines=n_lines)
        P = workloads.make_problem(atm, workloads.synthetic_linelist(
            seed, n_lines, wl[0] - 2, wl[1] + 2, (-4.0, 0.5), (0.0, 5.0)), workloads.Wavelengths(wl))
    else:
        from korg_line.atmosphere import Atmosphere
        atm = Atmosphere.scaled(Teff, 4.2, -0.3)
        P = workloads.make_problem(atm, workloads.synthetic_linelist(
            seed, n_lines, wl[0] - 2, wl[1] + 2, (-4.0, 0.5), (0.0, 5.0)), workloads.Wavelengths(wl))
    rng = np.random.default_rng(seed)
    # data: the oracle's flux of slightly different lines plus noise, so residuals are non-zero
    sigma = 0.01
    _, Fn, _ = oracle_step(P, atm, np.zeros(P.n_wl), 1.0)
    data = Fn * (1 + 0.02 * rng.standard_normal(P.n_wl))
    return P, atm, data, sigma


def target(P, atm, data, sigma):
    tau_ref, alpha_ref = _reference(atm)
    return FitTarget(P, tau_ref, alpha_ref, data, sigma)


def test_fit_step_fp64_matches_oracle_chain(ctx):
    P, atm, data, sigma = make_star(seed=3)
    D = DeviceProblem(P)
    a = D.new_alpha("float64")
    t = target(P, atm, data, sigma)
    ctx.fit_gradient_batched([a], [D], [t], "fp64")
    chi2, Fn, grads = oracle_step(P, atm, data, sigma)
    np.testing.assert_allclose(t.flux(), Fn, rtol=1e-9)
    assert t.chi2() == pytest.approx(chi2, rel=1e-9)
    got = t.grads()
    for k in NAMES:
        ref = grads[k]
        scale = np.max(np.abs(ref)) if np.any(ref) else 1.0
        np.testing.assert_allclose(got[k].reshape(ref.shape), ref, rtol=1e-8, atol=1e-8 * scale,
                                   err_msg=k)
    # alpha holds the step's line opacity
    al = np.zeros((P.n_layers, P.n_wl), order="F")
    oracle.line_absorption(al, P)
    np.testing.assert_allclose(a.numpy().T, al, rtol=1e-12, atol=1e-300)


def test_fit_step_batched_equals_single_stars(ctx):
    stars = [make_star(seed=s, Teff=T) for s, T in ((5, 5200.0), (6, 6100.0), (7, 4600.0))]
    single = []
    for P, atm, data, sigma in stars:
        D = DeviceProblem(P)
        a = D.new_alpha("float32")
        t = target(P, atm, data, sigma)
        ctx.fit_gradient_batched([a], [D], [t], "fp32")
        single.append((t.chi2(), t.flux(), t.grads()))
    Ds = [DeviceProblem(P) for P, _, _, _ in stars]
    As = [d.new_alpha("float32") for d in Ds]
    Ts = [target(*s) for s in stars]
    ctx.fit_gradient_batched(As, Ds, Ts, "fp32")
    for k, (t, (c, f, g)) in enumerate(zip(Ts, single)):
        assert t.chi2() == c, k
        assert np.array_equal(t.flux(), f), k
        for q in NAMES:
            assert np.array_equal(t.grads()[q], g[q]), (k, q)


def test_fit_step_ragged_layer_counts(ctx):
    """Stars of 56, 40 and 56 layers in one fit step: chi^2, flux and gradients of each equal the
    single-star step's (bitwise)."""
    from test_gpu_batched import _layers
    stars = [make_star(seed=s, Teff=T) for s, T in ((5, 5200.0), (6, 6100.0), (7, 4600.0))]
    cut = []
    for (P, atm, data, sigma), n in zip(stars, (56, 40, 56)):
        tau_ref, alpha_ref = _reference(atm)
        cut.append((_layers(P, n), tau_ref[:n], alpha_ref[:n], data, sigma))
    single = []
    for Q, tr, ar, data, sigma in cut:
        D = DeviceProblem(Q)
        t = FitTarget(Q, tr, ar, data, sigma)
        ctx.fit_gradient_batched([D.new_alpha("float32")], [D], [t], "fp32")
        single.append((t.chi2(), t.flux(), t.grads()))
    Ds = [DeviceProblem(c[0]) for c in cut]
    Ts = [FitTarget(*c) for c in cut]
    ctx.fit_gradient_batched([d.new_alpha("float32") for d in Ds], Ds, Ts, "fp32")
    for k, (t, (c, f, g)) in enumerate(zip(Ts, single)):
        assert t.chi2() == c, k
        assert np.array_equal(t.flux(), f), k
        for q in NAMES:
            assert np.array_equal(t.grads()[q], g[q]), (k, q)


def test_fit_plan_equals_direct_and_reads_new_inputs(ctx):
    stars = [make_star(seed=s, Teff=T) for s, T in ((8, 5500.0), (9, 6300.0))]
    Ds = [DeviceProblem(P) for P, _, _, _ in stars]
    As = [d.new_alpha("float32") for d in Ds]
    Ts = [target(*s) for s in stars]
    ctx.fit_gradient_batched(As, Ds, Ts, "fp32")
    direct = [(t.chi2(), t.flux(), t.grads()) for t in Ts]
    Ds2 = [DeviceProblem(P) for P, _, _, _ in stars]
    As2 = [d.new_alpha("float32") for d in Ds2]
    Ts2 = [target(*s) for s in stars]
    with ctx.plan_fit(As2, Ds2, Ts2, "fp32") as plan:
        for rep in range(2):  # a replay recomputes everything: replays agree bitwise
            plan.launch()
            for k, (t, (c, f, g)) in enumerate(zip(Ts2, direct)):
                assert t.chi2() == c, (rep, k)
                assert np.array_equal(t.flux(), f), (rep, k)
                for q in NAMES:
                    assert np.array_equal(t.grads()[q], g[q]), (rep, k, q)
        # new inputs in the captured buffers
        new = []
        for (P, atm, data, sigma), d in zip(stars, Ds2):
            P2 = Problem(P.all_wls, P.lines, P.species_mass, P.species_is_mol, P.n_div_U * 1.3,
                         P.temps * 1.01, P.n_e, P.n_HI, P.n_HeI, P.n_H2, P.xi, P.cntm_mode,
                         P.cntm_values, P.cntm_knots, P.cutoff_threshold)
            d.update(P2)
            new.append((P2, atm, data, sigma))
        plan.launch()
        Ds3 = [DeviceProblem(s[0]) for s in new]
        As3 = [d.new_alpha("float32") for d in Ds3]
        Ts3 = [target(*s) for s in new]
        ctx.fit_gradient_batched(As3, Ds3, Ts3, "fp32")
        for t2, t3 in zip(Ts2, Ts3):
            assert t2.chi2() == t3.chi2()
            for q in NAMES:
                assert np.array_equal(t2.grads()[q], t3.grads()[q]), q


@pytest.mark.parametrize("residuals,tol", [("systematic", 1e-5), ("noise", 1e-4)])
def test_fit_step_fp32_forward_close_to_fp64(ctx, residuals, tol):
    """The Float32 step against the Float64 step: flux within 1e-4, chi^2 within 1e-6, and each
    quantity's per-layer gradient (2-norm) within ``tol``.

    With systematic residuals (data: the lines 10% deeper, plus noise -- an iterate away from the
    optimum) the gradient meets the 1e-5 of the north star.  With pure noise residuals (data: the
    model itself with 2% noise, the gradient at the optimum) the gradient is a sum of terms of random
    sign that cancel, and its relative error is set by the Float32 forward's flux error over sigma
    (about 2e-7 / 1e-2) -- whatever the adjoint's precision (a Float64 near-zone adjoint or tighter
    series radii leave it unchanged) -- so the bound stated for it is 1e-4."""
    P, atm, data, sigma = make_star(seed=10, n_lines=600)
    if residuals == "systematic":
        D = DeviceProblem(P)
        t = target(P, atm, data, sigma)
        ctx.fit_gradient_batched([D.new_alpha("float64")], [D], [t], "fp64")
        rng = np.random.default_rng(10)
        f = t.flux()
        data = 1 - 1.1 * (1 - f) + sigma * rng.standard_normal(f.size)
    res = {}
    for prec, dt in (("fp64", "float64"), ("fp32", "float32")):
        D = DeviceProblem(P)
        a = D.new_alpha(dt)
        t = target(P, atm, data, sigma)
        ctx.fit_gradient_batched([a], [D], [t], prec)
        res[prec] = (t.chi2(), t.flux(), t.grads())
    c64, f64, g64 = res["fp64"]
    c32, f32, g32 = res["fp32"]
    assert np.max(np.abs(f32 - f64)) <= 1e-4  # normalized flux
    assert c32 == pytest.approx(c64, rel=1e-6)
    for q in ("T", "n_e", "xi", "n_HI", "n_div_U"):
        err = np.linalg.norm(g32[q] - g64[q]) / np.linalg.norm(g64[q])
        print(f"fit step fp32 vs fp64 ({residuals} residuals): {q} {err:.2e}")
        assert err <= tol, (q, err)


def test_fit_step_gradient_matches_central_differences(ctx):
    """<dchi2/dx, d> against central differences of the device's Float64 chi^2 for inputs that
    reach chi^2 only through the line opacity (xi, n_e, n/U; T also moves the source function),
    with steps that move no window (the windows are held fixed, as ForwardDiff holds them)."""
    P, atm, data, sigma = make_star(seed=12, n_lines=200)
    tau_ref, alpha_ref = _reference(atm)

    def step(**kw):
        args = dict(xi=P.xi, n_e=P.n_e, n_div_U=P.n_div_U)
        args.update(kw)
        P2 = Problem(P.all_wls, P.lines, P.species_mass, P.species_is_mol, args["n_div_U"],
                     P.temps, args["n_e"], P.n_HI, P.n_HeI, P.n_H2, args["xi"], P.cntm_mode,
                     P.cntm_values, P.cntm_knots, P.cutoff_threshold)
        D = DeviceProblem(P2)
        a = D.new_alpha("float64")
        t = FitTarget(P2, tau_ref, alpha_ref, data, sigma)
        ctx.fit_gradient_batched([a], [D], [t], "fp64")
        return t.chi2(), t.grads(), ctx.windows(len(P.lines))

    _, g, _ = step()
    rng = np.random.default_rng(2)
    # steps: the largest that moves no window, below the ones that move v or a across a Hunger
    # branch or Harris piece boundary (H jumps slightly there; xi moves v directly).  n_e reaches
    # chi^2 only weakly (Stark broadening), so its small steps are rounding-limited.
    cases = (("xi", P.xi, (1e-6, 1e-7, 1e-8)), ("n_e", P.n_e, (1e-4, 1e-5, 1e-6, 1e-7)),
             ("n_div_U", P.n_div_U, (1e-6, 1e-7, 1e-8)))
    for name, base, steps in cases:
        gk = name
        d = rng.uniform(0.5, 1.5, base.shape) * base
        for h in steps:
            cp, _, (lbp, ubp) = step(**{name: base + h * d})
            cm, _, (lbm, ubm) = step(**{name: base - h * d})
            if np.array_equal(lbp, lbm) and np.array_equal(ubp, ubm):
                break
        else:
            pytest.fail("every step moved a window")
        fd = (cp - cm) / (2 * h)
        pred = float(np.sum(g[gk].reshape(d.shape) * d))
        assert pred == pytest.approx(fd, rel=1e-5, abs=1e-9 * abs(fd)), (name, pred, fd)


def _refine(P, tau_ref, alpha_ref, n):
    """P's atmosphere resampled to n layers (temperatures and microturbulence linear, densities,
    continuum and the reference depth scale geometric in the layer index)."""
    x = np.linspace(0.0, P.n_layers - 1.0, n)
    k = np.arange(P.n_layers)
    lin = lambda v: np.interp(x, k, v)
    geo = lambda v: np.exp(np.interp(x, k, np.log(np.maximum(v, 1e-300))))
    c = np.ascontiguousarray
    nd = c(np.array([geo(r) for r in P.n_div_U]))
    cv = c(np.array([geo(P.cntm_values[:, j]) for j in range(P.cntm_values.shape[1])]).T)
    Q = Problem(P.all_wls, P.lines, P.species_mass, P.species_is_mol, nd, c(lin(P.temps)),
                c(geo(P.n_e)), c(geo(P.n_HI)), c(geo(P.n_HeI)), c(geo(P.n_H2)), c(lin(P.xi)),
                P.cntm_mode, cv, P.cntm_knots, P.cutoff_threshold)
    return Q, geo(tau_ref), geo(alpha_ref)


@pytest.mark.parametrize("n_layers", [60, 64, 72])
def test_fit_step_fp32_gradient_layer_counts(ctx, n_layers):
    """The Float32 step's gradient against the Float64 step's at 60 and 64 layers (the near-zone
    adjoint with 8 layer groups of 8) and at 72 (two 64-layer blocks: lanes over layers), with
    systematic residuals: each quantity's per-layer gradient within 1e-5 (2-norm)."""
    P0, atm, data, sigma = make_star(seed=14, n_lines=400)
    tau_ref, alpha_ref = _reference(atm)
    P, tr, ar = _refine(P0, tau_ref, alpha_ref, n_layers)
    D = DeviceProblem(P)
    t = FitTarget(P, tr, ar, data, sigma)
    ctx.fit_gradient_batched([D.new_alpha("float64")], [D], [t], "fp64")
    f = t.flux()
    data = 1 - 1.1 * (1 - f) + sigma * np.random.default_rng(14).standard_normal(f.size)
    res = {}
    for prec, dt in (("fp64", "float64"), ("fp32", "float32")):
        D = DeviceProblem(P)
        t = FitTarget(P, tr, ar, data, sigma)
        ctx.fit_gradient_batched([D.new_alpha(dt)], [D], [t], prec)
        res[prec] = t.grads()
    for q in ("T", "n_e", "xi", "n_HI", "n_div_U"):
        err = np.linalg.norm(res["fp32"][q] - res["fp64"][q]) / np.linalg.norm(res["fp64"][q])
        print(f"{n_layers} layers: {q} {err:.2e}")
        assert err <= 1e-5, (n_layers, q, err)
