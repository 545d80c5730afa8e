"""The three benchmark workloads at their full sizes, through the C ABI.  Needs a GPU; each test
takes up to a couple of minutes (the oracle runs on all host threads).

* solar_wide (300,000 lines, 4000-7000 A at 0.01 A, 56 layers): the whole Float64 alpha equals
  the oracle's to 1e-12 per element, the Float32 form is within 1e-4 of it, and the normalized
  flux of continuum + lines (korg_continuum_batched, korg_flux_batched: Korg's planar transfer)
  is within 1e-4 of the flux of continuum + the oracle's lines.
* batched (256 stars, 19.2 M lines, 13.6 M points): every star's windows equal the oracle's
  window pass; one Float32 plan replay under the shipped tuning is within 1e-4 of the direct
  Float64 batched call per star; the Float64 alpha of every 32nd star equals the oracle's to 1e-12.
* fit (64 stars, 100,000 shared lines, 100,001 points): one Float32 plan replay of the fit step
  against the all-Float64 step: chi^2 within 1e-6 and every star's gradient of each per-layer
  quantity within 1e-5 (2-norm over layers); the Float64 step of two stars against the oracle
  chain (line opacity, transfer, its reverse mode, line adjoint).
"""
import os

import numpy as np
import pytest

from oracle import oracle
from korg_line import Context, DeviceProblem, FitTarget, FluxTarget, workloads
from korg_line.hipmem import DeviceArray

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(900)]

TUNING = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "korg.jl_amd",
                      "korg_line", "tuning_mi355x.txt")


def rel_err(a, ref, floor_frac=1e-8):
    ref = np.asarray(ref, dtype=np.float64)
    a = np.asarray(a, dtype=np.float64)
    m = np.abs(ref) >= floor_frac * np.abs(ref).max()
    return float(np.max(np.abs(a[m] - ref[m]) / np.abs(ref[m]))) if m.any() else 0.0


def log(msg):
    print(msg, flush=True)


# ----------------------------------------------------------------------------- solar_wide
def test_solar_wide_full_alpha_and_flux_match_oracle():
    P, atm, _ = workloads.solar_wide()
    L, n = P.n_layers, P.n_wl
    ref = np.zeros((L, n), order="F")
    oracle.line_absorption(ref, P)  # all host threads, deterministic schedule
    with Context(0) as ctx:
        D = DeviceProblem(P)
        a64 = D.new_alpha("float64")
        ctx.line_absorption_device(a64, D, "fp64")
        got = a64.numpy().T
        log(f"solar_wide fp64: max rel err {rel_err(got, ref, 0.0):.2e}")
        np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-300)
        del got
        ctx.autotune_batched([D.new_alpha("float32")], [D], "fp32", cache_path=TUNING)
        a32 = D.new_alpha("float32")
        ctx.line_absorption_device(a32, D, "fp32")
        err = rel_err(a32.numpy().T, ref)
        log(f"solar_wide fp32: max rel err {err:.2e}")
        assert err <= 1e-4
        # normalized flux: continuum on the grid + lines, Korg's planar transfer on the device
        tau, aref = workloads.fit_reference(atm)
        c = D.new_alpha("float64")
        ctx.continuum_batched([c], [D])
        tot = DeviceArray.from_numpy((c.numpy().T + ref).T)
        t32 = D.new_alpha("float32")
        ctx.continuum_batched([t32], [D])
        ctx.line_absorption_device(t32, D, "fp32")
        fr, f3 = FluxTarget(P, tau, aref), FluxTarget(P, tau, aref)
        ctx.flux_batched([tot], [D], [fr])
        ctx.flux_batched([t32], [D], [f3])
        Fr, F3 = fr.flux() / fr.cntm(), f3.flux() / f3.cntm()
        ferr = float(np.max(np.abs(F3 - Fr)))
        log(f"solar_wide normalized flux: max abs err {ferr:.2e} (depth {1 - Fr.min():.3f})")
        assert ferr <= 1e-4


# ----------------------------------------------------------------------------- batched
@pytest.fixture(scope="module")
def batched():
    stars = workloads.batched_stars(n_stars=256, seed=3)
    probs = [P for P, _, _ in stars]
    ctx = Context(0)
    dprobs = [DeviceProblem(P) for P in probs]
    a64 = [d.new_alpha("float64") for d in dprobs]
    ctx.line_absorption_batched(a64, dprobs, "fp64")
    counts = [len(P.lines) for P in probs]
    lb, ub = ctx.windows(int(sum(counts)))
    off = np.concatenate([[0], np.cumsum(counts)])
    wins = [(lb[off[k]:off[k + 1]], ub[off[k]:off[k + 1]]) for k in range(len(probs))]
    yield {"probs": probs, "ctx": ctx, "dprobs": dprobs, "a64": a64, "wins": wins,
           "n_evals": ctx.stats()["n_evals"]}
    ctx.close()


def test_batched_256_windows_match_oracle(batched):
    total = 0.0
    for k, P in enumerate(batched["probs"]):
        lb, ub, nev = oracle.line_absorption(np.zeros((P.n_layers, P.n_wl), order="F"), P,
                                             mode=oracle.WINDOWS_ONLY, return_windows=True)
        glb, gub = batched["wins"][k]
        skip = lb > ub
        assert np.array_equal(skip, glb > gub), k
        assert np.array_equal(lb[~skip], glb[~skip]) and np.array_equal(ub[~skip], gub[~skip]), k
        total += nev
    assert batched["n_evals"] == total


def test_batched_256_fp32_plan_matches_fp64(batched):
    ctx, dprobs = batched["ctx"], batched["dprobs"]
    a32 = [d.new_alpha("float32") for d in dprobs]
    ctx.autotune_batched(a32, dprobs, "fp32", cache_path=TUNING)
    with ctx.plan_batched(a32, dprobs, "fp32") as plan:
        plan.launch()
    worst = 0.0
    for k, (x, y) in enumerate(zip(a32, batched["a64"])):
        err = rel_err(x.numpy().T, y.numpy().T)
        worst = max(worst, err)
        assert err <= 1e-4, (k, err)
    log(f"batched 256: fp32 plan vs fp64, worst star {worst:.2e}")
    ctx.set_tuning()


def test_batched_256_fp64_matches_oracle_every_32nd_star(batched):
    for k in range(0, len(batched["probs"]), 32):
        P = batched["probs"][k]
        ref = np.zeros((P.n_layers, P.n_wl), order="F")
        oracle.line_absorption(ref, P)
        got = batched["a64"][k].numpy().T
        np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-300, err_msg=f"star {k}")


# ----------------------------------------------------------------------------- fit step
QUANTITIES = ("T", "n_e", "xi", "n_HI", "n_div_U")


@pytest.fixture(scope="module")
def fit():
    probs, atms, _, _ = workloads.fit_stars(64)
    refs = [workloads.fit_reference(a) for a in atms]
    ctx = Context(0)
    dprobs = [DeviceProblem(P) for P in probs]
    a32 = [d.new_alpha("float32") for d in dprobs]
    ctx.autotune_batched(a32, dprobs, "fp32", cache_path=TUNING)
    targets = [FitTarget(P, tr, ar, np.ones(P.n_wl), 0.01) for P, (tr, ar) in zip(probs, refs)]
    # data: the model flux with lines 10% deeper, plus noise (an iterate away from the optimum)
    ctx.fit_gradient_batched(a32, dprobs, targets, "fp32")
    rng = np.random.default_rng(7)
    for tg in targets:
        f = tg.flux()
        tg.set_data(1 - 1.1 * (1 - f) + 0.01 * rng.standard_normal(f.size))
    with ctx.plan_fit(a32, dprobs, targets, "fp32") as plan:
        plan.launch()
        plan.launch()  # a replay recomputes everything (alpha is zeroed first)
    g32 = [tg.grads() for tg in targets]
    c32 = np.array([tg.chi2() for tg in targets])
    ctx.set_tuning()
    a64 = [d.new_alpha("float64") for d in dprobs]
    t64 = [FitTarget(P, tr, ar, tg.t["data"].numpy(), 0.01)
           for P, (tr, ar), tg in zip(probs, refs, targets)]
    ctx.fit_gradient_batched(a64, dprobs, t64, "fp64")
    yield {"probs": probs, "atms": atms, "g32": g32, "c32": c32, "t64": t64}
    ctx.close()


def test_fit_64_fp32_step_matches_fp64_step(fit):
    c64 = np.array([t.chi2() for t in fit["t64"]])
    cerr = float(np.max(np.abs(fit["c32"] - c64) / np.abs(c64)))
    worst = {}
    for k, (ga, t) in enumerate(zip(fit["g32"], fit["t64"])):
        gb = t.grads()
        for q in QUANTITIES:
            nb = np.linalg.norm(gb[q])
            if nb > 0:
                worst[q] = max(worst.get(q, 0.0), float(np.linalg.norm(ga[q] - gb[q]) / nb))
    log(f"fit 64: chi2 {cerr:.2e}; gradients " + ", ".join(f"{q} {v:.2e}" for q, v in worst.items()))
    assert cerr <= 1e-6
    for q, v in worst.items():
        assert v <= 1e-5, (q, v)


def test_fit_64_fp64_step_matches_oracle_chain(fit):
    from test_gpu_fit_step import NAMES, oracle_step
    for k in (0, 63):
        P, atm, t = fit["probs"][k], fit["atms"][k], fit["t64"][k]
        chi2, Fn, grads = oracle_step(P, atm, t.t["data"].numpy(), 0.01)
        assert t.chi2() == pytest.approx(chi2, rel=1e-9)
        np.testing.assert_allclose(t.flux(), Fn, rtol=1e-9, atol=0)
        g = t.grads()
        for q in NAMES:
            scale = np.max(np.abs(grads[q]))
            if scale > 0:
                assert np.max(np.abs(g[q] - grads[q])) <= 1e-8 * scale, (k, q)
