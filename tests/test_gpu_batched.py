"""Batched line opacity (korg_line_absorption_batched): one pipeline over many stars with uneven
line counts and wavelength ranges.  Each star's alpha must equal, bit for bit, what the single-star
entry point adds for that star alone (a star's lines keep their order in every tile, and tiles
never mix stars).  Batches of several layer counts (one pipeline per count) are held to the same
bar, with their windows, statistics, continuum and flux.  Needs a GPU.
"""
import numpy as np
import pytest

from oracle import oracle
from korg_line import DeviceProblem, Problem, Wavelengths, workloads
from korg_line.atmosphere import Atmosphere

pytestmark = pytest.mark.gpu


def _single(ctx, dprobs, precision, dtype):
    out = []
    for d in dprobs:
        a = d.new_alpha(dtype)
        ctx.line_absorption_device(a, d, precision)
        out.append(a.numpy())
    return out


@pytest.mark.parametrize("precision,dtype", [("fp32", "float32"), ("fp64", "float64"),
                                             ("fp32", "float64")])
def test_batched_equals_single_star_calls(ctx, precision, dtype):
    stars = workloads.batched_stars(n_stars=7, seed=5, scale=0.04)
    dprobs = [DeviceProblem(P) for P, _, _ in stars]
    ref = _single(ctx, dprobs, precision, dtype)
    alphas = [d.new_alpha(dtype) for d in dprobs]
    ctx.line_absorption_batched(alphas, dprobs, precision)
    for k, (a, r) in enumerate(zip(alphas, ref)):
        assert np.array_equal(a.numpy(), r), k
    assert ctx.stats()["n_evals"] > 0


def test_batched_accumulates_and_handles_edge_stars(ctx):
    """A star with no lines, a one-point grid and a star whose alpha already holds values."""
    stars = workloads.batched_stars(n_stars=3, seed=8, scale=0.03)
    P0 = stars[0][0]
    empty = Problem(P0.all_wls, P0.lines.take(np.arange(0)), P0.species_mass, P0.species_is_mol,
                    P0.n_div_U, P0.temps, P0.n_e, P0.n_HI, P0.n_HeI, P0.n_H2, P0.xi,
                    P0.cntm_mode, P0.cntm_values, P0.cntm_knots, P0.cutoff_threshold)
    atm = Atmosphere.from_model("sun")
    one_pt = workloads.make_problem(atm, workloads.synthetic_linelist(9, 50, 4990.0, 5010.0,
                                                                      (-4.0, 0.5), (0.0, 5.0)),
                                    Wavelengths([5000.0]))
    probs = [P for P, _, _ in stars] + [empty, one_pt]
    dprobs = [DeviceProblem(P) for P in probs]
    rng = np.random.default_rng(2)
    bases = [rng.uniform(0, 1e-6, (P.n_wl, P.n_layers)).astype(np.float32) for P in probs]
    ref = []
    for d, b in zip(dprobs, bases):
        a = d.new_alpha("float32")
        a.copy_from(b)
        ctx.line_absorption_device(a, d, "fp32")
        ref.append(a.numpy())
    alphas = []
    for d, b in zip(dprobs, bases):
        a = d.new_alpha("float32")
        a.copy_from(b)
        alphas.append(a)
    ctx.line_absorption_batched(alphas, dprobs, "fp32")
    for k, (a, r) in enumerate(zip(alphas, ref)):
        assert np.array_equal(a.numpy(), r), k
    assert np.array_equal(alphas[3].numpy(), bases[3])  # the empty star is untouched


def test_batched_fp64_matches_oracle(ctx):
    stars = workloads.batched_stars(n_stars=4, seed=11, scale=0.02)
    dprobs = [DeviceProblem(P) for P, _, _ in stars]
    alphas = [d.new_alpha("float64") for d in dprobs]
    ctx.line_absorption_batched(alphas, dprobs, "fp64")
    lb, ub = ctx.windows(sum(len(P.lines) for P, _, _ in stars))  # 1-based within each star
    off = 0
    for (P, _, _), a in zip(stars, alphas):
        ref = np.zeros((P.n_layers, P.n_wl), order="F")
        rlb, rub, _ = oracle.line_absorption(ref, P, return_windows=True)
        np.testing.assert_allclose(a.numpy().T, ref, rtol=1e-12, atol=1e-300)
        n = len(P.lines)
        used = rlb <= rub
        assert np.array_equal(used, lb[off:off + n] <= ub[off:off + n])
        assert np.array_equal(rlb[used], lb[off:off + n][used])
        assert np.array_equal(rub[used], ub[off:off + n][used])
        off += n


def test_batched_fp32_random_atmospheres_match_oracle(ctx):
    """Random atmospheres reach damping parameters of 1e6 and more in deep layers (star 5 here):
    every star's Float32 alpha is finite and within 1e-4 of the Float64 oracle."""
    from test_gpu_parity import FP32_RTOL, rel_err
    stars = workloads.batched_stars(n_stars=6, seed=21, scale=0.04)
    dprobs = [DeviceProblem(P) for P, _, _ in stars]
    alphas = [d.new_alpha("float32") for d in dprobs]
    ctx.line_absorption_batched(alphas, dprobs, "fp32")
    for k, ((P, _, _), a) in enumerate(zip(stars, alphas)):
        ref = np.zeros((P.n_layers, P.n_wl), order="F")
        oracle.line_absorption(ref, P)
        got = a.numpy().T
        assert np.all(np.isfinite(got)), k
        assert rel_err(got, ref) <= FP32_RTOL, (k, rel_err(got, ref))


def _layers(P, n):
    """P on its first n layers (Korg's interpolate_marcs drops layers: counts differ by star)."""
    c = np.ascontiguousarray
    return Problem(P.all_wls, P.lines, P.species_mass, P.species_is_mol, c(P.n_div_U[:, :n]),
                   c(P.temps[:n]), c(P.n_e[:n]), c(P.n_HI[:n]), c(P.n_HeI[:n]), c(P.n_H2[:n]),
                   c(P.xi[:n]), P.cntm_mode, c(P.cntm_values[:n]), P.cntm_knots,
                   P.cutoff_threshold)


@pytest.mark.parametrize("precision,dtype", [("fp32", "float32"), ("fp64", "float64")])
def test_batched_ragged_layer_counts(ctx, precision, dtype):
    """Stars of 56, 40, 56, 21 and 40 layers in one call: each star's alpha is what the single-star
    entry point adds for it (bitwise), the windows come back in the call's star order, and the
    statistics cover every star."""
    stars = workloads.batched_stars(n_stars=5, seed=12, scale=0.03)
    counts = [56, 40, 56, 21, 40]
    probs = [_layers(P, n) for (P, _, _), n in zip(stars, counts)]
    dprobs = [DeviceProblem(P) for P in probs]
    ref, ref_lb, ref_ub, n_evals = [], [], [], 0.0
    for d in dprobs:
        a = d.new_alpha(dtype)
        ctx.line_absorption_device(a, d, precision)
        ref.append(a.numpy())
        lb, ub = ctx.windows(len(d.problem.lines))
        ref_lb.append(lb)
        ref_ub.append(ub)
        n_evals += ctx.stats()["n_evals"]
    alphas = [d.new_alpha(dtype) for d in dprobs]
    ctx.line_absorption_batched(alphas, dprobs, precision)
    for k, (a, r) in enumerate(zip(alphas, ref)):
        assert a.shape == (probs[k].n_wl, counts[k])
        assert np.any(r) and np.array_equal(a.numpy(), r), k
    lb, ub = ctx.windows(sum(len(P.lines) for P in probs))
    assert np.array_equal(lb, np.concatenate(ref_lb)) and np.array_equal(ub, np.concatenate(ref_ub))
    assert ctx.stats()["n_evals"] == n_evals


def test_ragged_continuum_flux_and_plans(ctx):
    """Continuum and flux of a ragged batch equal the per-star calls; a plan refuses one."""
    from korg_line import FluxTarget
    from korg_line._lib import KorgError
    stars = workloads.batched_stars(n_stars=3, seed=13, scale=0.02)
    counts = [56, 30, 56]
    probs = [_layers(P, n) for (P, _, _), n in zip(stars, counts)]
    atms = [a for _, a, _ in stars]
    dprobs = [DeviceProblem(P) for P in probs]

    # per-layer references cut to each star's layers
    refs = [workloads.fit_reference(a) for a in atms]
    out = []
    for d, (tr, ar), n in zip(dprobs, refs, counts):
        a = d.new_alpha("float64")
        ctx.continuum_batched([a], [d])
        ctx.line_absorption_batched([a], [d], "fp64")
        t = FluxTarget(d.problem, tr[:n], ar[:n])
        ctx.flux_batched([a], [d], [t])
        out.append((a.numpy(), t.flux()))
    alphas = [d.new_alpha("float64") for d in dprobs]
    ctx.continuum_batched(alphas, dprobs)
    ctx.line_absorption_batched(alphas, dprobs, "fp64")
    tg = [FluxTarget(d.problem, tr[:n], ar[:n]) for d, (tr, ar), n in zip(dprobs, refs, counts)]
    ctx.flux_batched(alphas, dprobs, tg)
    for k in range(3):
        assert np.array_equal(alphas[k].numpy(), out[k][0]), k
        assert np.array_equal(tg[k].flux(), out[k][1]), k
    with pytest.raises(KorgError, match="layer count"):
        ctx.plan_batched([d.new_alpha("float32") for d in dprobs], dprobs, "fp32")
