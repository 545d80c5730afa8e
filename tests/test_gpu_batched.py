"""Batched line opacity (korg_line_absorption_batched): one pipeline over many stars with uneven
line counts and wavelength ranges.  Each star's alpha must equal, bit for bit, what the single-star
entry point adds for that star alone (a star's lines keep their order in every tile, and tiles
never mix stars).  Needs a GPU.
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


def test_batched_rejects_mismatched_layers(ctx):
    stars = workloads.batched_stars(n_stars=2, seed=12, scale=0.02)
    P1 = stars[1][0]
    short = Problem(P1.all_wls, P1.lines, P1.species_mass, P1.species_is_mol,
                    P1.n_div_U[:, :10], P1.temps[:10], P1.n_e[:10], P1.n_HI[:10],
                    P1.n_HeI[:10], P1.n_H2[:10], P1.xi[:10], P1.cntm_mode,
                    P1.cntm_values[:10], P1.cntm_knots, P1.cutoff_threshold)
    d0, d1 = DeviceProblem(stars[0][0]), DeviceProblem(short)
    from korg_line._lib import KorgError
    with pytest.raises(KorgError, match="layers"):
        ctx.line_absorption_batched([d0.new_alpha("float32"), d1.new_alpha("float32")],
                                    [d0, d1], "fp32")
