"""Plans (batched calls captured in HIP graphs: korg_plan_batched / korg_plan_launch), the tuning
variants of the far-wing kernel, and the autotuner.  Needs a GPU.

A replay must add exactly what the direct batched call adds for the same buffers' contents
(bitwise): the graph holds the same kernels in the same order.  Replays are checked one by one,
with direct calls and host synchronisation between them.  The tile-shape variants change the
summation order of the far wing only, so each is checked against the Float64 form at the
Float32 tolerance of the parity tests (1e-4 relative).  The near zone's two lane layouts add the
same terms in the same order, so they are checked for identical alpha.
"""
import numpy as np
import pytest

from korg_line import Context, DeviceProblem, Problem, workloads
from test_gpu_parity import FP32_RTOL, rel_err

pytestmark = pytest.mark.gpu


def _batch(n_stars=6, seed=21, scale=0.04):
    return [P for P, _, _ in workloads.batched_stars(n_stars=n_stars, seed=seed, scale=scale)]


def _eager(ctx, probs, precision, dtype):
    dprobs = [DeviceProblem(P) for P in probs]
    alphas = [d.new_alpha(dtype) for d in dprobs]
    ctx.line_absorption_batched(alphas, dprobs, precision)
    return [a.numpy() for a in alphas]


def _layers(P, n):
    """P on n layers resampled from its own (repeats when n exceeds them)."""
    idx = np.round(np.linspace(0, P.n_layers - 1, n)).astype(int)
    c = np.ascontiguousarray
    return Problem(P.all_wls, P.lines, P.species_mass, P.species_is_mol, c(P.n_div_U[:, idx]),
                   c(P.temps[idx]), c(P.n_e[idx]), c(P.n_HI[idx]), c(P.n_HeI[idx]), c(P.n_H2[idx]),
                   c(P.xi[idx]), P.cntm_mode, c(P.cntm_values[idx]), P.cntm_knots,
                   P.cutoff_threshold)


def _zero(alphas):
    for a in alphas:
        a.copy_from(np.zeros(a.shape, a.dtype))


def _scaled(P, temps=1.0, n_e=1.0, n_div_U=1.0):
    return Problem(P.all_wls, P.lines, P.species_mass, P.species_is_mol, P.n_div_U * n_div_U,
                   P.temps * temps, P.n_e * n_e, P.n_HI, P.n_HeI, P.n_H2, P.xi, P.cntm_mode,
                   P.cntm_values, P.cntm_knots, P.cutoff_threshold)


@pytest.mark.parametrize("precision,dtype", [("fp32", "float32"), ("fp64", "float64")])
def test_replays_equal_direct_call(ctx, precision, dtype):
    probs = _batch()
    ref = _eager(ctx, probs, precision, dtype)
    dprobs = [DeviceProblem(P) for P in probs]
    alphas = [d.new_alpha(dtype) for d in dprobs]
    with ctx.plan_batched(alphas, dprobs, precision) as plan:
        for rep in range(3):
            _zero(alphas)
            plan.launch()
            for k, (a, r) in enumerate(zip(alphas, ref)):
                assert np.array_equal(a.numpy(), r), (rep, k)
            assert ctx.stats()["n_evals"] > 0
            _eager(ctx, probs[:2], precision, dtype)  # other work on the context between replays
        info = plan.info()
        assert info["n_captures"] == 1 and info["n_stars"] == len(probs)
        assert info["n_lines"] == sum(len(P.lines) for P in probs)
        if precision == "fp32":
            assert 0 < info["tile_entries"] <= info["tile_capacity"]


def test_replay_accumulates_and_async_wait(ctx):
    probs = _batch(n_stars=3, seed=22)
    ref = _eager(ctx, probs, "fp32", "float32")
    dprobs = [DeviceProblem(P) for P in probs]
    rng = np.random.default_rng(5)
    bases = [rng.uniform(0, 1e-6, (P.n_wl, P.n_layers)).astype(np.float32) for P in probs]
    alphas = [d.new_alpha("float32") for d in dprobs]
    direct = [d.new_alpha("float32") for d in dprobs]
    for a, c, b in zip(alphas, direct, bases):
        a.copy_from(b)
        c.copy_from(b)
    ctx.line_absorption_batched(direct, dprobs, "fp32")
    with ctx.plan_batched(alphas, dprobs, "fp32") as plan:
        plan.launch(synchronize=False)
        plan.wait()
        for a, c in zip(alphas, direct):
            assert np.array_equal(a.numpy(), c.numpy())
    assert not np.array_equal(ref[0], bases[0])


def test_replay_reads_new_inputs(ctx):
    """DeviceProblem.update writes new inputs into the captured buffers (a fit's next iterate)."""
    probs = _batch(n_stars=4, seed=23)
    dprobs = [DeviceProblem(P) for P in probs]
    alphas = [d.new_alpha("float32") for d in dprobs]
    with ctx.plan_batched(alphas, dprobs, "fp32") as plan:
        new = [_scaled(P, temps=1.02, n_e=1.1) for P in probs]
        for d, P in zip(dprobs, new):
            d.update(P)
        _zero(alphas)
        plan.launch()
        ref = _eager(ctx, new, "fp32", "float32")
        for k, (a, r) in enumerate(zip(alphas, ref)):
            assert np.array_equal(a.numpy(), r), k


def test_replay_regrows_tile_lists(ctx):
    """Inputs that need more tile-list entries than the plan holds: the replay writes nothing, the
    lists grow, the graph is captured again, and the finished launch has the full result."""
    probs = _batch(n_stars=4, seed=24)
    dprobs = [DeviceProblem(P) for P in probs]
    alphas = [d.new_alpha("float32") for d in dprobs]
    with ctx.plan_batched(alphas, dprobs, "fp32", headroom=1.0) as plan:
        cap0 = plan.info()["tile_capacity"]
        stronger = [_scaled(P, n_div_U=30.0) for P in probs]  # wider windows
        for d, P in zip(dprobs, stronger):
            d.update(P)
        _zero(alphas)
        plan.launch()
        info = plan.info()
        assert info["n_captures"] == 2
        assert info["tile_entries"] > cap0 and info["tile_capacity"] >= info["tile_entries"]
        ref = _eager(ctx, stronger, "fp32", "float32")
        for k, (a, r) in enumerate(zip(alphas, ref)):
            assert np.array_equal(a.numpy(), r), k
        _zero(alphas)
        plan.launch()  # the regrown graph replays as captured
        assert plan.info()["n_captures"] == 2
        for k, (a, r) in enumerate(zip(alphas, ref)):
            assert np.array_equal(a.numpy(), r), k


def test_plan_of_empty_stars(ctx):
    P = _batch(n_stars=1, seed=25)[0]
    E = Problem(P.all_wls, P.lines.take(np.arange(0)), P.species_mass, P.species_is_mol,
                P.n_div_U, P.temps, P.n_e, P.n_HI, P.n_HeI, P.n_H2, P.xi, P.cntm_mode,
                P.cntm_values, P.cntm_knots, P.cutoff_threshold)
    d = DeviceProblem(E)
    a = d.new_alpha("float32")
    with ctx.plan_batched([a], [d], "fp32") as plan:
        plan.launch()
        assert plan.info()["n_lines"] == 0
    assert not np.any(a.numpy())


SHAPES = [(32, 1), (32, 2), (32, 4), (64, 1), (64, 2), (64, 4), (128, 1), (128, 2)]


@pytest.mark.parametrize("tile_points,waves_per_tile", SHAPES)
def test_tuning_variants_match_fp64(tile_points, waves_per_tile):
    probs = _batch(n_stars=3, seed=26, scale=0.03)
    with Context(0) as c:
        ref = _eager(c, probs, "fp64", "float64")
        c.set_tuning(tile_points=tile_points, waves_per_tile=waves_per_tile, waves_per_cu=8,
                     near_waves_per_cu=8)
        got = _eager(c, probs, "fp32", "float64")
        for k, (g, r) in enumerate(zip(got, ref)):
            assert rel_err(g.T, r.T) <= FP32_RTOL, (k, rel_err(g.T, r.T))
        # the single-star call under the same tuning adds the same bits
        d = DeviceProblem(probs[1])
        a = d.new_alpha("float64")
        c.line_absorption_device(a, d, "fp32")
        assert np.array_equal(a.numpy(), got[1])


@pytest.mark.parametrize("n_layers", [56, 20, 70, 130])
def test_near_layouts_identical(n_layers):
    """The near zone's two lane layouts (lanes over 64 layers, one point at a time; lanes over
    8 layers x 8 points), run after the far wing or beside it on a second stream (its sums added
    afterwards), accumulate the same terms in the same order: identical alpha, for layer counts
    below, above and at several 64-layer blocks, on a batch of stars."""
    probs = _batch(n_stars=3, seed=28, scale=0.03)
    if n_layers != probs[0].n_layers:
        probs = [_layers(P, n_layers) for P in probs]
    with Context(0) as c:
        out = []
        for layout, conc in ((1, 1), (2, 1), (2, 2)):
            c.set_tuning(near_layout=layout, near_concurrent=conc)
            out.append(_eager(c, probs, "fp32", "float32"))
        for k, (x, y, z) in enumerate(zip(*out)):
            assert np.any(x) and np.array_equal(x, y) and np.array_equal(x, z), k


def test_invalid_tuning_rejected():
    from korg_line._lib import KorgError
    with Context(0) as c:
        for kw in ({"tile_points": 128, "waves_per_tile": 4}, {"tile_points": 48},
                   {"waves_per_tile": 3}, {"near_waves_per_cu": 65}, {"near_layout": 3},
                   {"near_concurrent": 3}):
            with pytest.raises(KorgError, match="tuning"):
                c.set_tuning(**kw)


def test_autotune_and_cache(tmp_path):
    probs = _batch(n_stars=4, seed=27, scale=0.03)
    cache = tmp_path / "tuning.txt"
    with Context(0) as c:
        dprobs = [DeviceProblem(P) for P in probs]
        alphas = [d.new_alpha("float32") for d in dprobs]
        best = c.autotune_batched(alphas, dprobs, "fp32", cache_path=cache)
        assert all(not np.any(a.numpy()) for a in alphas)  # timing used scratch alphas
        assert best["tile_points"] in (32, 64, 128) and best["waves_per_tile"] in (1, 2, 4)
        assert c.tuning() == best
        rows = [r for r in cache.read_text().splitlines() if r and not r.startswith("#")]
        assert len(rows) == 1 and rows[0].split()[0].startswith("gfx")
        # a cached key is used without timing
        c.set_tuning()
        again = c.autotune_batched(alphas, dprobs, "fp32", cache_path=cache)
        assert again == best and c.tuning() == best
        assert len(cache.read_text().splitlines()) == 1
        # the tuned pipeline, direct and as a plan
        ref = _eager(c, probs, "fp64", "float64")
        c.line_absorption_batched(alphas, dprobs, "fp32")
        for a, r in zip(alphas, ref):
            assert rel_err(a.numpy().T.astype(np.float64), r.T) <= FP32_RTOL
        _zero(alphas)
        c.line_absorption_batched(alphas, dprobs, "fp32")
        direct = [a.numpy() for a in alphas]
        _zero(alphas)
        with c.plan_batched(alphas, dprobs, "fp32") as plan:
            assert plan.info()["tuning"]["tile_points"] == best["tile_points"]
            plan.launch()
            for a, r in zip(alphas, direct):
                assert np.array_equal(a.numpy(), r)
