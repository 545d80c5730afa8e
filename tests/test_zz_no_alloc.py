"""The allocation-free entry point, korg_line_absorption_device_noalloc, against
korg_line_absorption_device (the entry point it is a variant of), through the C ABI.

Both run the same kernels in the same fixed summation order, so alpha, the windows and the
evaluation counts must agree exactly (bitwise), in both precisions, on two calls in a row -- the
second with a smaller input, so that it runs on scratch the first call left behind -- and the
Float64 result must match the oracle to 1e-12.  A problem larger than the fixed scratch is refused
on the host with KORG_ERR_CAPACITY and leaves alpha unchanged.
"""
import numpy as np
import pytest

from oracle import oracle
from korg_line import Context, DeviceProblem, Wavelengths, _lib, workloads
from korg_line.atmosphere import Atmosphere

pytestmark = pytest.mark.gpu


def _problems():
    big, _, _ = workloads.solar_small(seed=11, n_lines=2000)          # 2,000 lines, 10,001 points
    atm = Atmosphere.from_model("sun")
    ll = workloads.synthetic_linelist(12, 400, 5000.0, 5030.0, (-4.0, 0.5), (0.0, 5.0))
    small = workloads.make_problem(atm, ll, Wavelengths((5000, 5030)), 1.0)  # 400 lines, 3,001
    return big, small


def _run(ctx, dp, precision, noalloc, base):
    a = dp.new_alpha("float64", fill=base)
    if noalloc:
        ctx.line_absorption_device_noalloc(a, dp, precision)
    else:
        ctx.line_absorption_device(a, dp, precision)
    out = a.numpy().T
    lb, ub = ctx.windows(len(dp.problem.lines))
    return out, lb, ub, ctx.stats()["n_evals"]


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_noalloc_matches_device_entry_point(precision):
    big, small = _problems()
    with Context(0) as ctx:
        for P in (big, small):  # second call: smaller input on the scratch the first left
            print(f"noalloc {precision}: {len(P.lines)} lines x {P.n_layers} layers x {P.n_wl} "
                  f"points", flush=True)
            dp = DeviceProblem(P)
            rng = np.random.default_rng(len(P.lines))
            base = rng.uniform(0.0, 1e-9, (P.n_layers, P.n_wl))  # alpha accumulates
            ref, rlb, rub, rn = _run(ctx, dp, precision, False, base)
            got, glb, gub, gn = _run(ctx, dp, precision, True, base)
            assert np.array_equal(got, ref)
            assert np.array_equal(glb, rlb) and np.array_equal(gub, rub)
            assert gn == rn > 0
            if precision == "fp64":
                want = np.asfortranarray(base.copy())
                oracle.line_absorption(want, P)
                np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-300)


def test_noalloc_refuses_problems_beyond_its_scratch():
    atm = Atmosphere.from_model("sun")
    ll = workloads.synthetic_linelist(13, 40_000, 5000.0, 5100.0, (-4.0, 0.5), (0.0, 5.0))
    P = workloads.make_problem(atm, ll, Wavelengths((5000, 5100)), 1.0)
    # 40,000 lines x 56 layers: the Float64 records alone (32 B per line and layer) exceed 64 MiB
    assert 32 * len(P.lines) * 56 > _lib.NOALLOC_SCRATCH_BYTES
    with Context(0) as ctx:
        dp = DeviceProblem(P)
        a = dp.new_alpha("float32", fill=np.full((P.n_layers, P.n_wl), 2.0))
        with pytest.raises(_lib.KorgError) as e:
            ctx.line_absorption_device_noalloc(a, dp, "fp32")
        assert e.value.code == _lib.KORG_ERR_CAPACITY
        assert "fixed scratch" in str(e.value)
        assert np.all(a.numpy() == 2.0)
        # the context still works afterwards, through both entry points
        small, _, _ = workloads.solar_small(seed=14, n_lines=100)
        d2 = DeviceProblem(small)
        x = d2.new_alpha("float32")
        y = d2.new_alpha("float32")
        ctx.line_absorption_device(x, d2, "fp32")
        ctx.line_absorption_device_noalloc(y, d2, "fp32")
        assert np.array_equal(x.numpy(), y.numpy())
