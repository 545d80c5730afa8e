"""The fit-gradient chain (korg_line.fit) of the differentiated single star: d chi^2 /
d(Teff, log g, [M/H], vmic, [Fe/H], [Ti/H], [Ni/H], [Ca/H]) with the line opacity, Korg's planar
transfer and their reverse modes on the device (one fit step), for the 2,000-line solar workload at
5000-5100 A and data a seeded noisy copy (sigma = 0.01).  Needs a GPU.

The windows are frozen at the iterate's own (as ForwardDiff holds them: Korg's Fit differentiates
with duals, src/Fit/fit_via_synthesis.jl:345-355), which makes chi^2 a smooth function of theta
apart from the jumps of two piecewise approximations -- Hunger's Voigt function (up to 10% at its
branch boundaries) and Korg's E2 (up to 1.6%) -- that derivatives do not see.  Central
differences of it pin the whole chain (the pattern of test/autodiff.jl:23-45) once no evaluation
crosses a boundary: the step (1e-6 relative, then divided by 3) is shrunk until two successive
differences agree to 3e-6, which a crossing (its jump over the step) cannot do.
  * Float64 chain against that central difference: <= 1e-5 relative per parameter;
  * Float32 chain against the Float64 chain: <= 1e-5 of the gradient's norm, per parameter.
"""
import numpy as np
import pytest

from korg_line import Wavelengths, workloads
from korg_line.fit import PARAMS, StarModel, noisy_data

pytestmark = pytest.mark.gpu

TRUE = np.array([5777.0, 4.44, 0.0, 1.0, 0.0, 0.0, 0.0, 0.0])
AT = np.array([5820.0, 4.38, -0.04, 1.15, 0.03, -0.02, 0.02, -0.01])


@pytest.fixture(scope="module")
def star(ctx):
    ll = workloads.synthetic_linelist(1, 2000, 5000.0, 5100.0, (-4.0, 0.5), (0.0, 5.0))
    model = StarModel(ll, Wavelengths((5000, 5100)), ctx, precision="fp64")
    data = noisy_data(model, TRUE, sigma=0.01, seed=4)
    return ll, model, data, model.windows(AT)


def _scale(name, x):
    return abs(x) if name in ("Teff", "vmic") else 1.0


def test_chi2_gradient_matches_central_differences(star):
    _, model, data, W = star
    chi2, grad = model.chi2_and_grad(AT, data, 0.01, windows=W)
    assert chi2 == pytest.approx(model.chi2(AT, data, 0.01, windows=W), rel=1e-12)
    lines = []
    for k, name in enumerate(PARAMS):
        def fd(h):
            tp, tm = AT.copy(), AT.copy()
            tp[k] += h
            tm[k] -= h
            return (model.chi2(tp, data, 0.01, windows=W) -
                    model.chi2(tm, data, 0.01, windows=W)) / (2 * h)
        h = 1e-6 * _scale(name, AT[k])
        prev = fd(h)
        for _ in range(4):
            h /= 3.0
            cur = fd(h)
            if abs(cur - prev) <= 3e-6 * abs(cur):
                break
            prev = cur
        else:
            pytest.fail(f"{name}: no two successive steps agree")
        err = abs(grad[k] - cur) / abs(cur)
        lines.append(f"{name:5s} adjoint {grad[k]:+.9e} central {cur:+.9e} rel err {err:.2e}")
        assert err <= 1e-5, (name, grad[k], cur, err)
    print("\nchi2", chi2, "\n" + "\n".join(lines))


def test_fp32_chain_matches_fp64(star, ctx):
    ll, model, data, W = star
    _, g64 = model.chi2_and_grad(AT, data, 0.01, windows=W)
    m32 = StarModel(ll, model.wls, ctx, precision="fp32")
    _, g32 = m32.chi2_and_grad(AT, data, 0.01, windows=W)
    err = np.abs(g32 - g64) / np.linalg.norm(g64)
    print("\n" + "\n".join(f"{n:5s} fp64 {a:+.6e} fp32 {b:+.6e} err/|g| {e:.1e}"
                           for n, a, b, e in zip(PARAMS, g64, g32, err)))
    assert np.all(err <= 1e-5), err
