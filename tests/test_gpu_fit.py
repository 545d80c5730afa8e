"""The fit-gradient chain (korg_line.fit): d chi^2 / d(Teff, log g, [M/H], vmic, [Fe/H], [Ti/H],
[Ni/H], [Ca/H]) with the line opacity differentiated by the device adjoint, against central
differences of chi^2 (each recomputing the device line opacity), steps 1e-4 relative (Teff,
vmic) or 1e-4 dex.  The configuration of the differentiated single star: the 2,000-line solar
workload at 5000-5100 A, data a seeded noisy copy (sigma = 0.01).  Needs a GPU.

Tolerance 2e-3 relative: the central differences of chi^2 cross window edges (the windows are
integer functions of the inputs and jump), which biases them by up to ~1e-3 whatever the step;
with the windows held fixed the adjoint agrees with central differences to 1e-5
(tests/test_gpu_adjoint.py, tests/test_adjoint.py).
"""
import numpy as np
import pytest

from korg_line import Wavelengths, workloads
from korg_line.fit import PARAMS, StarModel, noisy_data

pytestmark = pytest.mark.gpu

TRUE = np.array([5777.0, 4.44, 0.0, 1.0, 0.0, 0.0, 0.0, 0.0])
AT = np.array([5820.0, 4.38, -0.04, 1.15, 0.03, -0.02, 0.02, -0.01])


def test_chi2_gradient_matches_central_differences(ctx):
    ll = workloads.synthetic_linelist(1, 2000, 5000.0, 5100.0, (-4.0, 0.5), (0.0, 5.0))
    model = StarModel(ll, Wavelengths((5000, 5100)), ctx, precision="fp64")
    data = noisy_data(model, TRUE, sigma=0.01, seed=4)
    chi2, grad = model.chi2_and_grad(AT, data, 0.01)
    assert chi2 == pytest.approx(model.chi2(AT, data, 0.01), rel=1e-12)
    fd = np.zeros_like(grad)
    for k, name in enumerate(PARAMS):
        h = 1e-4 * abs(AT[k]) if name in ("Teff", "vmic") else 1e-4
        tp, tm = AT.copy(), AT.copy()
        tp[k] += h
        tm[k] -= h
        fd[k] = (model.chi2(tp, data, 0.01) - model.chi2(tm, data, 0.01)) / (2 * h)
    print("\nchi2", chi2, "\n" + "\n".join(f"{n:5s} adjoint {g:+.9e} central {d:+.9e} rel {abs(g - d) / abs(d):.2e}"
                                        for n, g, d in zip(PARAMS, grad, fd)))
    for k, name in enumerate(PARAMS):
        assert grad[k] == pytest.approx(fd[k], rel=2e-3, abs=1e-6 * np.max(np.abs(fd))), \
            (name, grad[k], fd[k])
