"""Pins of the Float64 radiative-transfer oracle (oracle/rt_oracle.c): Korg's plane-parallel
default (anchored optical depth, linear source, exponential-integral flux,
src/RadiativeTransfer/RadiativeTransfer.jl) and its reverse mode.  CPU only.
"""
import numpy as np
import pytest
from scipy.integrate import quad
from scipy.special import expn

from oracle import oracle

H, C, KB = 6.62607015e-27, 2.99792458e10, 1.380649e-16


def test_expint2_matches_exact_on_reference_range():
    """test/transfer.jl:4-7: exponential_integral_2 ≈ expint(2, x), rtol 1e-3, x in 2:0.01:8."""
    xs = np.arange(2.0, 8.0 + 1e-9, 0.01)
    e = np.array([oracle.expint2(x)[0] for x in xs])
    np.testing.assert_allclose(e, expn(2, xs), rtol=1e-3)


def test_expint2_within_one_percent_everywhere():
    """The approximation's stated accuracy (src/RadiativeTransfer/RadiativeTransfer.jl:405-407)."""
    xs = np.concatenate([np.geomspace(1e-8, 1.0, 300), np.linspace(1.0, 60.0, 600)])
    e = np.array([oracle.expint2(x)[0] for x in xs])
    np.testing.assert_allclose(e, expn(2, xs), rtol=1e-2)
    assert oracle.expint2(0.0)[0] == 1.0


@pytest.mark.parametrize("x", [1e-3, 0.3, 1.0, 1.7, 3.1, 4.2, 5.0, 6.3, 7.0, 8.4, 12.0, 30.0])
def test_expint2_derivative_by_central_differences(x):
    h = 1e-6 * x
    fd = (oracle.expint2(x + h)[0] - oracle.expint2(x - h)[0]) / (2 * h)
    assert oracle.expint2(x)[1] == pytest.approx(fd, rel=1e-6, abs=1e-12)
    # the approximation's derivative tracks the exact one, dE2/dx = -E1
    assert oracle.expint2(x)[1] == pytest.approx(-expn(1, x), rel=5e-2)


@pytest.mark.parametrize("tau,m,b", [(0.3, 2.0, 1.0), (1.7, -0.5, 3.0), (6.0, 1.0, 0.2),
                                     (15.0, 0.7, -0.4)])
def test_transfer_core_partials(tau, m, b):
    v, dt, dm, db = oracle.transfer_core(tau, m, b)
    for k, d in ((0, dt), (1, dm), (2, db)):
        x = [tau, m, b]
        h = 1e-6 * max(1.0, abs(x[k]))
        xp, xm = list(x), list(x)
        xp[k] += h
        xm[k] -= h
        fd = (oracle.transfer_core(*xp)[0] - oracle.transfer_core(*xm)[0]) / (2 * h)
        assert d == pytest.approx(fd, rel=1e-6, abs=1e-12)


def test_transfer_core_is_an_antiderivative():
    """With the exact E2 the core is an antiderivative of (m tau + b) E2(tau); with Korg's E2 the
    integral over an interval agrees to the approximation's accuracy."""
    m, b = 0.8, 1.5
    for t0, t1 in ((0.1, 0.9), (1.5, 3.0), (4.0, 9.5)):
        exact = quad(lambda t: (m * t + b) * expn(2, t), t0, t1)[0]
        core = oracle.transfer_core(t1, m, b)[0] - oracle.transfer_core(t0, m, b)[0]
        assert core == pytest.approx(exact, rel=1e-2)


def test_blackbody_integrates_to_stefan_boltzmann():
    """∫ B_λ dλ = σ T^4 / π (src/synthesize.jl:412-417)."""
    sigma_sb = 2 * np.pi ** 5 * KB ** 4 / (15 * H ** 3 * C ** 2)
    for T in (3500.0, 5777.0, 9000.0):
        lo, hi = 1e-6, 1e-2  # cm
        xs = np.geomspace(lo, hi, 20001)
        B = np.array([oracle.blackbody(T, x) for x in xs])
        integral = np.trapezoid(B * xs, np.log(xs))
        assert integral == pytest.approx(sigma_sb * T ** 4 / np.pi, rel=1e-4)


def _temps_for_source(S, lam):
    """Temperatures whose blackbody at lam equals S (inverse Planck function)."""
    return H * C / (lam * KB * np.log1p(2 * H * C ** 2 / (lam ** 5 * S)))


def test_anchored_tau_with_reference_alpha():
    """alpha = alpha_ref: tau_i = trapezoid of tau_ref over ln tau_ref -> tau_ref_i - tau_ref_1, and
    a constant source S gives F = 2 pi S (1/2 - E3(tau_max)) = pi S for a deep atmosphere."""
    L = 400
    tau_ref = np.geomspace(1e-6, 60.0, L)
    alpha_ref = np.geomspace(1e-9, 1e-5, L)
    lam = 5e-5
    S = 2.3e14
    temps = np.full(L, _temps_for_source(S, lam))
    F = oracle.rt_flux(alpha_ref[:, None], np.array([lam]), temps, tau_ref, alpha_ref)[0]
    assert F == pytest.approx(np.pi * S, rel=1e-2)  # E2 approximation accuracy


def test_linear_source_deep_flux():
    """S = b + m tau on a deep atmosphere: F = 2 pi (b/2 + m/3) (∫E2 = 1/2, ∫tau E2 = 1/3)."""
    L = 600
    tau_ref = np.geomspace(1e-7, 80.0, L)
    alpha_ref = np.full(L, 1e-7)
    lam = 6e-5
    # tau of alpha = alpha_ref: the trapezoid over ln tau_ref
    lt = np.log(tau_ref)
    tau = np.concatenate([[0.0], np.cumsum(0.5 * (tau_ref[1:] + tau_ref[:-1]) * np.diff(lt))])
    b, m = 1.0e14, 3.0e13
    temps = _temps_for_source(b + m * tau, lam)
    F = oracle.rt_flux(alpha_ref[:, None], np.array([lam]), temps, tau_ref, alpha_ref)[0]
    assert F == pytest.approx(2 * np.pi * (b / 2 + m / 3), rel=1e-2)


def test_rt_flux_vjp_matches_central_differences():
    rng = np.random.default_rng(4)
    L, n = 14, 5
    tau_ref = np.geomspace(1e-4, 30.0, L)
    alpha_ref = np.geomspace(1e-8, 1e-4, L)
    alpha = alpha_ref[:, None] * rng.uniform(0.5, 3.0, (L, n))
    wls = np.linspace(5000e-8, 5001e-8, n)
    temps = np.linspace(4000.0, 9000.0, L)
    gF = rng.standard_normal(n)
    g = oracle.rt_flux_vjp(alpha, wls, temps, tau_ref, alpha_ref, gF)
    # a step of 1e-4 relative: smaller ones are rounding-limited (a surface layer moves F by a
    # tiny fraction), and none of these steps moves a tau across a piece boundary of E2
    for l in range(L):
        for j in range(n):
            h = 1e-4 * alpha[l, j]
            ap, am = alpha.copy(), alpha.copy()
            ap[l, j] += h
            am[l, j] -= h
            fd = np.dot(gF, oracle.rt_flux(ap, wls, temps, tau_ref, alpha_ref) -
                        oracle.rt_flux(am, wls, temps, tau_ref, alpha_ref)) / (2 * h)
            assert g[l, j] == pytest.approx(fd, rel=1e-5, abs=1e-9 * np.abs(g).max()), (l, j)
