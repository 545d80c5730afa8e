"""Emergent flux from alpha for the flux-error metric (host, Float64).

A compact form of Korg's default radiative transfer (``src/RadiativeTransfer/RadiativeTransfer.jl``):
optical depth by the anchored scheme -- dtau = (alpha / alpha_ref) dtau_ref integrated along the
layers' reference optical depth (``:244-253``) -- and the flux-only linear-source solution
F = 2 pi int S(tau) E2(tau) dtau with S piecewise linear in tau (``:379-430``).  The source
function is the Planck function of the layer temperature.  It is used to compare two alphas on the
same atmosphere, so its own approximations cancel in the comparison.
"""
import numpy as np
from scipy.special import expn

from .constants import c_cgs, hplanck_cgs, kboltz_cgs


def planck(T, wl_cm):
    T = np.asarray(T)[:, None]
    lam = np.asarray(wl_cm)[None, :]
    return 2 * hplanck_cgs * c_cgs ** 2 / lam ** 5 / np.expm1(hplanck_cgs * c_cgs /
                                                            (lam * kboltz_cgs * T))


def flux(alpha, alpha_ref, tau_ref, temps, wl_cm):
    """alpha: (n_layers, n_wl) [cm^-1]; alpha_ref, tau_ref, temps: (n_layers,).  Returns F(wl)."""
    alpha = np.asarray(alpha, dtype=np.float64)
    ratio = alpha / np.asarray(alpha_ref)[:, None]
    dref = np.diff(tau_ref)[:, None]
    tau = np.zeros_like(ratio)
    tau[1:] = np.cumsum(0.5 * (ratio[1:] + ratio[:-1]) * dref, axis=0)
    tau += ratio[:1] * tau_ref[0]
    S = planck(temps, wl_cm)
    # int over [t0, t1] of (a + b t) E2(t) dt with E2' = -E1: use E3, E4 antiderivatives
    t0, t1 = tau[:-1], tau[1:]
    s0, s1 = S[:-1], S[1:]
    dt = np.maximum(t1 - t0, 1e-300)
    b = (s1 - s0) / dt
    a = s0 - b * t0

    def prim(t):  # antiderivative of (a + b t) E2(t)
        return -a * expn(3, t) - b * (t * expn(3, t) + expn(4, t))

    F = np.sum(prim(t1) - prim(t0), axis=0)
    return 2 * np.pi * F


def normalized_flux(alpha_total, alpha_cntm, alpha_ref, tau_ref, temps, wl_cm):
    return flux(alpha_total, alpha_ref, tau_ref, temps, wl_cm) / \
        flux(alpha_cntm, alpha_ref, tau_ref, temps, wl_cm)
