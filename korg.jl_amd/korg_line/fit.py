"""Gradient of a spectral-fit objective through the device line opacity and its adjoint.

The chain of a fit: stellar parameters theta -> per-layer atmosphere inputs (T, n_e, n_s / U_s,
xi, perturber densities) -> line opacity alpha_lines (the device) + continuum -> normalized flux
-> chi^2 = sum ((F - data) / sigma)^2.  ``StarModel.chi2_and_grad`` computes d chi^2 / d theta with

* the line-opacity path by the device adjoint: G = d chi^2 / d alpha_lines -> dL/d(per-layer
  inputs) (``korg_line_absorption_vjp``), then chained to theta through the atmosphere's
  per-layer inputs, whose derivatives are central differences of the stand-in atmosphere (cheap:
  no line opacity is evaluated);
* everything else -- the continuum, the reference opacity and the Planck source of the flux,
  with alpha_lines held fixed -- by central differences (no line opacity either).

d F / d alpha (the radiative transfer, a host-side stand-in) is taken by central differences per
layer, all wavelengths at once (F at a wavelength depends only on that wavelength's column).

The reference obtains the same gradient with ForwardDiff through ``synthesize``
(``test/autodiff.jl:23-45`` pattern); here the atmosphere and chemistry are the stand-ins of
``korg_line.atmosphere``, so only the line-opacity part is the reference's arithmetic.
"""
import numpy as np

from . import rt
from .atmosphere import Atmosphere, continuum_grid
from .problem import ContinuumGrid, Problem
from .species import Species

PARAMS = ("Teff", "logg", "M_H", "vmic", "Fe_H", "Ti_H", "Ni_H", "Ca_H")
# relative (Teff, vmic) or absolute (log g and abundances, in dex) central-difference steps
_STEP = {"Teff": ("rel", 1e-5), "logg": ("abs", 1e-5), "M_H": ("abs", 1e-5),
         "vmic": ("rel", 1e-5), "Fe_H": ("abs", 1e-5), "Ti_H": ("abs", 1e-5),
         "Ni_H": ("abs", 1e-5), "Ca_H": ("abs", 1e-5)}


def _step(name, x):
    kind, h = _STEP[name]
    return h * abs(x) if kind == "rel" else h


class StarModel:
    """One star: a sorted line list on a ``Wavelengths`` grid, a ``Context`` and a precision."""

    def __init__(self, linelist, wls, ctx, precision="fp64", cutoff_threshold=3e-4):
        self.ll = linelist if linelist.is_sorted() else linelist.sorted()
        self.wls = wls
        self.wl = wls.all_wls
        self.ctx = ctx
        self.precision = precision
        self.cutoff = cutoff_threshold

    # ------------------------------------------------------------------ inputs of the line opacity
    def atmosphere(self, th):
        t = dict(zip(PARAMS, th))
        ab = {"Fe": t["Fe_H"], "Ti": t["Ti_H"], "Ni": t["Ni_H"], "Ca": t["Ca_H"]}
        return Atmosphere.scaled(t["Teff"], t["logg"], t["M_H"], abundances=ab), t["vmic"] * 1e5

    def problem(self, th):
        atm, xi = self.atmosphere(th)
        knots, vals = continuum_grid(atm, float(self.wl[0]), float(self.wl[-1]))
        P = Problem.from_korg_args(self.ll, self.wls, atm.temps, atm.n_e, atm.n_densities,
                                   atm.partition_fns, xi, ContinuumGrid(knots, vals),
                                   self.cutoff)
        return atm, P, knots, vals

    @staticmethod
    def layer_inputs(P):
        """The per-layer inputs the adjoint differentiates, as one flat vector."""
        return np.concatenate([P.temps, P.n_e, P.xi, P.n_HI, P.n_HeI, P.n_H2, P.n_div_U.ravel()])

    # ------------------------------------------------------------------ forward pieces
    def line_opacity(self, P):
        a = np.zeros((P.n_layers, P.n_wl), order="F")
        self.ctx.line_absorption(a, P, self.precision)
        return a

    def _rt_inputs(self, atm, knots, vals):
        cntm = np.array([np.interp(self.wl, knots, vals[l]) for l in range(atm.n_layers)])
        aref = atm.continuum(np.array([5000e-8]))[:, 0]
        return cntm, aref, 10.0 ** atm.lgTau5, atm.temps

    def flux_from(self, alpha_lines, atm, knots, vals):
        cntm, aref, tau, T = self._rt_inputs(atm, knots, vals)
        return rt.normalized_flux(cntm + alpha_lines, cntm, aref, tau, T, self.wl)

    def flux(self, th):
        atm, P, knots, vals = self.problem(th)
        return self.flux_from(self.line_opacity(P), atm, knots, vals)

    def chi2(self, th, data, sigma):
        r = (self.flux(th) - data) / sigma
        return float(np.sum(r * r))

    # ------------------------------------------------------------------ gradient
    def chi2_and_grad(self, th, data, sigma):
        th = np.asarray(th, dtype=np.float64)
        atm, P, knots, vals = self.problem(th)
        alpha_lines = self.line_opacity(P)
        cntm, aref, tau, T = self._rt_inputs(atm, knots, vals)
        Fc = rt.flux(cntm, aref, tau, T, self.wl)
        Ft = rt.flux(cntm + alpha_lines, aref, tau, T, self.wl)
        F = Ft / Fc
        r = (F - data) / sigma
        chi2 = float(np.sum(r * r))
        dchi_dF = 2.0 * r / sigma
        # G = d chi2 / d alpha_lines: dF/d alpha by central differences, one layer at a time
        tot = cntm + alpha_lines
        G = np.zeros_like(tot)
        for l in range(tot.shape[0]):
            h = 1e-6 * tot[l]
            up, dn = tot.copy(), tot.copy()
            up[l] += h
            dn[l] -= h
            dF = (rt.flux(up, aref, tau, T, self.wl) - rt.flux(dn, aref, tau, T, self.wl)) / (2 * h)
            G[l] = dchi_dF * dF / Fc
        # line-opacity path: the adjoint, then d(per-layer inputs)/d theta
        raw = self.ctx.line_absorption_vjp(np.asfortranarray(G), P)
        g_inputs = np.concatenate([raw["T"], raw["n_e"], raw["xi"], raw["n_HI"], raw["n_HeI"],
                                   raw["n_H2"], raw["n_div_U"].ravel()])
        grad = np.zeros(len(PARAMS))
        for k, name in enumerate(PARAMS):
            h = _step(name, th[k])
            tp, tm = th.copy(), th.copy()
            tp[k] += h
            tm[k] -= h
            _, Pp, kp, vp = self.problem(tp)
            atm_p, _ = self.atmosphere(tp)
            _, Pm, km, vm = self.problem(tm)
            atm_m, _ = self.atmosphere(tm)
            d_inputs = (self.layer_inputs(Pp) - self.layer_inputs(Pm)) / (2 * h)
            ok = np.isfinite(d_inputs) & np.isfinite(g_inputs)
            grad_lines = float(np.sum(g_inputs[ok] * d_inputs[ok]))
            # the rest of the chain, alpha_lines held fixed
            Fp = self.flux_from(alpha_lines, atm_p, kp, vp)
            Fm = self.flux_from(alpha_lines, atm_m, km, vm)
            c_p = float(np.sum(((Fp - data) / sigma) ** 2))
            c_m = float(np.sum(((Fm - data) / sigma) ** 2))
            grad[k] = grad_lines + (c_p - c_m) / (2 * h)
        return chi2, grad


def noisy_data(model, th, sigma=0.01, seed=1):
    """A seeded noisy copy of the model flux at theta (Gaussian noise of width sigma)."""
    rng = np.random.default_rng(seed)
    F = model.flux(th)
    return F + sigma * rng.standard_normal(F.shape)
