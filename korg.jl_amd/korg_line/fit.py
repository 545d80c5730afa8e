"""Gradient of a spectral-fit objective through the device line opacity, transfer and adjoint.

The chain of a fit: stellar parameters theta -> per-layer atmosphere inputs (T, n_e, n_s / U_s,
xi, perturber densities) -> line opacity alpha_lines + continuum -> normalized flux (Korg's planar
transfer) -> chi^2 = sum ((F - data) / sigma)^2.  ``StarModel.chi2_and_grad`` computes
d chi^2 / d theta with

* the line-opacity path on the device: one fit step (``korg_fit_gradient_batched``) gives chi^2
  and d chi^2 / d(per-layer inputs) through alpha_lines -- the transfer's reverse mode gives
  G = d chi^2 / d alpha_lines and the line-opacity adjoint turns it into per-layer gradients --
  chained to theta through the atmosphere's per-layer inputs, whose derivatives are central
  differences of the stand-in atmosphere (cheap: no line opacity is evaluated);
* everything else -- the continuum, the reference opacity and optical depth, and the Planck
  source, with alpha_lines held fixed -- by central differences of the device transfer
  (``korg_flux_batched``; no line opacity either).

Windows: the reference's windows are integer functions of the inputs, and ForwardDiff (which Korg's
Fit uses, ``src/Fit/fit_via_synthesis.jl:345-355``) carries no derivative through them.  Passing
``windows=(lb, ub)`` freezes them (``Problem.with_windows``), which makes chi^2 a smooth function
of theta -- the function whose derivative the chain computes -- so that central differences of it
check the gradient (``test/autodiff.jl:23-45`` pattern).  The atmosphere and chemistry are the
stand-ins of ``korg_line.atmosphere``; only the line opacity and the transfer are the reference's
arithmetic.
"""
import numpy as np

from .atmosphere import Atmosphere, continuum_grid
from .hipmem import DeviceArray
from .line_absorption import DeviceProblem, FitTarget, FluxTarget
from .problem import ContinuumGrid, Problem

PARAMS = ("Teff", "logg", "M_H", "vmic", "Fe_H", "Ti_H", "Ni_H", "Ca_H")
# relative (Teff, vmic) or absolute (log g and abundances, in dex) central-difference steps of the
# stand-in atmosphere (smooth apart from its piecewise-linear partition functions)
_STEP = {"Teff": ("rel", 1e-6), "logg": ("abs", 1e-6), "M_H": ("abs", 1e-6),
         "vmic": ("rel", 1e-6), "Fe_H": ("abs", 1e-6), "Ti_H": ("abs", 1e-6),
         "Ni_H": ("abs", 1e-6), "Ca_H": ("abs", 1e-6)}


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
        self.dtype = "float64" if precision == "fp64" else "float32"
        self.cutoff = cutoff_threshold

    # ------------------------------------------------------------------ inputs of the line opacity
    def atmosphere(self, th):
        t = dict(zip(PARAMS, th))
        ab = {"Fe": t["Fe_H"], "Ti": t["Ti_H"], "Ni": t["Ni_H"], "Ca": t["Ca_H"]}
        return Atmosphere.scaled(t["Teff"], t["logg"], t["M_H"], abundances=ab), t["vmic"] * 1e5

    def problem(self, th, windows=None):
        atm, xi = self.atmosphere(th)
        knots, vals = continuum_grid(atm, float(self.wl[0]), float(self.wl[-1]))
        P = Problem.from_korg_args(self.ll, self.wls, atm.temps, atm.n_e, atm.n_densities,
                                   atm.partition_fns, xi, ContinuumGrid(knots, vals),
                                   self.cutoff)
        if windows is not None:
            P = P.with_windows(*windows)
        return atm, P

    @staticmethod
    def layer_inputs(P):
        """The per-layer inputs the adjoint differentiates, as one flat vector."""
        return np.concatenate([P.temps, P.n_e, P.xi, P.n_HI, P.n_HeI, P.n_H2, P.n_div_U.ravel()])

    @staticmethod
    def anchor(atm):
        """The anchored scheme's reference optical depth and absorption (5000 A) per layer."""
        return 10.0 ** atm.lgTau5, atm.continuum(np.array([5000e-8]))[:, 0]

    # ------------------------------------------------------------------ forward pieces
    def windows(self, th):
        """The windows (1-based lb, ub per line) the line opacity computes at theta."""
        _, P = self.problem(th)
        D = DeviceProblem(P)
        self.ctx.line_absorption_device(D.new_alpha(self.dtype), D, self.precision)
        return self.ctx.windows(len(self.ll))

    def _normalized_flux(self, atm, D, alpha):
        ft = FluxTarget(D.problem, *self.anchor(atm))
        self.ctx.flux_batched([alpha], [D], [ft])
        return ft.flux() / ft.cntm()

    def flux(self, th, windows=None):
        """Normalized flux at theta: continuum on the grid + lines, Korg's planar transfer, all
        on the device."""
        atm, P = self.problem(th, windows)
        D = DeviceProblem(P)
        a = D.new_alpha(self.dtype)
        self.ctx.continuum_batched([a], [D])
        self.ctx.line_absorption_device(a, D, self.precision)
        return self._normalized_flux(atm, D, a)

    def chi2(self, th, data, sigma, windows=None):
        r = (self.flux(th, windows) - data) / sigma
        return float(np.sum(r * r))

    # ------------------------------------------------------------------ gradient
    def chi2_and_grad(self, th, data, sigma, windows=None):
        """chi^2 at theta and d chi^2 / d theta (the windows held at ``windows`` when given, else
        at theta's own)."""
        th = np.asarray(th, dtype=np.float64)
        atm, P = self.problem(th, windows)
        D = DeviceProblem(P)
        a = D.new_alpha(self.dtype)
        tg = FitTarget(P, *self.anchor(atm), data, sigma)
        self.ctx.fit_gradient_batched([a], [D], [tg], self.precision)
        chi2 = tg.chi2()
        raw = tg.grads()
        alpha_lines = a.numpy().astype(np.float64)  # (n_wl, n_layers): the device layout
        g_inputs = np.concatenate([raw["T"], raw["n_e"], raw["xi"], raw["n_HI"], raw["n_HeI"],
                                   raw["n_H2"], raw["n_div_U"].ravel()])
        grad = np.zeros(len(PARAMS))
        for k, name in enumerate(PARAMS):
            h = _step(name, th[k])
            # the line path: the adjoint's gradients along d(per-layer inputs)/d theta
            plus, minus = th.copy(), th.copy()
            plus[k] += h
            minus[k] -= h
            d_inputs = (self.layer_inputs(self.problem(plus, windows)[1]) -
                        self.layer_inputs(self.problem(minus, windows)[1])) / (2 * h)
            ok = np.isfinite(d_inputs) & np.isfinite(g_inputs)
            grad[k] = float(np.sum(g_inputs[ok] * d_inputs[ok])) + \
                self._rest_derivative(th, k, h, windows, alpha_lines, data, sigma)
        return chi2, grad

    def _chi2_lines_fixed(self, th, windows, alpha_lines, data, sigma):
        """chi^2 with alpha_lines held fixed and the continuum, anchor and source at theta."""
        atm, P = self.problem(th, windows)
        D = DeviceProblem(P)
        tot = D.new_alpha("float64")
        self.ctx.continuum_batched([tot], [D])
        tot = DeviceArray.from_numpy(tot.numpy() + alpha_lines)
        F = self._normalized_flux(atm, D, tot)
        return float(np.sum(((F - data) / sigma) ** 2))

    def _rest_derivative(self, th, k, h, windows, alpha_lines, data, sigma, rtol=3e-7):
        """d chi^2 / d theta_k with alpha_lines held fixed, by central differences.  Korg's E2
        approximation is piecewise (src/RadiativeTransfer/RadiativeTransfer.jl:408-480) and jumps
        by up to 1.6% at its piece boundaries, so a step that moves an optical depth across one
        is off by the jump over the step; derivatives have no such term.  The step is divided by
        3 until two successive differences agree to rtol: both are then free of crossings."""
        def fd(step):
            plus, minus = th.copy(), th.copy()
            plus[k] += step
            minus[k] -= step
            return (self._chi2_lines_fixed(plus, windows, alpha_lines, data, sigma) -
                    self._chi2_lines_fixed(minus, windows, alpha_lines, data, sigma)) / (2 * step)
        prev = fd(h)
        for _ in range(4):
            h /= 3.0
            cur = fd(h)
            if abs(cur - prev) <= rtol * max(abs(cur), abs(prev)):
                return cur
            prev = cur
        return prev


def noisy_data(model, th, sigma=0.01, seed=1):
    """A seeded noisy copy of the model flux at theta (Gaussian noise of width sigma)."""
    rng = np.random.default_rng(seed)
    F = model.flux(th)
    return F + sigma * rng.standard_normal(F.shape)
