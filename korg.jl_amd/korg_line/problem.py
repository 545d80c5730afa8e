"""Flattening of one ``line_absorption!`` call into the arrays the C ABI consumes.

``Problem.from_korg_args`` mirrors the argument handling of ``line_absorption!``
(``src/line_absorption.jl:33-60``):

* ``n_div_U[s, l] = n_densities[s][l] / partition_fns[s](log(temps[l]))`` for every species in the
  line list (``:42-44``), evaluated on the host because partition functions are arbitrary
  callables;
* the vdW perturber densities n(H I), n(He I), n(H2) are passed through; the device forms
  ``n_eff_vdW`` itself (``:54-60``);
* ``xi`` may be a scalar or a per-layer vector (cm/s);
* ``alpha_cntm`` is either a list of per-layer callables (evaluated at each line centre, the
  generic form of ``:108``) or a ``ContinuumGrid`` (the knots and values of the piecewise-linear
  interpolators that ``synthesize`` builds, ``src/synthesize.jl:226-229``), which the device
  interpolates itself.
"""
import logging

import numpy as np

from .species import Species

CNTM_AT_LINES = 0
CNTM_GRID = 1

log = logging.getLogger("korg_line")


class ContinuumGrid:
    """Per-layer piecewise-linear continuum: ``values[l, k]`` at ``knots[k]`` (cm, ascending)."""

    def __init__(self, knots, values):
        self.knots = np.ascontiguousarray(knots, dtype=np.float64)
        self.values = np.ascontiguousarray(values, dtype=np.float64)
        if self.values.ndim != 2 or self.values.shape[1] != self.knots.size:
            raise ValueError("values must have shape (n_layers, n_knots)")
        if self.knots.size < 2 or np.any(np.diff(self.knots) <= 0):
            raise ValueError("knots must be strictly ascending, at least 2")

    def layer(self, l):
        k, v = self.knots, self.values[l]

        def f(x):
            i = int(np.clip(np.searchsorted(k, x, side="right") - 1, 0, k.size - 2))
            t = (x - k[i]) / (k[i + 1] - k[i])
            return (1.0 - t) * v[i] + t * v[i + 1]
        return f


class Problem:
    """ABI-ready arrays of one ``line_absorption!`` call (all float64 unless noted).

    ``all_wls`` [n_wl] cm; ``lines`` a ``LineList``; ``species_mass``/``species_is_mol`` [S];
    ``n_div_U`` [S, L]; ``temps``, ``n_e``, ``n_HI``, ``n_HeI``, ``n_H2``, ``xi`` [L];
    ``cntm_mode`` and ``cntm_values`` ([N, L] at lines, or [L, n_knots] on ``cntm_knots``).
    """

    def __init__(self, all_wls, lines, species_mass, species_is_mol, n_div_U, temps, n_e, n_HI,
                 n_HeI, n_H2, xi, cntm_mode, cntm_values, cntm_knots, cutoff_threshold):
        c = np.ascontiguousarray
        self.all_wls = c(all_wls, dtype=np.float64)
        self.lines = lines
        self.species_mass = c(species_mass, dtype=np.float64)
        self.species_is_mol = c(species_is_mol, dtype=np.uint8)
        self.n_div_U = c(n_div_U, dtype=np.float64)
        self.temps = c(temps, dtype=np.float64)
        self.n_e = c(n_e, dtype=np.float64)
        self.n_HI = c(n_HI, dtype=np.float64)
        self.n_HeI = c(n_HeI, dtype=np.float64)
        self.n_H2 = c(n_H2, dtype=np.float64)
        self.xi = c(xi, dtype=np.float64)
        self.cntm_mode = int(cntm_mode)
        self.cntm_values = c(cntm_values, dtype=np.float64)
        self.cntm_knots = c(cntm_knots if cntm_knots is not None else np.zeros(0),
                            dtype=np.float64)
        self.n_knots = self.cntm_knots.size
        self.cutoff_threshold = float(cutoff_threshold)
        # frozen windows (1-based int64, as Context.windows returns them), or None: computed
        self.window_lb = None
        self.window_ub = None

    def with_windows(self, lb, ub):
        """A copy of this problem whose line windows are frozen at ``lb``/``ub`` (1-based, one
        per line; lb > ub skips the line) instead of computed from the inputs."""
        import copy
        lb = np.ascontiguousarray(lb, dtype=np.int64)
        ub = np.ascontiguousarray(ub, dtype=np.int64)
        if lb.shape != (len(self.lines),) or ub.shape != (len(self.lines),):
            raise ValueError("one window per line")
        P = copy.copy(self)
        P.window_lb, P.window_ub = lb, ub
        return P

    @property
    def n_layers(self):
        return self.temps.size

    @property
    def n_wl(self):
        return self.all_wls.size

    @classmethod
    def from_korg_args(cls, linelist, wls, temps, n_e, n_densities, partition_fns, xi,
                       alpha_cntm, cutoff_threshold=3e-4):
        temps = np.asarray(temps, dtype=np.float64)
        L = temps.size
        lnT = np.log(temps)
        species = linelist.species
        n_div_U = np.zeros((len(species), L))
        used = np.unique(linelist.species_index)
        for s in used:
            sp = species[s]
            n_div_U[s] = np.asarray(n_densities[sp]) / np.asarray(partition_fns[sp](lnT))
        if Species("H I") in [species[s] for s in used]:
            log.error("Atomic hydrogen should not be in the linelist. Korg has built-in hydrogen "
                      "lines.")

        def dens(code):
            sp = Species(code)
            return np.asarray(n_densities[sp], dtype=np.float64) if sp in n_densities \
                else np.full(L, np.nan)

        xi = np.broadcast_to(np.asarray(xi, dtype=np.float64), (L,)).copy()
        if isinstance(alpha_cntm, ContinuumGrid):
            if alpha_cntm.values.shape[0] != L:
                raise ValueError("continuum grid has the wrong number of layers")
            mode, vals, knots = CNTM_GRID, alpha_cntm.values, alpha_cntm.knots
        else:
            if len(alpha_cntm) != L:
                raise ValueError("alpha_cntm must have one callable per layer")
            vals = np.empty((len(linelist), L))
            for l, f in enumerate(alpha_cntm):
                vals[:, l] = [f(w) for w in linelist.wl]
            mode, knots = CNTM_AT_LINES, None
        all_wls = wls.all_wls if hasattr(wls, "all_wls") else np.asarray(wls, dtype=np.float64)
        return cls(all_wls, linelist, [sp.get_mass() for sp in species],
                   [sp.ismolecule() for sp in species], n_div_U, temps, np.asarray(n_e),
                   dens("H I"), dens("He I"), dens("H2"), xi, mode, vals, knots,
                   cutoff_threshold)
