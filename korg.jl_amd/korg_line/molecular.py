"""Precomputed molecular cross-sections on the device: Korg's ``MolecularCrossSection``
(``src/molecular_cross_sections.jl:3-83``) and ``interpolate_molecular_cross_sections!``
(``:99-114``).

``MolecularCrossSection.compute`` is the precompute of ``:48-83``: for every vmic knot one
``line_absorption!`` call whose "layers" are the log T knots -- nₑ = 0, the molecule's density
1/cutoff_alpha, NaN perturber densities (molecules have no vdW broadening), continuum 1 and cutoff
1.0, so that each line is cut where its cross-section falls below cutoff_alpha -- run as one batched
device call over the vmic knots; the cross-section is that alpha times cutoff_alpha.  The table's
layout is the library's (``korg_mol_xsection``): ``sigma[iv, j, iT]`` (C order), which is exactly
the column-major alpha of star ``iv`` of the batch.

``interpolate_molecular_cross_sections_`` adds sigma(vmic, log10 T, lambda) * n_species into alpha
per layer through the library (linear in each gridded dimension, zero outside the table).
"""
import ctypes

import numpy as np

from . import _lib
from .line_absorption import DeviceProblem, default_context
from .problem import CNTM_GRID, Problem
from .wavelengths import Wavelengths

# the precompute's default knots (src/molecular_cross_sections.jl:48-50)
DEFAULT_VMIC = np.array([0.0, 1 / 3, 2 / 3, 1.0, 1.5, 2.0, 2 + 2 / 3, 2 + 4 / 3, 4.0, 4 + 2 / 3,
                         5 + 1 / 3])
DEFAULT_LOG_T = 3.0 + 0.04 * np.arange(51)  # 3:0.04:5


class MolecularCrossSection:
    """The cross-section of one molecular species: ``sigma[iv, j, iT]`` (cm^2) at ``vmic_vals``
    (km/s), ``log_temp_vals`` (log10 K) and the wavelengths ``wls`` (a ``Wavelengths``)."""

    def __init__(self, wls, vmic_vals, log_temp_vals, sigma, species):
        self.wls = wls
        self.vmic_vals = np.ascontiguousarray(vmic_vals, dtype=np.float64)
        self.log_temp_vals = np.ascontiguousarray(log_temp_vals, dtype=np.float64)
        self.sigma = np.ascontiguousarray(sigma, dtype=np.float64)
        self.species = species
        shape = (self.vmic_vals.size, wls.all_wls.size, self.log_temp_vals.size)
        if self.sigma.shape != shape:
            raise ValueError(f"sigma has shape {self.sigma.shape}, expected {shape}")
        for knots in (self.vmic_vals, self.log_temp_vals, wls.all_wls):
            if knots.size < 2 or np.any(np.diff(knots) <= 0):
                raise ValueError("each dimension needs at least 2 ascending knots")

    @staticmethod
    def precompute_problems(linelist, wls, partition_fn, cutoff_alpha=1e-32, vmic_vals=None,
                            log_temp_vals=None):
        """The precompute's line_absorption! calls, one ``Problem`` per vmic knot
        (``src/molecular_cross_sections.jl:56-77``)."""
        vmic_vals = DEFAULT_VMIC if vmic_vals is None else np.asarray(vmic_vals, dtype=np.float64)
        log_T = DEFAULT_LOG_T if log_temp_vals is None else np.asarray(log_temp_vals, dtype=np.float64)
        species = {linelist.species[s] for s in np.unique(linelist.species_index)}
        if len(species) != 1:
            raise ValueError("All lines must be of the same species")  # :53-55
        sp = species.pop()
        ll = linelist if linelist.is_sorted() else linelist.sorted()
        wls = wls if isinstance(wls, Wavelengths) else Wavelengths(wls)
        T = 10.0 ** log_T
        L = T.size
        S = len(ll.species)
        n_div_U = np.zeros((S, L))
        for s in np.unique(ll.species_index):
            n_div_U[s] = (1.0 / cutoff_alpha) / np.asarray(partition_fn(np.log(T)), dtype=np.float64)
        nan = np.full(L, np.nan)
        # a constant continuum of 1 on two knots around the lines and the grid
        lo = min(float(wls.all_wls[0]), float(ll.wl.min())) * 0.5
        hi = max(float(wls.all_wls[-1]), float(ll.wl.max())) * 2.0
        knots = np.array([lo, hi])
        probs = [Problem(wls.all_wls, ll, [s.get_mass() for s in ll.species],
                         [s.ismolecule() for s in ll.species], n_div_U, T, np.zeros(L), nan, nan,
                         nan, np.full(L, v * 1e5), CNTM_GRID, np.ones((L, 2)), knots, 1.0)
                 for v in vmic_vals]
        return probs, sp, wls, vmic_vals, log_T

    @classmethod
    def compute(cls, linelist, wls, partition_fn, cutoff_alpha=1e-32, vmic_vals=None,
                log_temp_vals=None, ctx=None):
        """The precompute on the device (``MolecularCrossSection(linelist, wl_params; ...)``):
        one batched Float64 line-opacity call over the vmic knots."""
        probs, sp, wls, vmic_vals, log_T = cls.precompute_problems(
            linelist, wls, partition_fn, cutoff_alpha, vmic_vals, log_temp_vals)
        ctx = ctx or default_context()
        dprobs = [DeviceProblem(P) for P in probs]
        alphas = [d.new_alpha("float64") for d in dprobs]
        ctx.line_absorption_batched(alphas, dprobs, "fp64")
        sigma = np.stack([a.numpy() for a in alphas]) * cutoff_alpha  # (n_vmic, n_wl, n_logT)
        return cls(wls, vmic_vals, log_T, sigma, sp)


def _struct(table, n_species):
    keep = [table.vmic_vals, table.log_temp_vals, table.wls.all_wls, table.sigma, n_species]
    return _lib.KorgMolXsection(table.vmic_vals.size, table.log_temp_vals.size,
                                table.wls.all_wls.size, *[a.ctypes.data for a in keep]), keep


def interpolate_molecular_cross_sections_(alpha, tables, wls, temps, vmic, n_densities, ctx=None):
    """``interpolate_molecular_cross_sections!`` (``src/molecular_cross_sections.jl:99-114``) on
    the GPU: for each table, alpha[l, :] += sigma(vmic_l, log10 T_l, wls) * n_densities[species][l];
    ``vmic`` in km/s (scalar or per layer); alpha (n_layers, n_wl) Fortran-ordered float64/float32,
    updated."""
    if len(tables) == 0:
        return None
    temps = np.ascontiguousarray(temps, dtype=np.float64)
    L = temps.size
    if alpha.shape[0] != L or not alpha.flags.f_contiguous:
        raise ValueError("alpha must be (n_layers, n_wl), column-major")
    all_wls = np.ascontiguousarray(wls.all_wls if hasattr(wls, "all_wls") else wls,
                                   dtype=np.float64)
    if alpha.shape[1] != all_wls.size:
        raise ValueError("alpha must have one column per wavelength")
    dt = {np.dtype(np.float64): _lib.DTYPE_F64, np.dtype(np.float32): _lib.DTYPE_F32}
    vm = np.ascontiguousarray(np.broadcast_to(np.asarray(vmic, dtype=np.float64), (L,)))
    keep = [temps, vm, all_wls]
    structs = (_lib.KorgMolXsection * len(tables))()
    for k, t in enumerate(tables):
        n = np.ascontiguousarray(np.broadcast_to(np.asarray(n_densities[t.species],
                                                            dtype=np.float64), (L,)))
        structs[k], kk = _struct(t, n)
        keep += kk
    grid = _lib.KorgGrid(all_wls.size, all_wls.ctypes.data)
    ctx = ctx or default_context()
    rc = ctx._L.korg_molecular_cross_sections(ctx._h, ctypes.c_void_p(alpha.ctypes.data),
                                              dt[alpha.dtype], ctypes.byref(grid), L,
                                              ctypes.c_void_p(temps.ctypes.data),
                                              ctypes.c_void_p(vm.ctypes.data), len(tables),
                                              structs)
    del keep
    _lib.check(ctx._h, rc)
    return None
