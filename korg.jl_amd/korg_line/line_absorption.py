"""Host-side mirror of Korg.jl's ``line_absorption!`` over the device library.

``line_absorption_(alpha, linelist, wls, temps, n_e, n_densities, partition_fns, xi, alpha_cntm;
cutoff_threshold)`` has the argument meaning of ``src/line_absorption.jl:33-34``: it ADDS the
line opacity [cm^-1] into ``alpha`` (shape (n_layers, n_wavelengths), Julia's column-major
layout: pass a Fortran-ordered array, or the transposed view of a C-ordered one), leaves it
untouched for an empty line list, and logs an error (continuing) if H I is in the line list.

Device-resident use (inputs already in HBM) goes through ``Context.line_absorption_device`` with
a ``DeviceProblem``; its buffers are ``hipmem.DeviceArray``s.
"""
import ctypes

import numpy as np

from . import _lib
from .problem import Problem

PRECISIONS = {"fp64": _lib.PRECISION_FP64, "fp32": _lib.PRECISION_FP32_FIXUP,
              "fp32_fixup": _lib.PRECISION_FP32_FIXUP, "fp64_series": _lib.PRECISION_FP64_SERIES}


def _ptr(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


def _host_structs(P):
    ll = P.lines
    keep = [P.all_wls, ll.wl, ll.log_gf, ll.E_lower, ll.gamma_rad, ll.gamma_stark, ll.vdW1,
            ll.vdW2, ll.species_index, P.species_mass, P.species_is_mol, P.n_div_U, P.temps,
            P.n_e, P.n_HI, P.n_HeI, P.n_H2, P.xi, P.cntm_values, P.cntm_knots]
    frozen = P.window_lb is not None and P.window_ub is not None
    if frozen:
        keep += [P.window_lb, P.window_ub]
    for a in keep:
        assert a.flags.c_contiguous
    g = _lib.KorgGrid(P.n_wl, P.all_wls.ctypes.data)
    lines = _lib.KorgLines(len(ll), ll.wl.ctypes.data, ll.log_gf.ctypes.data,
                           ll.E_lower.ctypes.data, ll.gamma_rad.ctypes.data,
                           ll.gamma_stark.ctypes.data, ll.vdW1.ctypes.data, ll.vdW2.ctypes.data,
                           ll.species_index.ctypes.data,
                           P.window_lb.ctypes.data if frozen else None,
                           P.window_ub.ctypes.data if frozen else None)
    sp = _lib.KorgSpecies(P.species_mass.size, P.species_mass.ctypes.data,
                          P.species_is_mol.ctypes.data)
    ly = _lib.KorgLayers(P.n_layers, P.temps.ctypes.data, P.n_e.ctypes.data,
                         P.n_div_U.ctypes.data, P.n_HI.ctypes.data, P.n_HeI.ctypes.data,
                         P.n_H2.ctypes.data, P.xi.ctypes.data)
    cn = _lib.KorgContinuum(P.cntm_mode, P.n_knots,
                            P.cntm_knots.ctypes.data if P.n_knots else None,
                            P.cntm_values.ctypes.data)
    return (g, lines, sp, ly, cn), keep


_LINE_FIELDS = ("wl", "log_gf", "E_lower", "gamma_rad", "gamma_stark", "vdW1", "vdW2", "species")


class DeviceProblem:
    """A ``Problem`` uploaded once to device memory (HBM-resident inputs).

    ``lines``: optionally a ``korg_lines`` of device pointers already in HBM (one star's list from
    ``Context.filter_linelist_batched``) used in place of the problem's own lines, which are then
    not uploaded; the problem's line list only supplies the species table, which must be the one
    those lines index, and its continuum must be given on knots (per-line values would belong to
    the unfiltered list).  ``keep``: the object that owns that memory, held as long as this
    problem."""

    def __init__(self, problem, lines=None, keep=None):
        from .hipmem import DeviceArray
        P = problem
        ll = P.lines
        self.problem = P
        self._owner = keep
        if lines is not None and P.cntm_mode != _lib.CNTM_GRID:
            raise ValueError("device lines need a continuum on knots")
        if lines is not None and P.window_lb is not None:
            raise ValueError("device lines carry no frozen windows")
        self.t = {k: DeviceArray.from_numpy(v) for k, v in self._arrays(P).items()
                  if lines is None or k not in _LINE_FIELDS}
        t = self.t
        d = lambda k: t[k].ptr
        self.grid = _lib.KorgGrid(P.n_wl, d("all_wls"))
        frozen = "window_lb" in t
        if lines is not None:
            self.lines = lines
        else:
            self.lines = _lib.KorgLines(len(ll), d("wl"), d("log_gf"), d("E_lower"),
                                        d("gamma_rad"), d("gamma_stark"), d("vdW1"), d("vdW2"),
                                        d("species"), d("window_lb") if frozen else None,
                                        d("window_ub") if frozen else None)
        self.species = _lib.KorgSpecies(P.species_mass.size, d("mass"), d("is_mol"))
        self.layers = _lib.KorgLayers(P.n_layers, d("temps"), d("n_e"), d("n_div_U"), d("n_HI"),
                                      d("n_HeI"), d("n_H2"), d("xi"))
        self.cntm = _lib.KorgContinuum(P.cntm_mode, P.n_knots, d("knots") if P.n_knots else None,
                                       d("cntm"))

    def _arrays(self, P):
        ll = P.lines
        out = {"all_wls": P.all_wls, "wl": ll.wl, "log_gf": ll.log_gf, "E_lower": ll.E_lower,
               "gamma_rad": ll.gamma_rad, "gamma_stark": ll.gamma_stark, "vdW1": ll.vdW1,
               "vdW2": ll.vdW2, "species": ll.species_index, "mass": P.species_mass,
               "is_mol": P.species_is_mol, "n_div_U": P.n_div_U, "temps": P.temps, "n_e": P.n_e,
               "n_HI": P.n_HI, "n_HeI": P.n_HeI, "n_H2": P.n_H2, "xi": P.xi,
               "cntm": P.cntm_values, "knots": P.cntm_knots if P.n_knots else np.zeros(1)}
        if P.window_lb is not None and P.window_ub is not None:
            out["window_lb"], out["window_ub"] = P.window_lb, P.window_ub
        return out

    def update(self, problem):
        """Writes ``problem`` (same sizes and continuum mode) into this problem's device buffers,
        so that a ``Plan`` captured over them sees the new inputs at its next replay."""
        P = problem
        new = self._arrays(P)
        if (P.n_wl, P.n_layers, len(P.lines), P.cntm_mode, P.n_knots) != (
                self.problem.n_wl, self.problem.n_layers, len(self.problem.lines),
                self.problem.cntm_mode, self.problem.n_knots):
            raise ValueError("update needs a problem of the same sizes and continuum mode")
        if new.keys() != self.t.keys():
            raise ValueError("update cannot freeze or unfreeze the windows")
        for k, v in new.items():
            if v.shape != self.t[k].shape or v.dtype != self.t[k].dtype:
                raise ValueError(f"update: {k} changes shape or dtype")
        for k, v in new.items():
            self.t[k].copy_from(v)
        self.problem = P

    def new_alpha(self, dtype="float32", fill=None):
        """Device alpha in the library's layout: a C-ordered (n_wl, n_layers) buffer is the
        column-major (n_layers, n_wl) matrix.  ``fill``: optional (n_layers, n_wl) start value."""
        from .hipmem import DeviceArray
        P = self.problem
        dt = np.float32 if dtype in ("float32", "f32") else np.float64
        if fill is None:
            return DeviceArray.zeros((P.n_wl, P.n_layers), dt)
        return DeviceArray.from_numpy(np.asarray(fill, dtype=dt).T)


class DeviceLineList:
    """A wavelength-sorted ``LineList`` in device memory: the master list of
    ``Context.filter_linelist_batched``."""

    def __init__(self, linelist):
        from .hipmem import DeviceArray
        ll = linelist
        if not ll.is_sorted():
            raise ValueError("the master list must be sorted by wavelength")
        self.linelist = ll
        cols = dict(wl=ll.wl, log_gf=ll.log_gf, E_lower=ll.E_lower, gamma_rad=ll.gamma_rad,
                    gamma_stark=ll.gamma_stark, vdW1=ll.vdW1, vdW2=ll.vdW2,
                    species=ll.species_index.astype(np.int32))
        self.t = {k: DeviceArray.from_numpy(np.ascontiguousarray(v)) for k, v in cols.items()}
        self.lines = _lib.KorgLines(len(ll), *[self.t[k].ptr for k in _LINE_FIELDS], None, None)


class FilteredLines:
    """The per-star line lists of ``Context.filter_linelist_batched``: ``lines[k]`` (a
    ``korg_lines`` of device pointers into the master list or ``pool``), valid while this object
    lives (it holds the master and the pool)."""

    def __init__(self, master, lines, pool):
        self.master = master
        self.lines = lines
        self.pool = pool

    def __len__(self):
        return len(self.lines)

    def n_lines(self, k):
        return int(self.lines[k].n_lines)

    def to_host(self, k):
        """Star k's list copied back to the host, as a ``LineList`` over the master's species."""
        from .hipmem import read
        from .linelist import LineList
        L = self.lines[k]
        n = L.n_lines
        f = [read(getattr(L, name), n, np.float64) for name in _LINE_FIELDS[:-1]]
        sp = read(L.species, n, np.int32)
        return LineList(*f, sp, self.master.linelist.species)

    def problem(self, k, problem):
        """A ``DeviceProblem`` of ``problem`` (built over the master list's species table, its
        continuum on knots) with star k's filtered lines."""
        return DeviceProblem(problem, lines=self.lines[k], keep=self)


class FitTarget:
    """The fit step's target for one star, in device memory: the anchored scheme's reference
    optical depth and absorption (per layer), the data and its uncertainties (per grid point), and
    the outputs -- the model's normalized flux, chi^2 and dchi^2/d(per-layer inputs)."""

    GRADS = ("T", "n_e", "xi", "n_HI", "n_HeI", "n_H2", "n_div_U")

    def __init__(self, problem, tau_ref, alpha_ref, data, sigma):
        from .hipmem import DeviceArray
        P = problem
        L, n, S = P.n_layers, P.n_wl, P.species_mass.size
        sigma = np.broadcast_to(np.asarray(sigma, dtype=np.float64), (n,))
        self.t = {"tau_ref": DeviceArray.from_numpy(np.asarray(tau_ref, np.float64)),
                  "alpha_ref": DeviceArray.from_numpy(np.asarray(alpha_ref, np.float64)),
                  "data": DeviceArray.from_numpy(np.asarray(data, np.float64)),
                  "sigma": DeviceArray.from_numpy(np.ascontiguousarray(sigma)),
                  "flux": DeviceArray.zeros((n,), np.float64),
                  "chi2": DeviceArray.zeros((1,), np.float64)}
        self.g = {k: DeviceArray.zeros((S, L) if k == "n_div_U" else (L,), np.float64)
                  for k in self.GRADS}
        for k in ("tau_ref", "alpha_ref"):
            assert self.t[k].shape == (L,), k
        assert self.t["data"].shape == (n,)
        t = self.t
        self.struct = _lib.KorgFitTarget(t["tau_ref"].ptr, t["alpha_ref"].ptr, t["data"].ptr,
                                         t["sigma"].ptr, t["flux"].ptr, t["chi2"].ptr,
                                         _lib.KorgLayerGrads(*[self.g[k].ptr for k in self.GRADS]))

    def chi2(self):
        return float(self.t["chi2"].numpy()[0])

    def flux(self):
        return self.t["flux"].numpy()

    def grads(self):
        return {k: v.numpy() for k, v in self.g.items()}

    def set_data(self, data):
        self.t["data"].copy_from(np.asarray(data, np.float64))


class FluxTarget:
    """korg_flux_batched's target for one star, in device memory: the anchored scheme's reference
    optical depth and absorption (per layer), and the outputs -- the flux of the total absorption
    and of the continuum alone (per grid point, erg/s/cm^4/A)."""

    def __init__(self, problem, tau_ref, alpha_ref):
        from .hipmem import DeviceArray
        P = problem
        self.t = {"tau_ref": DeviceArray.from_numpy(np.asarray(tau_ref, np.float64)),
                  "alpha_ref": DeviceArray.from_numpy(np.asarray(alpha_ref, np.float64)),
                  "flux": DeviceArray.zeros((P.n_wl,), np.float64),
                  "cntm": DeviceArray.zeros((P.n_wl,), np.float64)}
        for k in ("tau_ref", "alpha_ref"):
            if self.t[k].shape != (P.n_layers,):
                raise ValueError(f"{k} has one value per layer")
        self.struct = _lib.KorgFluxTarget(*[self.t[k].ptr for k in ("tau_ref", "alpha_ref", "flux", "cntm")])

    def flux(self):
        return self.t["flux"].numpy()

    def cntm(self):
        return self.t["cntm"].numpy()


class Context:
    """One library context (HIP stream + device scratch) on one device."""

    def __init__(self, device=0):
        L = _lib.load()
        h = ctypes.c_void_p()
        rc = L.korg_create(ctypes.byref(h), int(device))
        if rc != _lib.KORG_OK:
            raise _lib.KorgError(rc, "korg_create failed (no HIP device?)")
        self._h = h
        self._L = L

    def close(self):
        if self._h:
            self._L.korg_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_tuning(self, points_per_lane=0, waves_per_cu=0, count_branches=0, tile_points=0,
                   waves_per_tile=0, near_waves_per_cu=0, near_layout=0, near_concurrent=0,
                   far_field=0):
        """Kernel tuning (0 = library default); see ``korg_tuning`` in include/korg_line.h."""
        t = _lib.KorgTuning(points_per_lane, waves_per_tile, waves_per_cu, count_branches,
                            tile_points, near_waves_per_cu, near_layout, near_concurrent,
                            far_field)
        _lib.check(self._h, self._L.korg_set_tuning(self._h, ctypes.byref(t)))

    def tuning(self):
        t = _lib.KorgTuning()
        _lib.check(self._h, self._L.korg_get_tuning(self._h, ctypes.byref(t)))
        return {k: getattr(t, k) for k, _ in t._fields_}

    def set_stream(self, stream_handle):
        _lib.check(self._h, self._L.korg_set_stream(self._h, ctypes.c_void_p(stream_handle)))

    def stats(self):
        s = _lib.KorgStats()
        _lib.check(self._h, self._L.korg_get_stats(self._h, ctypes.byref(s)))
        out = {k: getattr(s, k) for k, _ in s._fields_}
        out["n_evals_branch"] = list(s.n_evals_branch)
        return out

    def windows(self, n_lines):
        lb = np.zeros(n_lines, dtype=np.int64)
        ub = np.zeros(n_lines, dtype=np.int64)
        _lib.check(self._h, self._L.korg_get_windows(self._h, _ptr(lb), _ptr(ub), n_lines))
        return lb, ub

    def line_absorption(self, alpha, problem, precision="fp64"):
        """Host arrays: ``alpha`` float64/float32 (n_layers, n_wl), Fortran-ordered, updated."""
        if not isinstance(problem, Problem):
            raise TypeError("problem must be a korg_line.Problem")
        if alpha.shape != (problem.n_layers, problem.n_wl):
            raise ValueError(f"alpha has shape {alpha.shape}, expected "
                             f"{(problem.n_layers, problem.n_wl)}")
        if not alpha.flags.f_contiguous:
            raise ValueError("alpha must be column-major (Fortran order), like Julia's Matrix")
        dt = {np.dtype(np.float64): _lib.DTYPE_F64, np.dtype(np.float32): _lib.DTYPE_F32}
        if alpha.dtype not in dt:
            raise TypeError("alpha must be float64 or float32")
        structs, keep = _host_structs(problem)
        g, lines, sp, ly, cn = structs
        rc = self._L.korg_line_absorption(self._h, _ptr(alpha), dt[alpha.dtype], ctypes.byref(g),
                                          ctypes.byref(lines), ctypes.byref(sp), ctypes.byref(ly),
                                          ctypes.byref(cn), problem.cutoff_threshold,
                                          PRECISIONS[precision])
        del keep
        _lib.check(self._h, rc)

    def line_absorption_device(self, alpha_dev, dprob, precision="fp32", synchronize=True):
        """Device arrays: ``alpha_dev`` a ``DeviceArray`` from ``DeviceProblem.new_alpha``."""
        dt = {np.dtype(np.float64): _lib.DTYPE_F64, np.dtype(np.float32): _lib.DTYPE_F32}
        P = dprob.problem
        if tuple(alpha_dev.shape) != (P.n_wl, P.n_layers) or alpha_dev.dtype not in dt:
            raise ValueError("alpha_dev must be a (n_wl, n_layers) float32/float64 device array")
        rc = self._L.korg_line_absorption_device(
            self._h, ctypes.c_void_p(alpha_dev.ptr), dt[alpha_dev.dtype], ctypes.byref(dprob.grid),
            ctypes.byref(dprob.lines), ctypes.byref(dprob.species), ctypes.byref(dprob.layers),
            ctypes.byref(dprob.cntm), P.cutoff_threshold, PRECISIONS[precision],
            1 if synchronize else 0)
        _lib.check(self._h, rc)

    def line_absorption_device_noalloc(self, alpha_dev, dprob, precision="fp32",
                                       synchronize=True):
        """``line_absorption_device`` through the entry point that allocates no device memory
        (its scratch is the library's fixed per-GPU array of ``_lib.NOALLOC_SCRATCH_BYTES``)."""
        dt = {np.dtype(np.float64): _lib.DTYPE_F64, np.dtype(np.float32): _lib.DTYPE_F32}
        P = dprob.problem
        if tuple(alpha_dev.shape) != (P.n_wl, P.n_layers) or alpha_dev.dtype not in dt:
            raise ValueError("alpha_dev must be a (n_wl, n_layers) float32/float64 device array")
        rc = self._L.korg_line_absorption_device_noalloc(
            self._h, ctypes.c_void_p(alpha_dev.ptr), dt[alpha_dev.dtype], ctypes.byref(dprob.grid),
            ctypes.byref(dprob.lines), ctypes.byref(dprob.species), ctypes.byref(dprob.layers),
            ctypes.byref(dprob.cntm), P.cutoff_threshold, PRECISIONS[precision],
            1 if synchronize else 0)
        _lib.check(self._h, rc)

    @staticmethod
    def _stars(alphas, dprobs):
        """The korg_star array of a batch, and its alpha dtype code."""
        if len(alphas) != len(dprobs) or not dprobs:
            raise ValueError("one alpha per problem, at least one star")
        dt = {np.dtype(np.float64): _lib.DTYPE_F64, np.dtype(np.float32): _lib.DTYPE_F32}
        dtypes = {a.dtype for a in alphas}
        if len(dtypes) != 1 or next(iter(dtypes)) not in dt:
            raise TypeError("all alphas must share one dtype, float32 or float64")
        P0 = dprobs[0].problem
        for d, a in zip(dprobs, alphas):
            P = d.problem
            if not (np.array_equal(P.species_mass, P0.species_mass) and
                    np.array_equal(P.species_is_mol, P0.species_is_mol)):
                raise ValueError("every star must use the same species table")
            if tuple(a.shape) != (P.n_wl, P.n_layers):
                raise ValueError("alpha shape does not match its problem")
        stars = (_lib.KorgStar * len(dprobs))()
        for k, (d, a) in enumerate(zip(dprobs, alphas)):
            stars[k] = _lib.KorgStar(a.ptr, d.grid, d.lines, d.layers, d.cntm,
                                     d.problem.cutoff_threshold)
        return stars, dt[next(iter(dtypes))]

    def filter_linelist_batched(self, master, wls_list, line_buffer):
        """``filter_linelist`` (``src/synthesize.jl:331-350``) on the device for many stars drawn
        from one master list (a ``DeviceLineList``): star k keeps the master's lines within
        ``line_buffer`` (cm; a scalar or one per star) of each range of ``wls_list[k]`` (a
        ``Wavelengths``), each line once, in order.  Returns a ``FilteredLines``.  Warns, as the
        reference does, for a star that keeps nothing of a non-empty master."""
        import warnings
        from .hipmem import DeviceArray
        n = len(wls_list)
        buf = np.broadcast_to(np.asarray(line_buffer, dtype=np.float64), (n,))
        ranges = []
        stars = (_lib.KorgFilterStar * max(n, 1))()
        for k, w in enumerate(wls_list):
            r = np.ascontiguousarray(np.array(w.eachwindow(), dtype=np.float64).ravel())
            ranges.append(r)
            stars[k] = _lib.KorgFilterStar(r.size // 2, r.ctypes.data, float(buf[k]))
        out = (_lib.KorgLines * max(n, 1))()
        need = ctypes.c_int64(0)
        rc = self._L.korg_filter_linelist_batched(self._h, ctypes.byref(master.lines), n, stars,
                                                  None, 0, out, ctypes.byref(need))
        pool = None
        if rc == _lib.KORG_ERR_CAPACITY:
            pool = DeviceArray((_lib.filter_pool_bytes(need.value),), np.uint8)
            rc = self._L.korg_filter_linelist_batched(self._h, ctypes.byref(master.lines), n,
                                                      stars, ctypes.c_void_p(pool.ptr),
                                                      need.value, out, ctypes.byref(need))
        _lib.check(self._h, rc)
        del ranges
        lines = [out[k] for k in range(n)]
        if len(master.linelist) and any(L.n_lines == 0 for L in lines):
            warnings.warn("The provided linelist was not empty, but none of the lines were "
                          "within the provided wavelength range.")
        return FilteredLines(master, lines, pool)

    def line_absorption_batched(self, alphas, dprobs, precision="fp32", synchronize=True):
        """Many stars in one pipeline: ``alphas[k]`` (a ``DeviceProblem.new_alpha`` array) receives
        the line opacity of ``dprobs[k]``.  Every problem must share one species table and one
        layer count; all alphas have one dtype."""
        stars, dtc = self._stars(alphas, dprobs)
        rc = self._L.korg_line_absorption_batched(self._h, len(dprobs), stars,
                                                  ctypes.byref(dprobs[0].species), dtc,
                                                  PRECISIONS[precision], 1 if synchronize else 0)
        _lib.check(self._h, rc)

    def plan_batched(self, alphas, dprobs, precision="fp32", headroom=0.0):
        """A ``Plan``: the batched call captured in a HIP graph for these device buffers."""
        return Plan(self, alphas, dprobs, precision, headroom)

    def fit_gradient_batched(self, alphas, dprobs, targets, precision="fp32", synchronize=True):
        """The fit step (korg_fit_gradient_batched): alphas[k] receives star k's line opacity;
        targets[k] (a ``FitTarget``) its normalized flux, chi^2 and gradient."""
        stars, dtc = self._stars(alphas, dprobs)
        tg = self._targets(targets, dprobs)
        rc = self._L.korg_fit_gradient_batched(self._h, len(dprobs), stars, tg,
                                               ctypes.byref(dprobs[0].species), dtc,
                                               PRECISIONS[precision], 1 if synchronize else 0)
        _lib.check(self._h, rc)

    def plan_fit(self, alphas, dprobs, targets, precision="fp32", headroom=0.0):
        """A ``Plan`` of the fit step (korg_plan_fit)."""
        return Plan(self, alphas, dprobs, precision, headroom, targets=targets)

    @staticmethod
    def _targets(targets, dprobs):
        if len(targets) != len(dprobs):
            raise ValueError("one FitTarget per star")
        tg = (_lib.KorgFitTarget * len(targets))()
        for k, t in enumerate(targets):
            tg[k] = t.struct
        return tg

    def continuum_batched(self, alphas, dprobs, synchronize=True):
        """korg_continuum_batched: alphas[k] set to star k's continuum on its grid (the problems'
        continuum on knots), the alpha that synthesize starts the lines from."""
        stars, dtc = self._stars(alphas, dprobs)
        rc = self._L.korg_continuum_batched(self._h, len(dprobs), stars, dtc, 1 if synchronize else 0)
        _lib.check(self._h, rc)

    def flux_batched(self, alphas, dprobs, targets, synchronize=True):
        """korg_flux_batched: the emergent flux of alphas[k] taken as star k's total absorption,
        and of its continuum alone (erg/s/cm^4/A, as synthesize), into targets[k] (a
        ``FluxTarget``)."""
        stars, dtc = self._stars(alphas, dprobs)
        if len(targets) != len(dprobs):
            raise ValueError("one FluxTarget per star")
        tg = (_lib.KorgFluxTarget * len(targets))()
        for k, t in enumerate(targets):
            tg[k] = t.struct
        rc = self._L.korg_flux_batched(self._h, len(dprobs), stars, tg, dtc, 1 if synchronize else 0)
        _lib.check(self._h, rc)

    def set_hydrogen_tables(self, tables=None):
        """korg_set_hydrogen_tables: the Stark tables (a ``hydrogen.StarkTables``; default: the
        shipped file) to the device, once per context."""
        from .hydrogen import StarkTables
        t = tables if tables is not None else StarkTables()
        _lib.check(self._h, self._L.korg_set_hydrogen_tables(self._h, ctypes.byref(t.struct)))
        self._htables = t

    def hydrogen_line_absorption(self, alpha, dprob, U_HI, window_size=150e-8, use_MHD=None,
                                 synchronize=True):
        """korg_hydrogen_line_absorption: synthesize's hydrogen lines (src/synthesize.jl:293-301)
        added into ``alpha`` (a device alpha of ``dprob``) with the problem's grid and layers;
        ``U_HI``: the H I partition function at each layer's temperature; ``use_MHD`` defaults as
        synthesize does (grids ending below 13,000 A, src/synthesize.jl:166-168)."""
        from .hipmem import DeviceArray
        if getattr(self, "_htables", None) is None:
            self.set_hydrogen_tables()
        P = dprob.problem
        if use_MHD is None:
            use_MHD = bool(P.all_wls[-1] < 13_000e-8)
        U = np.ascontiguousarray(U_HI, dtype=np.float64)
        if U.shape != (P.n_layers,):
            raise ValueError(f"U_HI must have one value per layer ({P.n_layers})")
        self._hl_U = DeviceArray.from_numpy(U)  # held until the next call
        t = dprob.t
        lay = _lib.KorgHLayers(P.n_layers, t["temps"].ptr, t["n_e"].ptr, t["n_HI"].ptr,
                               t["n_HeI"].ptr, self._hl_U.ptr, t["xi"].ptr)
        dt = _lib.DTYPE_F32 if alpha.dtype == np.float32 else _lib.DTYPE_F64
        rc = self._L.korg_hydrogen_line_absorption(self._h, alpha.ptr, dt, ctypes.byref(dprob.grid),
                                                   ctypes.byref(lay), float(window_size),
                                                   1 if use_MHD else 0, 1 if synchronize else 0)
        _lib.check(self._h, rc)

    def autotune_batched(self, alphas, dprobs, precision="fp32", cache_path=None):
        """Times the tuning candidates on these stars (into scratch alphas), makes the fastest
        this context's tuning and returns it; ``cache_path``: optional results file."""
        stars, dtc = self._stars(alphas, dprobs)
        t = _lib.KorgTuning()
        rc = self._L.korg_autotune_batched(self._h, len(dprobs), stars,
                                           ctypes.byref(dprobs[0].species), dtc,
                                           PRECISIONS[precision],
                                           None if cache_path is None else str(cache_path).encode(),
                                           ctypes.byref(t))
        _lib.check(self._h, rc)
        return {k: getattr(t, k) for k, _ in t._fields_}

    def line_absorption_vjp(self, G, problem):
        """Adjoint on host arrays: given ``G`` = dL/dalpha (float64/float32, (n_layers, n_wl),
        Fortran-ordered), returns dL/dx for the per-layer inputs: a dict with ``T``, ``n_e``,
        ``xi``, ``n_HI``, ``n_HeI``, ``n_H2`` (length n_layers) and ``n_div_U``
        (n_species x n_layers).  Windows are held fixed; Float64; deterministic."""
        if G.shape != (problem.n_layers, problem.n_wl):
            raise ValueError(f"G has shape {G.shape}, expected {(problem.n_layers, problem.n_wl)}")
        if not G.flags.f_contiguous:
            raise ValueError("G must be column-major (Fortran order), like alpha")
        dt = {np.dtype(np.float64): _lib.DTYPE_F64, np.dtype(np.float32): _lib.DTYPE_F32}
        if G.dtype not in dt:
            raise TypeError("G must be float64 or float32")
        L, S = problem.n_layers, problem.species_mass.size
        out = {k: np.zeros(L) for k in ("T", "n_e", "xi", "n_HI", "n_HeI", "n_H2")}
        out["n_div_U"] = np.zeros((S, L))
        grads = _lib.KorgLayerGrads(*[out[k].ctypes.data for k in
                                      ("T", "n_e", "xi", "n_HI", "n_HeI", "n_H2", "n_div_U")])
        structs, keep = _host_structs(problem)
        g, lines, sp, ly, cn = structs
        rc = self._L.korg_line_absorption_vjp(self._h, _ptr(G), dt[G.dtype], ctypes.byref(g),
                                              ctypes.byref(lines), ctypes.byref(sp),
                                              ctypes.byref(ly), ctypes.byref(cn),
                                              problem.cutoff_threshold, ctypes.byref(grads))
        del keep
        _lib.check(self._h, rc)
        return out

    def line_absorption_vjp_device(self, G_dev, dprob, out_dev, synchronize=True):
        """Adjoint on device arrays: ``G_dev`` a (n_wl, n_layers) device array (the layout of
        ``DeviceProblem.new_alpha``); ``out_dev`` a dict of device arrays keyed as the host form's
        result (missing keys are not written)."""
        dt = {np.dtype(np.float64): _lib.DTYPE_F64, np.dtype(np.float32): _lib.DTYPE_F32}
        P = dprob.problem
        if tuple(G_dev.shape) != (P.n_wl, P.n_layers) or G_dev.dtype not in dt:
            raise ValueError("G_dev must be a (n_wl, n_layers) float32/float64 device array")
        names = ("T", "n_e", "xi", "n_HI", "n_HeI", "n_H2", "n_div_U")
        grads = _lib.KorgLayerGrads(*[out_dev[k].ptr if k in out_dev else None for k in names])
        rc = self._L.korg_line_absorption_vjp_device(
            self._h, ctypes.c_void_p(G_dev.ptr), dt[G_dev.dtype], ctypes.byref(dprob.grid),
            ctypes.byref(dprob.lines), ctypes.byref(dprob.species), ctypes.byref(dprob.layers),
            ctypes.byref(dprob.cntm), P.cutoff_threshold, ctypes.byref(grads),
            1 if synchronize else 0)
        _lib.check(self._h, rc)

    def line_absorption_jvp(self, problem, tangents, dalpha=None, dtype=np.float64):
        """Forward mode on host arrays: ``tangents`` is a list of K dicts of per-layer input
        tangents (``T``, ``n_e``, ``xi``, ``n_HI``, ``n_HeI``, ``n_H2``: length n_layers;
        ``n_div_U``: (n_species, n_layers); a missing key is a zero tangent).  Adds J x_k into
        ``dalpha[k]`` (a list of K (n_layers, n_wl) Fortran-ordered arrays; zeros when not
        given) and returns the list.  Windows held fixed; Float64; deterministic."""
        P = problem
        L, S, K = P.n_layers, P.species_mass.size, len(tangents)
        if dalpha is None:
            dalpha = [np.zeros((L, P.n_wl), dtype=dtype, order="F") for _ in range(K)]
        if len(dalpha) != K:
            raise ValueError("one dalpha per tangent")
        dt = {np.dtype(np.float64): _lib.DTYPE_F64, np.dtype(np.float32): _lib.DTYPE_F32}
        for a in dalpha:
            if a.shape != (L, P.n_wl) or not a.flags.f_contiguous or a.dtype != dalpha[0].dtype \
                    or a.dtype not in dt:
                raise ValueError("dalpha: (n_layers, n_wl) Fortran-ordered float64/float32 arrays "
                                 "of one dtype")
        names = ("T", "n_e", "xi", "n_HI", "n_HeI", "n_H2", "n_div_U")
        keep = []

        def arr(t, k):
            if k not in t:
                return None
            a = np.ascontiguousarray(np.broadcast_to(np.asarray(t[k], dtype=np.float64),
                                                     (S, L) if k == "n_div_U" else (L,)))
            keep.append(a)
            return a.ctypes.data
        tv = (_lib.KorgLayerTangents * max(K, 1))(*[_lib.KorgLayerTangents(*[arr(t, k) for k in names])
                                                    for t in tangents])
        ptrs = (ctypes.c_void_p * max(K, 1))(*[a.ctypes.data for a in dalpha])
        structs, keep2 = _host_structs(P)
        g, lines, sp, ly, cn = structs
        rc = self._L.korg_line_absorption_jvp(
            self._h, ptrs, dt[dalpha[0].dtype] if K else _lib.DTYPE_F64, ctypes.byref(g),
            ctypes.byref(lines), ctypes.byref(sp), ctypes.byref(ly), ctypes.byref(cn),
            P.cutoff_threshold, K, tv)
        del keep, keep2
        _lib.check(self._h, rc)
        return dalpha

    def line_absorption_jvp_device(self, dprob, tangents_dev, dalpha_dev, synchronize=True):
        """Forward mode on device arrays: ``tangents_dev`` a list of dicts of device arrays
        (keys as ``line_absorption_jvp``), ``dalpha_dev`` a list of (n_wl, n_layers) device arrays
        of one dtype (the layout of ``DeviceProblem.new_alpha``), added into."""
        P = dprob.problem
        K = len(tangents_dev)
        dt = {np.dtype(np.float64): _lib.DTYPE_F64, np.dtype(np.float32): _lib.DTYPE_F32}
        if len(dalpha_dev) != K:
            raise ValueError("one dalpha per tangent")
        for a in dalpha_dev:
            if tuple(a.shape) != (P.n_wl, P.n_layers) or a.dtype != dalpha_dev[0].dtype:
                raise ValueError("dalpha_dev: (n_wl, n_layers) device arrays of one dtype")
        names = ("T", "n_e", "xi", "n_HI", "n_HeI", "n_H2", "n_div_U")
        tv = (_lib.KorgLayerTangents * max(K, 1))(
            *[_lib.KorgLayerTangents(*[t[k].ptr if k in t else None for k in names])
              for t in tangents_dev])
        ptrs = (ctypes.c_void_p * max(K, 1))(*[a.ptr for a in dalpha_dev])
        rc = self._L.korg_line_absorption_jvp_device(
            self._h, ptrs, dt[dalpha_dev[0].dtype] if K else _lib.DTYPE_F64,
            ctypes.byref(dprob.grid), ctypes.byref(dprob.lines), ctypes.byref(dprob.species),
            ctypes.byref(dprob.layers), ctypes.byref(dprob.cntm), P.cutoff_threshold, K, tv,
            1 if synchronize else 0)
        _lib.check(self._h, rc)

    def synchronize(self):
        _lib.check(self._h, self._L.korg_synchronize(self._h))


_default_ctx = None


class Plan:
    """A batched forward call captured once in a HIP graph (``korg_plan_batched``) and replayed
    by ``launch``.  Holds its alphas, problems and context, whose device buffers the graph uses;
    new inputs are written into those buffers between replays."""

    def __init__(self, ctx, alphas, dprobs, precision="fp32", headroom=0.0, targets=None):
        stars, dtc = ctx._stars(alphas, dprobs)
        self.ctx, self.alphas, self.dprobs = ctx, list(alphas), list(dprobs)
        self.targets = None if targets is None else list(targets)
        h = ctypes.c_void_p()
        if targets is None:
            rc = ctx._L.korg_plan_batched(ctx._h, len(dprobs), stars,
                                          ctypes.byref(dprobs[0].species), dtc,
                                          PRECISIONS[precision], float(headroom), ctypes.byref(h))
        else:
            rc = ctx._L.korg_plan_fit(ctx._h, len(dprobs), stars, ctx._targets(targets, dprobs),
                                      ctypes.byref(dprobs[0].species), dtc, PRECISIONS[precision],
                                      float(headroom), ctypes.byref(h))
        _lib.check(ctx._h, rc)
        self._p = h

    def launch(self, synchronize=True):
        """Replays the graph (adds into the alphas); with synchronize, waits and reads the counters."""
        _lib.check(self.ctx._h, self.ctx._L.korg_plan_launch(self._p, 1 if synchronize else 0))

    def wait(self):
        _lib.check(self.ctx._h, self.ctx._L.korg_plan_wait(self._p))

    def info(self):
        i = _lib.KorgPlanInfo()
        _lib.check(self.ctx._h, self.ctx._L.korg_plan_get_info(self._p, ctypes.byref(i)))
        out = {k: getattr(i, k) for k, _ in i._fields_ if k not in ("pad", "tuning")}
        out["tuning"] = {k: getattr(i.tuning, k) for k, _ in i.tuning._fields_}
        return out

    def close(self):
        if getattr(self, "_p", None):
            self.ctx._L.korg_plan_destroy(self._p)
            self._p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def default_context():
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


def line_absorption_(alpha, linelist, wls, temps, n_e, n_densities, partition_fns, xi,
                     alpha_cntm, cutoff_threshold=3e-4, precision="fp64", ctx=None):
    """``line_absorption!`` (``src/line_absorption.jl:33-129``) on the GPU; alpha is updated.

    ``precision``: "fp64" (the reference's Float64 arithmetic) or "fp32" (Float32 profiles
    with Float64 fix-ups).  Unsorted line lists are sorted by wavelength first, as
    ``synthesize`` does (``src/synthesize.jl:191-195``).
    """
    if len(linelist) == 0:
        return None  # src/line_absorption.jl:35-37
    if not linelist.is_sorted():
        linelist = linelist.sorted()
    P = Problem.from_korg_args(linelist, wls, temps, n_e, n_densities, partition_fns, xi,
                               alpha_cntm, cutoff_threshold)
    (ctx or default_context()).line_absorption(alpha, P, precision)
    return None


def line_absorption_vjp_(G, linelist, wls, temps, n_e, n_densities, partition_fns, xi,
                         alpha_cntm, cutoff_threshold=3e-4, ctx=None):
    """Reverse-mode derivative of ``line_absorption_`` with respect to its per-layer arguments,
    given ``G`` = dL/dalpha (shape (n_layers, n_wl), column-major): the pullback a
    ChainRulesCore ``rrule`` of the line opacity returns.

    Returns a dict:

    * ``temps``: dL/dT per layer, including the partition functions' dependence on T
      (d(n/U)/dT = -n U'(ln T) / (U^2 T), U' by a central difference of the callable);
    * ``n_e`` and ``xi`` per layer (``xi`` summed to a scalar when ``xi`` is a scalar);
    * ``n_densities``: {species: dL/dn per layer} -- the line species through n/U, and H I,
      He I and H2 through the van der Waals perturber density;
    * ``n_div_U``: the raw derivative with respect to n/U (species x layers, line-list order).

    The windows are held fixed (integer functions of the inputs); the adjoint runs in Float64.
    """
    if not linelist.is_sorted():
        linelist = linelist.sorted()
    P = Problem.from_korg_args(linelist, wls, temps, n_e, n_densities, partition_fns, xi,
                               alpha_cntm, cutoff_threshold)
    G = np.asfortranarray(G)
    raw = (ctx or default_context()).line_absorption_vjp(G, P)
    temps = np.asarray(temps, dtype=np.float64)
    lnT = np.log(temps)
    dT = raw["T"].copy()
    dens = {}
    used = np.unique(linelist.species_index) if len(linelist) else []
    for s in used:
        sp = linelist.species[s]
        n = np.asarray(n_densities[sp], dtype=np.float64)
        U = np.asarray(partition_fns[sp](lnT), dtype=np.float64)
        h = 1e-6
        dU = (np.asarray(partition_fns[sp](lnT + h)) - np.asarray(partition_fns[sp](lnT - h))) / (2 * h)
        dens[sp] = raw["n_div_U"][s] / U
        dT += raw["n_div_U"][s] * (-n * dU / (U * U * temps))
    from .species import Species
    for code, key in (("H I", "n_HI"), ("He I", "n_HeI"), ("H2", "n_H2")):
        sp = Species(code)
        if sp in n_densities:
            dens[sp] = dens.get(sp, 0.0) + raw[key]
    dxi = raw["xi"] if np.ndim(xi) else float(np.sum(raw["xi"]))
    return {"temps": dT, "n_e": raw["n_e"], "xi": dxi, "n_densities": dens,
            "n_div_U": raw["n_div_U"]}



def line_absorption_jvp_(dalphas, linelist, wls, temps, n_e, n_densities, partition_fns, xi,
                         alpha_cntm, tangents, cutoff_threshold=3e-4, ctx=None):
    """Forward-mode derivative of ``line_absorption_``: what the partials of a ForwardDiff.Dual
    alpha receive (``src/synthesize.jl:242-247, 305-306`` under ``Fit``).

    ``tangents`` is a list of K dicts, each a direction of the per-layer arguments: ``temps``,
    ``n_e``, ``xi`` (arrays or scalars) and ``n_densities`` ({species: per-layer tangent});
    missing entries are zero.  For each k, adds J x_k into ``dalphas[k]`` (column-major
    (n_layers, n_wl) arrays).  n/U moves with both the densities and the partition functions:
    d(n/U) = dn/U - n U'(ln T) dT / (U^2 T), U' by a central difference of the callable, as the
    reverse mode.  Windows held fixed; Float64.
    """
    if len(linelist) == 0 or len(tangents) == 0:
        return None  # src/line_absorption.jl:35-37
    if not linelist.is_sorted():
        linelist = linelist.sorted()
    P = Problem.from_korg_args(linelist, wls, temps, n_e, n_densities, partition_fns, xi,
                               alpha_cntm, cutoff_threshold)
    from .species import Species
    L = P.n_layers
    temps = np.asarray(temps, dtype=np.float64)
    lnT = np.log(temps)
    used = np.unique(linelist.species_index)
    lay = lambda x: np.broadcast_to(np.asarray(x, dtype=np.float64), (L,)).copy()
    xs = []
    for t in tangents:
        dT = lay(t.get("temps", 0.0))
        dn = t.get("n_densities", {})
        dnU = np.zeros_like(P.n_div_U)
        for s in used:
            sp = linelist.species[s]
            n = np.asarray(n_densities[sp], dtype=np.float64)
            U = np.asarray(partition_fns[sp](lnT), dtype=np.float64)
            h = 1e-6
            dU = (np.asarray(partition_fns[sp](lnT + h)) -
                  np.asarray(partition_fns[sp](lnT - h))) / (2 * h)
            dnU[s] = lay(dn.get(sp, 0.0)) / U - n * dU * dT / (U * U * temps)
        x = {"T": dT, "n_e": lay(t.get("n_e", 0.0)), "xi": lay(t.get("xi", 0.0)),
             "n_div_U": dnU}
        for code, key in (("H I", "n_HI"), ("He I", "n_HeI"), ("H2", "n_H2")):
            if Species(code) in dn:
                x[key] = lay(dn[Species(code)])
        xs.append(x)
    (ctx or default_context()).line_absorption_jvp(P, xs, dalphas)
    return None
