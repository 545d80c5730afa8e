"""ORACLE -- test infrastructure only: ctypes access to the Float64 CPU restatement of Korg's
``line_absorption!`` (``line_oracle.c``; every function there cites the reference file:line it
follows).

Only the tests, the smoke check and the benchmark's CPU-baseline leg use this module, as the
checker of the device library; the product never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

DETERMINISTIC = 0
KORG_CHUNKS = 1
WINDOWS_ONLY = 2
CNTM_AT_LINES = 0
CNTM_GRID = 1

_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", HERE, "liboracle.so"])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        d, i64 = ctypes.c_double, ctypes.c_int64
        for name, args in [("oracle_voigt_hjerting", [d, d]),
                           ("oracle_line_profile", [d, d, d, d, d]),
                           ("oracle_inverse_gaussian_density", [d, d]),
                           ("oracle_inverse_lorentz_density", [d, d]),
                           ("oracle_doppler_width", [d, d, d, d]),
                           ("oracle_scaled_stark", [d, d]),
                           ("oracle_scaled_vdW", [d, d, d, d]),
                           ("oracle_sigma_line", [d])]:
            f = getattr(L, name)
            f.argtypes = args
            f.restype = d
        L.oracle_harris_series.argtypes = [d, ctypes.POINTER(d), ctypes.POINTER(d)]
        L.oracle_harris_series.restype = d
        for name in ("oracle_searchsorted_first", "oracle_searchsorted_last"):
            f = getattr(L, name)
            f.argtypes = [ctypes.c_void_p, i64, d]
            f.restype = i64
        L.oracle_line_absorption.restype = ctypes.c_int
        L.oracle_voigt_hjerting_grad.argtypes = [d, d, ctypes.POINTER(d), ctypes.POINTER(d)]
        L.oracle_voigt_hjerting_grad.restype = d
        L.oracle_line_absorption_vjp.restype = ctypes.c_int
        L.oracle_line_absorption_jvp.restype = ctypes.c_int
        L.oracle_blackbody.argtypes = [d, d]
        L.oracle_blackbody.restype = d
        L.oracle_expint2.argtypes = [d, ctypes.POINTER(d)]
        L.oracle_expint2.restype = d
        L.oracle_transfer_core.argtypes = [d, d, d, ctypes.POINTER(d), ctypes.POINTER(d),
                                           ctypes.POINTER(d)]
        L.oracle_transfer_core.restype = d
        vp = ctypes.c_void_p
        L.oracle_rt_flux.argtypes = [i64, i64, vp, vp, vp, vp, vp, vp]
        L.oracle_rt_flux.restype = ctypes.c_int
        L.oracle_rt_flux_vjp.argtypes = [i64, i64, vp, vp, vp, vp, vp, vp, vp]
        L.oracle_rt_flux_vjp.restype = ctypes.c_int
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def voigt_hjerting(a, v):
    return lib().oracle_voigt_hjerting(float(a), float(v))


def line_profile(lambda0, sigma, gamma, amplitude, lam):
    return lib().oracle_line_profile(float(lambda0), float(sigma), float(gamma), float(amplitude),
                                     float(lam))


def harris_series(v):
    h1, h2 = ctypes.c_double(), ctypes.c_double()
    h0 = lib().oracle_harris_series(float(v), ctypes.byref(h1), ctypes.byref(h2))
    return h0, h1.value, h2.value


def inverse_gaussian_density(rho, sigma):
    return lib().oracle_inverse_gaussian_density(float(rho), float(sigma))


def inverse_lorentz_density(rho, gamma):
    return lib().oracle_inverse_lorentz_density(float(rho), float(gamma))


def doppler_width(lambda0, T, m, xi):
    return lib().oracle_doppler_width(float(lambda0), float(T), float(m), float(xi))


def scaled_stark(gamma_stark, T):
    return lib().oracle_scaled_stark(float(gamma_stark), float(T))


def scaled_vdW(vdW, m, T):
    return lib().oracle_scaled_vdW(float(vdW[0]), float(vdW[1]), float(m), float(T))


def sigma_line(lam):
    return lib().oracle_sigma_line(float(lam))


def searchsortedfirst(all_wls, lam):
    a = np.ascontiguousarray(all_wls, dtype=np.float64)
    return lib().oracle_searchsorted_first(_p(a), a.size, float(lam))


def searchsortedlast(all_wls, lam):
    a = np.ascontiguousarray(all_wls, dtype=np.float64)
    return lib().oracle_searchsorted_last(_p(a), a.size, float(lam))


def _frozen(P):
    """The problem's frozen windows (1-based int64 arrays) or (None, None)."""
    lb, ub = getattr(P, "window_lb", None), getattr(P, "window_ub", None)
    if lb is None or ub is None:
        return None, None
    return (np.ascontiguousarray(lb, dtype=np.int64), np.ascontiguousarray(ub, dtype=np.int64))


def line_absorption(alpha, problem, nthreads=0, mode=DETERMINISTIC, return_windows=False):
    """Accumulate the lines of ``problem`` (a ``korg_line.problem.Problem``) into ``alpha``
    (float64, shape (L, n_wl), Fortran order == the reference's column-major layout).  The
    problem's frozen windows (``window_lb``/``window_ub``), when set, replace the computed ones.

    Returns ``(lb, ub, n_evals)`` (1-based windows) when ``return_windows``.
    """
    assert alpha.dtype == np.float64 and alpha.flags.f_contiguous
    P = problem
    L, nwl = alpha.shape
    n = len(P.lines)
    lb = np.zeros(n, dtype=np.int64)
    ub = np.zeros(n, dtype=np.int64)
    nev = ctypes.c_double(0.0)
    ll = P.lines
    rc = lib().oracle_line_absorption(
        _p(alpha), ctypes.c_int64(L), ctypes.c_int64(nwl), _p(P.all_wls), ctypes.c_int64(n),
        _p(ll.wl), _p(ll.log_gf), _p(ll.E_lower), _p(ll.gamma_rad), _p(ll.gamma_stark),
        _p(ll.vdW1), _p(ll.vdW2), _p(ll.species_index), _p(P.species_mass), _p(P.species_is_mol),
        _p(P.n_div_U), _p(P.temps), _p(P.n_e), _p(P.n_HI), _p(P.n_HeI), _p(P.n_H2), _p(P.xi),
        ctypes.c_int(P.cntm_mode), _p(P.cntm_values), ctypes.c_int64(P.n_knots),
        _p(P.cntm_knots), ctypes.c_double(P.cutoff_threshold), ctypes.c_int(nthreads),
        ctypes.c_int(mode), _p(lb), _p(ub), ctypes.byref(nev), *map(_p, _frozen(P)))
    if rc != 0:
        raise RuntimeError(f"oracle_line_absorption returned {rc}")
    if return_windows:
        return lb, ub, nev.value
    return None


def voigt_hjerting_grad(a, v):
    """(H, dH/da, dH/dv) of the branch of voigt_hjerting chosen at (a, v)."""
    da, dv = ctypes.c_double(), ctypes.c_double()
    h = lib().oracle_voigt_hjerting_grad(a, v, ctypes.byref(da), ctypes.byref(dv))
    return h, da.value, dv.value


GRAD_NAMES = ("T", "n_e", "xi", "n_HI", "n_HeI", "n_H2")


def line_absorption_vjp(G, problem, nthreads=0):
    """dL/d(per-layer inputs) given G = dL/dalpha (float64, shape (L, n_wl), Fortran order).

    Returns a dict: ``T``, ``n_e``, ``xi``, ``n_HI``, ``n_HeI``, ``n_H2`` (length L each) and
    ``n_div_U`` (n_species x L), the derivative with respect to n_s / U_s of each layer.
    """
    P = problem
    G = np.asfortranarray(G, dtype=np.float64)
    L, nwl = G.shape
    assert (L, nwl) == (P.n_layers, P.n_wl)
    n = len(P.lines)
    S = len(P.species_mass)
    out = {k: np.zeros(L) for k in GRAD_NAMES}
    out["n_div_U"] = np.zeros((S, L))
    ll = P.lines
    rc = lib().oracle_line_absorption_vjp(
        _p(G), ctypes.c_int64(L), ctypes.c_int64(nwl), _p(P.all_wls), ctypes.c_int64(n),
        _p(ll.wl), _p(ll.log_gf), _p(ll.E_lower), _p(ll.gamma_rad), _p(ll.gamma_stark),
        _p(ll.vdW1), _p(ll.vdW2), _p(ll.species_index), _p(P.species_mass), _p(P.species_is_mol),
        ctypes.c_int64(S), _p(P.n_div_U), _p(P.temps), _p(P.n_e), _p(P.n_HI), _p(P.n_HeI),
        _p(P.n_H2), _p(P.xi), ctypes.c_int(P.cntm_mode), _p(P.cntm_values),
        ctypes.c_int64(P.n_knots), _p(P.cntm_knots), ctypes.c_double(P.cutoff_threshold),
        ctypes.c_int(nthreads), *[_p(out[k]) for k in GRAD_NAMES], _p(out["n_div_U"]),
        *map(_p, _frozen(P)))
    if rc != 0:
        raise RuntimeError(f"oracle_line_absorption_vjp returned {rc}")
    return out


def line_absorption_jvp(problem, tangents, nthreads=0):
    """Jacobian-vector products of the line opacity: ``tangents`` is a list of K dicts, each
    giving tangents of the per-layer inputs (``T``, ``n_e``, ``xi``, ``n_HI``, ``n_HeI``,
    ``n_H2``: length L; ``n_div_U``: (n_species, L); a missing key is a zero tangent).  Returns a
    (K, L, n_wl) float64 array: dalpha_k = J x_k, windows held fixed (frozen ones if set)."""
    P = problem
    L, nwl, S = P.n_layers, P.n_wl, len(P.species_mass)
    K = len(tangents)
    n = len(P.lines)
    # each direction's alpha is column-major (L, n_wl): a C-ordered (K, n_wl, L) array
    out = np.zeros((K, nwl, L))
    if K == 0:
        return out.transpose(0, 2, 1)

    def stack(name, shape):
        if all(name not in t for t in tangents):
            return None
        return np.ascontiguousarray(np.stack([np.broadcast_to(np.asarray(t.get(name, 0.0),
                                                                         dtype=np.float64), shape)
                                              for t in tangents]))
    tv = [stack(k, (L,)) for k in GRAD_NAMES] + [stack("n_div_U", (S, L))]
    ll = P.lines
    rc = lib().oracle_line_absorption_jvp(
        _p(out), ctypes.c_int64(K), ctypes.c_int64(L), ctypes.c_int64(nwl), _p(P.all_wls),
        ctypes.c_int64(n), _p(ll.wl), _p(ll.log_gf), _p(ll.E_lower), _p(ll.gamma_rad),
        _p(ll.gamma_stark), _p(ll.vdW1), _p(ll.vdW2), _p(ll.species_index), _p(P.species_mass),
        _p(P.species_is_mol), ctypes.c_int64(S), _p(P.n_div_U), _p(P.temps), _p(P.n_e),
        _p(P.n_HI), _p(P.n_HeI), _p(P.n_H2), _p(P.xi), ctypes.c_int(P.cntm_mode),
        _p(P.cntm_values), ctypes.c_int64(P.n_knots), _p(P.cntm_knots),
        ctypes.c_double(P.cutoff_threshold), ctypes.c_int(nthreads), *map(_p, tv),
        *map(_p, _frozen(P)))
    if rc != 0:
        raise RuntimeError(f"oracle_line_absorption_jvp returned {rc}")
    return out.transpose(0, 2, 1)



# ---------------------------------------------------------------------------------------------
# radiative transfer (rt_oracle.c): src/RadiativeTransfer/RadiativeTransfer.jl, src/synthesize.jl


def blackbody(T, lam):
    return lib().oracle_blackbody(float(T), float(lam))


def expint2(x):
    """Korg's approximate E2(x) and its derivative (the derivative of the approximation)."""
    dE = ctypes.c_double()
    E = lib().oracle_expint2(float(x), ctypes.byref(dE))
    return E, dE.value


def transfer_core(tau, m, b):
    """expint_transfer_integral_core(tau, m, b) and its partials (d/dtau, d/dm, d/db)."""
    dt, dm, db = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    v = lib().oracle_transfer_core(float(tau), float(m), float(b), ctypes.byref(dt),
                                   ctypes.byref(dm), ctypes.byref(db))
    return v, dt.value, dm.value, db.value


def _rt_args(alpha, wls, temps, tau_ref, alpha_ref):
    alpha = np.asfortranarray(alpha, dtype=np.float64)  # (L, nwl), column-major
    c = lambda x: np.ascontiguousarray(x, dtype=np.float64)
    return alpha, c(wls), c(temps), c(tau_ref), c(alpha_ref)


def rt_flux(alpha, wls, temps, tau_ref, alpha_ref):
    """Astrophysical flux F[j] (plane-parallel, anchored tau, linear source, expint form) of
    alpha (n_layers, n_wl) with the blackbody source of each layer at wls (cm)."""
    alpha, wls, temps, tau_ref, alpha_ref = _rt_args(alpha, wls, temps, tau_ref, alpha_ref)
    F = np.zeros(wls.size)
    rc = lib().oracle_rt_flux(alpha.shape[0], wls.size, _p(alpha), _p(wls), _p(temps),
                              _p(tau_ref), _p(alpha_ref), _p(F))
    assert rc == 0
    return F


def rt_flux_vjp(alpha, wls, temps, tau_ref, alpha_ref, gF):
    """dL/dalpha (n_layers, n_wl) given gF = dL/dF (n_wl)."""
    alpha, wls, temps, tau_ref, alpha_ref = _rt_args(alpha, wls, temps, tau_ref, alpha_ref)
    gF = np.ascontiguousarray(gF, dtype=np.float64)
    g = np.zeros(alpha.shape, order="F")
    rc = lib().oracle_rt_flux_vjp(alpha.shape[0], wls.size, _p(alpha), _p(wls), _p(temps),
                                  _p(tau_ref), _p(alpha_ref), _p(gF), _p(g))
    assert rc == 0
    return g
