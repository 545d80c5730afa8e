"""ctypes binding of ``libkorgline.so`` (the C ABI declared in ``include/korg_line.h``).

The library is built in-tree (``korg.jl_amd/libkorgline.so``).  Loading fails loudly when it is
missing: there is no CPU fallback for the line-opacity path.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
PKG_ROOT = os.path.dirname(HERE)
LIB_PATH = os.path.join(PKG_ROOT, "libkorgline.so")

KORG_OK = 0
KORG_ERR_INVALID_ARGUMENT = 1
KORG_ERR_HIP = 2
KORG_ERR_OUT_OF_MEMORY = 3
KORG_ERR_NO_DEVICE = 4
KORG_ERR_UNSUPPORTED = 5
KORG_ERR_CAPACITY = 6
NOALLOC_SCRATCH_BYTES = 64 << 20  # KORG_NOALLOC_SCRATCH_BYTES

PRECISION_FP64 = 0
PRECISION_FP32_FIXUP = 1
PRECISION_FP64_SERIES = 2
DTYPE_F64 = 0
DTYPE_F32 = 1
CNTM_AT_LINES = 0
CNTM_GRID = 1

_c_dp = ctypes.c_void_p  # all array pointers are passed as raw addresses


class KorgLines(ctypes.Structure):
    _fields_ = [("n_lines", ctypes.c_int64), ("wl", _c_dp), ("log_gf", _c_dp),
                ("E_lower", _c_dp), ("gamma_rad", _c_dp), ("gamma_stark", _c_dp),
                ("vdW1", _c_dp), ("vdW2", _c_dp), ("species", _c_dp), ("window_lb", _c_dp),
                ("window_ub", _c_dp)]


class KorgSpecies(ctypes.Structure):
    _fields_ = [("n_species", ctypes.c_int32), ("mass", _c_dp), ("is_molecule", _c_dp)]


class KorgLayers(ctypes.Structure):
    _fields_ = [("n_layers", ctypes.c_int32), ("temps", _c_dp), ("n_e", _c_dp),
                ("n_div_U", _c_dp), ("n_HI", _c_dp), ("n_HeI", _c_dp), ("n_H2", _c_dp),
                ("xi", _c_dp)]


class KorgGrid(ctypes.Structure):
    _fields_ = [("n_wl", ctypes.c_int64), ("all_wls", _c_dp)]


class KorgContinuum(ctypes.Structure):
    _fields_ = [("mode", ctypes.c_int32), ("n_knots", ctypes.c_int64), ("knots", _c_dp),
                ("values", _c_dp)]


class KorgStats(ctypes.Structure):
    _fields_ = [("n_evals", ctypes.c_double), ("n_lines_used", ctypes.c_int64),
                ("setup_ms", ctypes.c_double), ("bin_ms", ctypes.c_double),
                ("profile_ms", ctypes.c_double), ("core_ms", ctypes.c_double),
                ("total_ms", ctypes.c_double), ("n_evals_branch", ctypes.c_double * 4),
                ("n_evals_near", ctypes.c_double), ("n_wing_mfma", ctypes.c_double),
                ("n_evals_tail", ctypes.c_double), ("n_evals_near_A", ctypes.c_double),
                ("far_ms", ctypes.c_double), ("n_far_entries", ctypes.c_double)]


class KorgLayerGrads(ctypes.Structure):
    _fields_ = [(k, ctypes.c_void_p) for k in ("T", "n_e", "xi", "n_HI", "n_HeI", "n_H2",
                                               "n_div_U")]


class KorgLayerTangents(ctypes.Structure):
    _fields_ = [(k, ctypes.c_void_p) for k in ("T", "n_e", "xi", "n_HI", "n_HeI", "n_H2",
                                               "n_div_U")]


class KorgMolXsection(ctypes.Structure):
    _fields_ = [("n_vmic", ctypes.c_int32), ("n_logT", ctypes.c_int32), ("n_wl", ctypes.c_int64),
                ("vmic", _c_dp), ("logT", _c_dp), ("wls", _c_dp), ("sigma", _c_dp),
                ("n_species", _c_dp)]


class KorgFilterStar(ctypes.Structure):
    _fields_ = [("n_ranges", ctypes.c_int32), ("ranges", _c_dp), ("line_buffer", ctypes.c_double)]


def filter_pool_bytes(n):
    """KORG_FILTER_POOL_BYTES: bytes of a filter pool of n lines."""
    return 60 * int(n)


class KorgStar(ctypes.Structure):
    _fields_ = [("alpha", ctypes.c_void_p), ("grid", KorgGrid), ("lines", KorgLines),
                ("layers", KorgLayers), ("cntm", KorgContinuum),
                ("cutoff_threshold", ctypes.c_double)]


class KorgTuning(ctypes.Structure):
    _fields_ = [("points_per_lane", ctypes.c_int32), ("waves_per_tile", ctypes.c_int32),
                ("waves_per_cu", ctypes.c_int32), ("count_branches", ctypes.c_int32),
                ("tile_points", ctypes.c_int32), ("near_waves_per_cu", ctypes.c_int32),
                ("near_layout", ctypes.c_int32), ("near_concurrent", ctypes.c_int32),
                ("far_field", ctypes.c_int32)]


class KorgFitTarget(ctypes.Structure):
    _fields_ = [("tau_ref", ctypes.c_void_p), ("alpha_ref", ctypes.c_void_p),
                ("data", ctypes.c_void_p), ("sigma", ctypes.c_void_p), ("flux", ctypes.c_void_p),
                ("chi2", ctypes.c_void_p), ("grads", KorgLayerGrads)]


class KorgFluxTarget(ctypes.Structure):
    _fields_ = [("tau_ref", ctypes.c_void_p), ("alpha_ref", ctypes.c_void_p),
                ("flux", ctypes.c_void_p), ("cntm", ctypes.c_void_p)]


class KorgPlanInfo(ctypes.Structure):
    _fields_ = [("n_stars", ctypes.c_int64), ("n_lines", ctypes.c_int64),
                ("n_points", ctypes.c_int64), ("tile_capacity", ctypes.c_int64),
                ("tile_entries", ctypes.c_int64), ("n_captures", ctypes.c_int32),
                ("pad", ctypes.c_int32), ("tuning", KorgTuning)]


class KorgHTransition(ctypes.Structure):
    _fields_ = [("lower", ctypes.c_int32), ("upper", ctypes.c_int32), ("n_T", ctypes.c_int32),
                ("n_ne", ctypes.c_int32), ("n_dnu", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("log_gf", ctypes.c_double), ("off_T", ctypes.c_int64), ("off_ne", ctypes.c_int64),
                ("off_dnu", ctypes.c_int64), ("off_profile", ctypes.c_int64),
                ("off_lambda0", ctypes.c_int64)]


class KorgHTables(ctypes.Structure):
    _fields_ = [("n_transitions", ctypes.c_int32), ("transitions", ctypes.POINTER(KorgHTransition)),
                ("temps", _c_dp), ("n_e", _c_dp), ("log_dnu", _c_dp), ("log_profile", _c_dp),
                ("lambda0", _c_dp), ("n_temps", ctypes.c_int64), ("n_n_e", ctypes.c_int64),
                ("n_log_dnu", ctypes.c_int64), ("n_log_profile", ctypes.c_int64),
                ("n_lambda0", ctypes.c_int64)]


class KorgHLayers(ctypes.Structure):
    _fields_ = [("n_layers", ctypes.c_int32), ("temps", _c_dp), ("n_e", _c_dp), ("n_HI", _c_dp),
                ("n_HeI", _c_dp), ("U_HI", _c_dp), ("xi", _c_dp)]


ABI_VERSION = 17  # KORG_LINE_ABI_VERSION of include/korg_line.h

EXPORTED = ["korg_create", "korg_destroy", "korg_last_error", "korg_abi_version",
            "korg_line_absorption", "korg_line_absorption_device", "korg_set_stream",
            "korg_synchronize", "korg_get_stats", "korg_get_windows", "korg_set_tuning",
            "korg_line_absorption_vjp", "korg_line_absorption_vjp_device",
            "korg_line_absorption_batched", "korg_get_tuning", "korg_plan_batched",
            "korg_plan_launch", "korg_plan_wait", "korg_plan_get_info", "korg_plan_destroy",
            "korg_autotune_batched", "korg_fit_gradient_batched", "korg_plan_fit",
            "korg_continuum_batched", "korg_flux_batched", "korg_set_hydrogen_tables",
            "korg_hydrogen_line_absorption", "korg_line_absorption_device_noalloc",
            "korg_line_absorption_jvp", "korg_line_absorption_jvp_device",
            "korg_molecular_cross_sections", "korg_molecular_cross_sections_device",
            "korg_filter_linelist_batched"]

_lib = None


class KorgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"korg_line error {code}: {msg}")
        self.code = code


def load():
    """Load the library (no GPU call is made)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build() or "
                          f"`make -C {PKG_ROOT}`")
    L = ctypes.CDLL(LIB_PATH)
    vp, i32, i64, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_double
    L.korg_create.argtypes = [ctypes.POINTER(vp), i32]
    L.korg_create.restype = i32
    L.korg_destroy.argtypes = [vp]
    L.korg_destroy.restype = None
    L.korg_last_error.argtypes = [vp]
    L.korg_last_error.restype = ctypes.c_char_p
    L.korg_abi_version.argtypes = []
    L.korg_abi_version.restype = i32
    common = [vp, vp, i32, ctypes.POINTER(KorgGrid), ctypes.POINTER(KorgLines),
              ctypes.POINTER(KorgSpecies), ctypes.POINTER(KorgLayers),
              ctypes.POINTER(KorgContinuum), d, i32]
    L.korg_line_absorption.argtypes = common
    L.korg_line_absorption.restype = i32
    L.korg_line_absorption_device.argtypes = common + [i32]
    L.korg_line_absorption_device.restype = i32
    L.korg_line_absorption_device_noalloc.argtypes = common + [i32]
    L.korg_line_absorption_device_noalloc.restype = i32
    L.korg_set_stream.argtypes = [vp, vp]
    L.korg_set_stream.restype = i32
    L.korg_synchronize.argtypes = [vp]
    L.korg_synchronize.restype = i32
    L.korg_get_stats.argtypes = [vp, ctypes.POINTER(KorgStats)]
    L.korg_get_stats.restype = i32
    L.korg_get_windows.argtypes = [vp, vp, vp, i64]
    L.korg_get_windows.restype = i32
    L.korg_set_tuning.argtypes = [vp, ctypes.POINTER(KorgTuning)]
    L.korg_set_tuning.restype = i32
    vjp = [vp, vp, i32, ctypes.POINTER(KorgGrid), ctypes.POINTER(KorgLines),
           ctypes.POINTER(KorgSpecies), ctypes.POINTER(KorgLayers), ctypes.POINTER(KorgContinuum),
           d, ctypes.POINTER(KorgLayerGrads)]
    L.korg_line_absorption_vjp.argtypes = vjp
    L.korg_line_absorption_vjp.restype = i32
    L.korg_line_absorption_vjp_device.argtypes = vjp + [i32]
    L.korg_line_absorption_vjp_device.restype = i32
    jvp = [vp, ctypes.POINTER(vp), i32, ctypes.POINTER(KorgGrid), ctypes.POINTER(KorgLines),
           ctypes.POINTER(KorgSpecies), ctypes.POINTER(KorgLayers), ctypes.POINTER(KorgContinuum),
           d, i64, ctypes.POINTER(KorgLayerTangents)]
    L.korg_line_absorption_jvp.argtypes = jvp
    L.korg_line_absorption_jvp.restype = i32
    L.korg_line_absorption_jvp_device.argtypes = jvp + [i32]
    L.korg_line_absorption_jvp_device.restype = i32
    mol = [vp, vp, i32, ctypes.POINTER(KorgGrid), ctypes.c_int32, vp, vp, i64,
           ctypes.POINTER(KorgMolXsection)]
    L.korg_molecular_cross_sections.argtypes = mol
    L.korg_molecular_cross_sections.restype = i32
    L.korg_molecular_cross_sections_device.argtypes = mol + [i32]
    L.korg_molecular_cross_sections_device.restype = i32
    L.korg_filter_linelist_batched.argtypes = [vp, ctypes.POINTER(KorgLines), i64,
                                               ctypes.POINTER(KorgFilterStar), vp, i64,
                                               ctypes.POINTER(KorgLines), ctypes.POINTER(i64)]
    L.korg_filter_linelist_batched.restype = i32
    L.korg_line_absorption_batched.argtypes = [vp, i64, ctypes.POINTER(KorgStar),
                                               ctypes.POINTER(KorgSpecies), i32, i32, i32]
    L.korg_line_absorption_batched.restype = i32
    L.korg_get_tuning.argtypes = [vp, ctypes.POINTER(KorgTuning)]
    L.korg_get_tuning.restype = i32
    L.korg_plan_batched.argtypes = [vp, i64, ctypes.POINTER(KorgStar), ctypes.POINTER(KorgSpecies),
                                    i32, i32, d, ctypes.POINTER(vp)]
    L.korg_plan_batched.restype = i32
    L.korg_plan_launch.argtypes = [vp, i32]
    L.korg_plan_launch.restype = i32
    L.korg_plan_wait.argtypes = [vp]
    L.korg_plan_wait.restype = i32
    L.korg_plan_get_info.argtypes = [vp, ctypes.POINTER(KorgPlanInfo)]
    L.korg_plan_get_info.restype = i32
    L.korg_plan_destroy.argtypes = [vp]
    L.korg_plan_destroy.restype = None
    L.korg_autotune_batched.argtypes = [vp, i64, ctypes.POINTER(KorgStar),
                                        ctypes.POINTER(KorgSpecies), i32, i32, ctypes.c_char_p,
                                        ctypes.POINTER(KorgTuning)]
    L.korg_autotune_batched.restype = i32
    L.korg_fit_gradient_batched.argtypes = [vp, i64, ctypes.POINTER(KorgStar),
                                            ctypes.POINTER(KorgFitTarget),
                                            ctypes.POINTER(KorgSpecies), i32, i32, i32]
    L.korg_fit_gradient_batched.restype = i32
    L.korg_plan_fit.argtypes = [vp, i64, ctypes.POINTER(KorgStar), ctypes.POINTER(KorgFitTarget),
                                ctypes.POINTER(KorgSpecies), i32, i32, d, ctypes.POINTER(vp)]
    L.korg_plan_fit.restype = i32
    L.korg_continuum_batched.argtypes = [vp, i64, ctypes.POINTER(KorgStar), i32, i32]
    L.korg_continuum_batched.restype = i32
    L.korg_flux_batched.argtypes = [vp, i64, ctypes.POINTER(KorgStar), ctypes.POINTER(KorgFluxTarget),
                                    i32, i32]
    L.korg_flux_batched.restype = i32
    L.korg_set_hydrogen_tables.argtypes = [vp, ctypes.POINTER(KorgHTables)]
    L.korg_set_hydrogen_tables.restype = i32
    L.korg_hydrogen_line_absorption.argtypes = [vp, vp, i32, ctypes.POINTER(KorgGrid),
                                                ctypes.POINTER(KorgHLayers), d, i32, i32]
    L.korg_hydrogen_line_absorption.restype = i32
    if L.korg_abi_version() != ABI_VERSION:
        raise ImportError(f"{LIB_PATH} has ABI version {L.korg_abi_version()}, these bindings "
                          f"expect {ABI_VERSION}; rebuild it")
    _lib = L
    return L


def check(ctx, rc):
    if rc != KORG_OK:
        msg = load().korg_last_error(ctx)
        raise KorgError(rc, msg.decode() if msg else "")
