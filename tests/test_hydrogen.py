"""Hydrogen lines (src/hydrogen_line_absorption.jl:75-169): the oracle (oracle/hydrogen_oracle.py)
and the device (korg_hydrogen_line_absorption, through the C ABI).

The oracle is pinned by the reference's own test (test/line_absorption.jl:19-39): the Lyman
profile of tests/golden/lyman_absorption.h5 (T = 9000 K, n_e = 1e11, n_HI = 1e13, no MHD) within
its absolute tolerance 5e-9 at every point.  Korg's H I partition function at 9000 K is not
available offline, so it enters as the one constant fitted on the line cores, which must be the
physical value (2 plus the excited levels' share, between 2.00001 and 2.0001).  Its helper
functions are checked against the C oracle's, which the line-opacity tests pin.  The device is
checked against the oracle at 1e-12 (Float64) on the golden case and on the solar atmosphere around
H alpha, H beta and the Paschen lines, with MHD on, and around the Brackett lines (their own pins
are in test_brackett_oracle.py).
"""
import ctypes
import math
import os

import numpy as np
import pytest

from oracle import hydrogen_oracle as H
from oracle import oracle
from korg_line import _lib
from korg_line.hdf5 import H5File
from korg_line.hydrogen import StarkTables, DEFAULT_TABLES
from korg_line.wavelengths import Wavelengths

HERE = os.path.dirname(os.path.abspath(__file__))
LYMAN = os.path.join(HERE, "golden", "lyman_absorption.h5")


def oracle_transitions(path=DEFAULT_TABLES):
    f = H5File(path)
    out = []
    for k in f.keys():
        a = f.attrs(k)
        out.append(dict(lower=int(a["lower"]), upper=int(a["upper"]), log_gf=float(a["log_gf"]),
                        temps=f.read(k + "/temps"), nes=f.read(k + "/electron_number_densities"),
                        dnu=f.read(k + "/delta_nu_over_F0"), profile=f.read(k + "/profile"),
                        lambda0=f.read(k + "/lambda0")))
    return out


@pytest.fixture(scope="module")
def transitions():
    return oracle_transitions()


def lyman_case():
    f = H5File(LYMAN)
    a = f.attrs("profile")
    wls = Wavelengths((a["start_wl"], a["stop_wl"], a["wl_step"]))
    return a, wls, f.read("profile")


def fit_U(ref, alpha_U1):
    """The H I partition function the golden implies: alpha scales as 1/U (one constant)."""
    core = ref > 1e-6
    return 1.0 / np.median(ref[core] / alpha_U1[core])


# ------------------------------------------------------------------------------------ CPU
def test_stark_tables_pack():
    t = StarkTables()
    assert len(t) == 84
    lowers = [t.transitions[k].lower for k in range(len(t))]
    assert sorted(set(lowers)) == [1, 2, 3]
    assert all(t.transitions[k].upper > t.transitions[k].lower for k in range(len(t)))
    assert np.all(np.isfinite(t.log_profile)) and t.log_profile.min() >= -700.0
    assert t.names == sorted(t.names)


def test_oracle_matches_lyman_golden(transitions):
    a, wls, ref = lyman_case()
    alpha = np.zeros(len(wls))
    H.hydrogen_line_absorption(alpha, wls.all_wls, a["T"], a["ne"], a["nH_I"], 0.0, 1.0, 0.0,
                               15e-7, transitions, use_MHD=False)
    U = fit_U(ref, alpha)
    assert 2.00001 < U < 2.0001, U
    assert np.max(np.abs(ref - alpha / U)) <= 5e-9


def test_oracle_helpers_match_c_oracle():
    rng = np.random.default_rng(4)
    for _ in range(200):
        a, v = 10 ** rng.uniform(-4, 1), rng.uniform(0, 8)
        assert H.voigt_hjerting(a, v) == pytest.approx(oracle.voigt_hjerting(a, v), rel=1e-14)
    assert H.sigma_line(5e-5) == pytest.approx(oracle.sigma_line(5e-5), rel=1e-15)
    assert H.doppler_width(6.5e-5, 5000.0, H.H_MASS, 1e5) == \
        pytest.approx(oracle.doppler_width(6.5e-5, 5000.0, H.H_MASS, 1e5), rel=1e-15)
    s = 1180.0 * H.bohr_radius_cgs ** 2
    assert H.scaled_vdW_abo(s, 0.677, H.H_MASS, 6000.0) == \
        pytest.approx(oracle.scaled_vdW((s, 0.677), H.H_MASS, 6000.0), rel=1e-13)
    assert H.line_profile(6.5e-5, 3e-10, 1e-10, 2.0, 6.5002e-5) == \
        pytest.approx(oracle.line_profile(6.5e-5, 3e-10, 1e-10, 2.0, 6.5002e-5), rel=1e-13)


def test_mhd_occupation_probability():
    """w in (0, 1], 1 in the low-density limit, falling with n, with n_e and with n_HI."""
    T = 6000.0
    assert H.hummer_mihalas_w(T, 2, 0.0, 0.0, 0.0) == 1.0
    w = [H.hummer_mihalas_w(T, n, 1e16, 1e15, 1e14) for n in range(1, 31)]
    assert all(0 < x <= 1 for x in w) and all(b < a for a, b in zip(w, w[1:]))
    assert H.hummer_mihalas_w(T, 10, 1e16, 1e15, 1e15) < H.hummer_mihalas_w(T, 10, 1e16, 1e15, 1e14)
    assert H.hummer_mihalas_w(T, 10, 1e17, 1e15, 1e14) < H.hummer_mihalas_w(T, 10, 1e16, 1e15, 1e14)
    # low n in a photosphere: essentially unperturbed
    assert H.hummer_mihalas_w(T, 2, 1e16, 1e15, 1e13) > 0.9999


# ------------------------------------------------------------------------------------ GPU
def device_hlines(ctx, wls, layers, window, use_MHD, dtype=np.float64):
    """alpha [L, n_wl] from korg_hydrogen_line_absorption, every array on the device."""
    from korg_line.hipmem import DeviceArray
    if getattr(ctx, "_htables", None) is None:
        ctx.set_hydrogen_tables()
    L = len(layers["T"])
    dw = DeviceArray.from_numpy(np.ascontiguousarray(wls, dtype=np.float64))
    arrs = {k: DeviceArray.from_numpy(np.ascontiguousarray(layers[k], dtype=np.float64))
            for k in ("T", "ne", "nH", "nHe", "U", "xi")}
    alpha = DeviceArray.zeros((len(wls), L), dtype)
    grid = _lib.KorgGrid(len(wls), dw.ptr)
    lay = _lib.KorgHLayers(L, arrs["T"].ptr, arrs["ne"].ptr, arrs["nH"].ptr, arrs["nHe"].ptr,
                           arrs["U"].ptr, arrs["xi"].ptr)
    dt = _lib.DTYPE_F32 if dtype == np.float32 else _lib.DTYPE_F64
    rc = ctx._L.korg_hydrogen_line_absorption(ctx._h, alpha.ptr, dt, ctypes.byref(grid),
                                              ctypes.byref(lay), window, int(use_MHD), 1)
    _lib.check(ctx._h, rc)
    return alpha.numpy().T


@pytest.mark.gpu
def test_device_lyman_golden(ctx, transitions):
    a, wls, ref = lyman_case()
    layers = dict(T=[a["T"]], ne=[a["ne"]], nH=[a["nH_I"]], nHe=[0.0], U=[1.0], xi=[0.0])
    got = device_hlines(ctx, wls.all_wls, layers, 15e-7, False)[0]
    alpha = np.zeros(len(wls))
    H.hydrogen_line_absorption(alpha, wls.all_wls, a["T"], a["ne"], a["nH_I"], 0.0, 1.0, 0.0,
                               15e-7, transitions, use_MHD=False)
    np.testing.assert_allclose(got, alpha, rtol=1e-12, atol=0)
    assert np.max(np.abs(ref - got / fit_U(ref, got))) <= 5e-9


def solar_layers():
    from korg_line.atmosphere import Atmosphere, partition_function
    from korg_line.species import Species
    atm = Atmosphere.from_model("sun")
    nd = atm.n_densities
    U = partition_function("H_I")(np.log(atm.temps))
    return atm, dict(T=atm.temps, ne=atm.n_e, nH=nd[Species("H I")], nHe=nd[Species("He I")],
                     U=U, xi=np.full(atm.n_layers, 1e5))


@pytest.mark.gpu
@pytest.mark.parametrize("rng_A", [(4855.0, 4868.0), (6556.0, 6570.0), (8590.0, 8602.0)])
def test_device_solar_matches_oracle(ctx, transitions, rng_A):
    """H beta, H alpha (resonance and Stark profiles) and Paschen lines near 8598 A, MHD on,
    56 solar layers; Float64 alpha at 1e-12, Float32 alpha within its rounding."""
    atm, lay = solar_layers()
    wls = Wavelengths((rng_A[0], rng_A[1], 0.01)).all_wls
    got = device_hlines(ctx, wls, lay, 150e-8, True)
    ref = np.zeros_like(got)
    for l in range(atm.n_layers):
        H.hydrogen_line_absorption(ref[l], wls, lay["T"][l], lay["ne"][l], lay["nH"][l],
                                   lay["nHe"][l], lay["U"][l], lay["xi"][l], 150e-8, transitions,
                                   use_MHD=True)
    assert np.count_nonzero(ref) > ref.size // 2
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=0)
    got32 = device_hlines(ctx, wls, lay, 150e-8, True, dtype=np.float32)
    np.testing.assert_allclose(got32, ref, rtol=1e-6, atol=0)


@pytest.mark.gpu
@pytest.mark.parametrize("rng_A,use_MHD", [((15000.0, 15300.0, 0.05), False),
                                           ((21640.0, 21680.0, 0.02), False),
                                           ((21655.0, 21668.0, 0.01), True)])
def test_device_brackett_matches_oracle(ctx, transitions, rng_A, use_MHD):
    """The Brackett series (src/hydrogen_line_absorption.jl:152-169) with the tabulated lines, 56
    solar layers: the high-m lines near the series limit (m = 19 .. 30 reach 15,000-15,300 A, Stark
    dominated in the deep layers) and Br gamma (its window clipped by a grid narrower than it, MHD
    on), in both regimes of the Doppler convolution; Float64 alpha at 1e-12, Float32 within its
    rounding."""
    atm, lay = solar_layers()
    wls = Wavelengths(rng_A).all_wls
    got = device_hlines(ctx, wls, lay, 150e-8, use_MHD)
    ref = np.zeros_like(got)
    br = np.zeros_like(got)
    for l in range(atm.n_layers):
        H.hydrogen_line_absorption(ref[l], wls, lay["T"][l], lay["ne"][l], lay["nH"][l],
                                   lay["nHe"][l], lay["U"][l], lay["xi"][l], 150e-8, transitions,
                                   use_MHD=use_MHD)
        ws = [H.hummer_mihalas_w(lay["T"][l], n, lay["nH"][l], lay["nHe"][l], lay["ne"][l])
              if use_MHD else 1.0 for n in range(1, 31)]
        H.brackett_lines(br[l], wls, lay["T"][l], lay["ne"][l], lay["nH"][l], lay["U"][l],
                         lay["xi"][l], ws, 30)
    # the Brackett lines are most of the hydrogen opacity here, and reach every layer
    assert np.all(br.max(axis=1) > 0) and np.sum(br) > 0.5 * np.sum(ref)
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=0)
    got32 = device_hlines(ctx, wls, lay, 150e-8, use_MHD, dtype=np.float32)
    np.testing.assert_allclose(got32, ref, rtol=1e-6, atol=0)


@pytest.mark.gpu
def test_device_rejects_missing_tables(ctx):
    from korg_line import Context
    atm, lay = solar_layers()
    fresh = Context(0)
    fresh._htables = True  # skip the upload: the call must report the missing tables
    with pytest.raises(_lib.KorgError):
        device_hlines(fresh, Wavelengths((4860.0, 4861.0)).all_wls, lay, 150e-8, True)


@pytest.mark.gpu
def test_context_wrapper_on_a_device_problem(ctx):
    """Context.hydrogen_line_absorption on a DeviceProblem (its grid and layers, synthesize's MHD
    default: on below 13,000 A) adds exactly what the direct call computes, onto existing alpha."""
    from korg_line import DeviceProblem, workloads
    from korg_line.atmosphere import Atmosphere, partition_function
    atm = Atmosphere.from_model("sun")
    P = workloads.make_problem(atm, workloads.synthetic_linelist(3, 20, 6540.0, 6590.0, (-3.0, 0.0),
                                                                 (0.0, 4.0)),
                               Wavelengths((6550.0, 6575.0)))
    D = DeviceProblem(P)
    base = np.random.default_rng(1).uniform(1.0, 2.0, (P.n_layers, P.n_wl))
    a = D.new_alpha("float64", fill=base)
    U = partition_function("H_I")(np.log(P.temps))
    ctx.hydrogen_line_absorption(a, D, U)
    lay = dict(T=P.temps, ne=P.n_e, nH=P.n_HI, nHe=P.n_HeI, U=U, xi=P.xi)
    direct = device_hlines(ctx, P.all_wls, lay, 150e-8, True)
    # alpha - base is exact to half an ulp of base (base in [1, 2): 1.1e-16)
    np.testing.assert_allclose(a.numpy().T - base, direct, rtol=1e-12, atol=4.5e-16)
    assert np.max(direct) > 0
