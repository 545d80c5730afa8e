"""Host-side logic of the Python mirror and the C-ABI library symbols.  CPU only."""
import ctypes
import os
import re

import numpy as np
import pytest

from korg_line import Line, LineList, Species, Wavelengths, approximate_gammas
from korg_line import _lib, workloads
from korg_line.constants import bohr_radius_cgs
from korg_line.problem import ContinuumGrid, Problem

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_line_known_answers():
    """test/linelist.jl:199-222 (the VALD fixture test/data/linelists/linelist.vald): Fe I
    3000.0414 A, E_lower 3.3014 eV, log gf -2.957, with some broadening fields missing."""
    wl = 1e-8 * 3000.0414
    full = Line(wl, -2.957, "Fe 1", 3.3014, 10 ** 7.280, 10 ** -3.910, -7.330)
    assert full.wl == pytest.approx(3000.0414e-8)
    assert full.gamma_rad == pytest.approx(1.905460717963248e7, rel=1e-15)
    assert full.gamma_stark == pytest.approx(0.0001230268770812381, rel=1e-15)
    assert full.vdW[0] == pytest.approx(4.6773514128719815e-8, rel=1e-15)
    no_rad = Line(wl, -2.957, "Fe 1", 3.3014, None, 10 ** -3.910, -7.330)
    assert no_rad.gamma_rad == pytest.approx(818252.5391161365, rel=1e-15)
    no_stark = Line(wl, -2.957, "Fe 1", 3.3014, None, None, -7.330)
    assert no_stark.gamma_stark == pytest.approx(0.00019044182974029873, rel=1e-15)
    no_vdw = Line(wl, -2.957, "Fe 1", 3.3014, 10 ** 7.280, None, None)
    assert no_vdw.vdW[0] == 9.953360714197118e-8
    assert no_vdw.vdW[1] == -1.0
    abo = Line(wl, -2.957, "Fe 1", 3.3014, 10 ** 7.280, 10 ** -3.910, 497.227)
    assert abo.vdW[0] == pytest.approx(1.3917417470792187e-14, rel=1e-12)
    assert abo.vdW[1] == pytest.approx(0.227, rel=1e-12)
    # test/linelist.jl:2-14: vdW -8.0 is log10(gamma), Angstrom input converts to cm
    assert Line(5000.0, 0.0, "Fe I", 1.0, vdW=-8.0).vdW[0] == pytest.approx(1e-8, rel=1e-15)
    assert Line(5001.0, 0.0, "Fe I", 1.0).wl == 5001.0 * 1e-8


def test_vdw_encodings():
    """src/linelist.jl:84-95: 0 -> none, (0, 20) -> Unsoeld fudge factor, >= 20 -> packed ABO."""
    assert Line(5000.0, 0.0, "Fe I", 1.0, vdW=0.0).vdW == (0.0, -1.0)
    base = 10 ** approximate_gammas(5000e-8, Species("Fe I"), 1.0)[1]
    assert Line(5000.0, 0.0, "Fe I", 1.0, vdW=1.5).vdW[0] == pytest.approx(1.5 * base, rel=1e-15)
    assert Line(5000.0, 0.0, "Fe I", 1.0, vdW=250.3).vdW[0] == \
        pytest.approx(250 * bohr_radius_cgs ** 2, rel=1e-15)


def test_approximate_gammas_special_cases():
    """Molecules and ionisation stages above III get no approximations; autoionizing lines no
    vdW (src/linelist.jl:150-152, 174-175)."""
    assert approximate_gammas(5000e-8, Species("CO"), 1.0) == (0.0, 0.0)
    assert approximate_gammas(5000e-8, Species("Fe IV"), 1.0) == (0.0, 0.0)
    # Ca I (chi = 6.1132 eV) at E_lower 5 eV, 5000 A (2.48 eV) is autoionizing
    assert approximate_gammas(5000e-8, Species("Ca I"), 5.0)[1] == 0.0


def test_vectorised_linelist_matches_scalar():
    rng = np.random.default_rng(3)
    n = 200
    wl = 4000 + 3000 * rng.random(n)
    sp = [Species(s) for s in ("Fe I", "Ti I", "Ni I", "Fe II", "Ca I")]
    idx = rng.integers(0, 5, n)
    gf = rng.uniform(-6, 0.5, n)
    El = rng.uniform(0, 8, n)
    grad = np.where(rng.random(n) < 0.5, np.nan, 10 ** rng.uniform(6, 9, n))
    ll = LineList.from_arrays(wl, gf, sp, idx, El, gamma_rad=grad)
    for i in range(n):
        g = None if np.isnan(grad[i]) else grad[i]
        l = Line(wl[i], gf[i], sp[idx[i]], El[i], g)
        assert ll.wl[i] == l.wl and ll.gamma_rad[i] == l.gamma_rad
        assert ll.gamma_stark[i] == l.gamma_stark
        assert (ll.vdW1[i], ll.vdW2[i]) == l.vdW


def test_species_parsing():
    """src/species.jl:181-262 docstring examples."""
    assert Species("H I") == Species("H 1") == Species("H_1") == Species("H.I") == Species("H")
    assert Species("H 2").charge == 1
    assert Species("H2").ismolecule() and Species("H2").charge == 0
    assert Species("01.00") == Species("H I")
    assert Species("02.01") == Species("He II")
    assert Species("02.1000") == Species("He II")
    assert Species("0608") == Species("CO")
    assert Species("Fe+") == Species("Fe II")
    assert Species("FeH").get_atoms() == (1, 26)
    assert repr(Species("Fe II")) == "Fe II"


def test_wavelength_grid_semantics():
    w = Wavelengths((5000, 5100))
    assert len(w) == 10001
    assert w.all_wls[0] == 5000 * 1e-8 and w.all_wls[-1] == 5100 * 1e-8
    # each point is the exact decimal wavelength times fl(1e-8), rounded once
    from fractions import Fraction
    for j in (1, 17, 5000, 9999):
        exact = Fraction(500000 + j, 100) * Fraction(1e-8)
        assert w.all_wls[j] == float(exact)
    m = Wavelengths([(5000, 5010), (5020, 5031, 1.0)])
    assert [len(r) for r in m.subspectrum_indices()] == [1001, 12]
    with pytest.raises(ValueError):
        Wavelengths([(5020, 5030), (5000, 5010)])


def test_rng_reproducible_and_pinned():
    """The counter-based generator: fixed values for a fixed (seed, stream, index)."""
    u = workloads.uniform(1, 1, 4)
    assert np.array_equal(u, workloads.uniform(1, 1, 4))
    assert np.array_equal(workloads.uniform(1, 1, 2, offset=2), u[2:])
    v = workloads.uniform(7, 3, 200000)
    assert 0 <= v.min() and v.max() < 1 and abs(v.mean() - 0.5) < 0.005
    golden = np.load(os.path.join(ROOT, "tests", "golden", "rng_pins.npy"))
    assert np.array_equal(u, golden)


def test_workload_shapes():
    P, atm, wls = workloads.solar_small()
    assert P.n_layers == 56 and P.n_wl == 10001 and len(P.lines) == 2000
    assert P.lines.is_sorted()
    frac = np.bincount(P.lines.species_index, minlength=5) / 2000
    assert np.allclose(frac, [0.6, 0.15, 0.10, 0.10, 0.05], atol=0.03)
    assert P.n_div_U.shape == (5, 56) and np.all(P.n_div_U > 0)
    assert P.cntm_mode == 1 and P.cntm_values.shape == (56, P.n_knots)
    fx = np.load(os.path.join(ROOT, "tests", "golden", "solar_small_lines.npz"))
    for k in ("wl", "log_gf", "E_lower", "gamma_rad", "gamma_stark", "vdW1", "vdW2"):
        assert np.array_equal(getattr(P.lines, k), fx[k]), k
    assert np.array_equal(P.lines.species_index, fx["species_index"])


def test_problem_flattening():
    """n_div_U = n / U(log T) per species (src/line_absorption.jl:42-44); xi broadcasts."""
    ll = LineList.from_lines([Line(5000.0, 0.0, "Fe I", 1.0), Line(5001.0, 0.0, "Ti I", 1.0)])
    T = np.array([4000.0, 5000.0])
    nd = {Species("Fe I"): np.array([1e10, 2e10]), Species("Ti I"): np.array([3e8, 4e8]),
          Species("H I"): np.array([1e16, 1e17])}
    pf = {Species("Fe I"): lambda x: 2 * x, Species("Ti I"): lambda x: x + 1}
    grid = ContinuumGrid([4000e-8, 6000e-8], [[1.0, 2.0], [3.0, 4.0]])
    P = Problem.from_korg_args(ll, Wavelengths((4999, 5002)), T, [1e12, 1e13], nd, pf, 1e5, grid)
    assert np.allclose(P.n_div_U[0], nd[Species("Fe I")] / (2 * np.log(T)))
    assert np.allclose(P.n_div_U[1], nd[Species("Ti I")] / (np.log(T) + 1))
    assert np.array_equal(P.xi, [1e5, 1e5])
    assert np.all(np.isnan(P.n_HeI)) and np.array_equal(P.n_HI, [1e16, 1e17])
    cb = [grid.layer(l) for l in range(2)]
    P2 = Problem.from_korg_args(ll, Wavelengths((4999, 5002)), T, [1e12, 1e13], nd, pf, 1e5, cb)
    assert P2.cntm_mode == 0 and P2.cntm_values.shape == (2, 2)
    assert P2.cntm_values[0, 0] == pytest.approx(1.5) and P2.cntm_values[0, 1] == pytest.approx(3.5)


def test_library_exports_declared_symbols():
    """libkorgline.so loads and exports every function include/korg_line.h declares (no GPU call)."""
    L = _lib.load()
    header = open(os.path.join(ROOT, "include", "korg_line.h")).read()
    declared = set(re.findall(r"^\s*(?:int|void|const char\*)\s+(korg_\w+)\s*\(", header, re.M))
    assert declared == set(_lib.EXPORTED)
    for name in declared:
        assert hasattr(L, name), name
    assert L.korg_abi_version() == _lib.ABI_VERSION == 17


def test_struct_layouts_match_header(tmp_path):
    """Every ctypes mirror of an ABI struct has the size and field offsets that the C compiler
    gives the struct declared in include/korg_line.h."""
    import subprocess
    structs = {"korg_lines": _lib.KorgLines, "korg_species": _lib.KorgSpecies,
               "korg_layers": _lib.KorgLayers, "korg_grid": _lib.KorgGrid,
               "korg_continuum": _lib.KorgContinuum, "korg_stats": _lib.KorgStats,
               "korg_tuning": _lib.KorgTuning, "korg_layer_grads": _lib.KorgLayerGrads,
               "korg_star": _lib.KorgStar, "korg_plan_info": _lib.KorgPlanInfo,
               "korg_fit_target": _lib.KorgFitTarget, "korg_flux_target": _lib.KorgFluxTarget,
               "korg_htransition": _lib.KorgHTransition, "korg_htables": _lib.KorgHTables,
               "korg_hlayers": _lib.KorgHLayers, "korg_layer_tangents": _lib.KorgLayerTangents,
               "korg_mol_xsection": _lib.KorgMolXsection,
               "korg_filter_star": _lib.KorgFilterStar}
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "korg_line.h"', "int main(void) {"]
    for cname, cls in structs.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in cls._fields_:
            lines.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append('printf("pool_bytes 1000 %lld\\n", (long long)KORG_FILTER_POOL_BYTES(1000));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    inc = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")
    subprocess.check_call(["gcc", "-std=c99", "-I", inc, str(src), "-o", str(exe)])
    got = {}
    for row in subprocess.check_output([str(exe)], text=True).splitlines():
        cname, what, val = row.split()
        got[(cname, what)] = int(val)
    assert got[("pool_bytes", "1000")] == _lib.filter_pool_bytes(1000)
    for cname, cls in structs.items():
        assert got[(cname, "size")] == ctypes.sizeof(cls), cname
        for fname, _ in cls._fields_:
            assert got[(cname, fname)] == getattr(cls, fname).offset, (cname, fname)


def test_device_path_fails_loudly_without_gpu():
    """No silent CPU fallback: creating a context without a HIP device raises."""
    from korg_line.hipmem import device_count
    if device_count() > 0:
        pytest.skip("a GPU is present")
    from korg_line import Context
    with pytest.raises(_lib.KorgError):
        Context(0)
