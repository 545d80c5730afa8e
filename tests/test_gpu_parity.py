"""Parity of the gfx950 library (through the C ABI) with the Float64 oracle.  Needs a GPU.

Tolerances (stated by the north star): Float64 form <= 1e-12 relative to the oracle per alpha
element; Float32-with-Float64-fix-ups form <= 1e-4 relative per line-opacity element above
1e-8 of the maximum, and <= 1e-4 absolute in normalized flux.  Windows (index work) bit-exact.
"""
import numpy as np
import pytest

from oracle import oracle
from korg_line import Context, DeviceProblem, Line, LineList, Problem, Species, Wavelengths
from korg_line import _lib, workloads
from korg_line.atmosphere import Atmosphere, continuum_grid
from korg_line.problem import ContinuumGrid

pytestmark = pytest.mark.gpu

FP64_RTOL = 1e-12
FP32_RTOL = 1e-4
FLUX_ATOL = 1e-4


def oracle_alpha(P, base=None, windows=False):
    a = np.zeros((P.n_layers, P.n_wl), order="F") if base is None else np.asfortranarray(base.copy())
    lb, ub, nev = oracle.line_absorption(a, P, return_windows=True)
    return (a, lb, ub, nev) if windows else a


def gpu_alpha(ctx, P, precision, dtype=np.float64, base=None):
    a = np.zeros((P.n_layers, P.n_wl), dtype=dtype, order="F") if base is None else \
        np.asfortranarray(base.astype(dtype))
    ctx.line_absorption(a, P, precision)
    return a


def rel_err(a, ref, floor_frac=1e-8):
    """Max |a - ref| / |ref| over elements with |ref| >= floor_frac * max|ref|."""
    ref = np.asarray(ref, dtype=np.float64)
    a = np.asarray(a, dtype=np.float64)
    m = np.abs(ref) >= floor_frac * np.abs(ref).max()
    return float(np.max(np.abs(a[m] - ref[m]) / np.abs(ref[m])))


def check_fp64(ctx, P):
    ref, lb, ub, nev = oracle_alpha(P, windows=True)
    got = gpu_alpha(ctx, P, "fp64")
    glb, gub = ctx.windows(len(P.lines))
    skip = lb > ub
    assert np.array_equal(skip, glb > gub)
    assert np.array_equal(lb[~skip], glb[~skip]) and np.array_equal(ub[~skip], gub[~skip])
    assert ctx.stats()["n_evals"] == nev
    np.testing.assert_allclose(got, ref, rtol=FP64_RTOL, atol=1e-300)
    return ref


def check_fp32(ctx, P, ref=None):
    if ref is None:
        ref = oracle_alpha(P)
    got = gpu_alpha(ctx, P, "fp32")
    err = rel_err(got, ref)
    assert err <= FP32_RTOL, err
    return got, err


# ----------------------------------------------------------------------------- real lines
def test_galah_lines_fp64_and_fp32(ctx):
    """Real lines: the GALAH DR3 Fe, Ti, Ni and Ca I-II lines of 4795-4825 A (tests/golden, raw
    file columns), built as Korg builds them, including 20+ packed-ABO van der Waals lines, on
    the 4800-4820 A grid of the solar atmosphere."""
    import os
    from korg_line.linelist import GALAH_COLUMNS, galah_linelist_from_columns
    z = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                             "galah_4800_fe_ti_ni_ca.npz"))
    ll = galah_linelist_from_columns(*(z[k] for k in GALAH_COLUMNS))
    assert len(ll) == 204 and np.count_nonzero(ll.vdW2 != -1.0) >= 20 and ll.is_sorted()
    atm = Atmosphere.from_model("sun")
    P = workloads.make_problem(atm, ll, Wavelengths((4800, 4820)))
    ref = check_fp64(ctx, P)
    check_fp32(ctx, P, ref)


# ----------------------------------------------------------------------------- solar, c1
def test_solar_small_fp64_and_fp32(ctx):
    P, atm, wls = workloads.solar_small()
    ref = check_fp64(ctx, P)
    got, err = check_fp32(ctx, P, ref)
    # normalized flux: continuum on the fine grid + lines, by the oracle's Korg planar transfer
    cntm = np.array([np.interp(wls.all_wls, P.cntm_knots, P.cntm_values[l])
                     for l in range(P.n_layers)])
    tau, aref = workloads.fit_reference(atm)
    fc = oracle.rt_flux(cntm, wls.all_wls, atm.temps, tau, aref)
    f_ref = oracle.rt_flux(cntm + ref, wls.all_wls, atm.temps, tau, aref) / fc
    f_got = oracle.rt_flux(cntm + got, wls.all_wls, atm.temps, tau, aref) / fc
    assert np.max(np.abs(f_got - f_ref)) <= FLUX_ATOL
    assert f_ref.min() < 0.5  # strong lines present: the comparison is not vacuous


def test_solar_small_at_line_continuum(ctx):
    """alpha_cntm given per line centre (callables), not as the coarse grid."""
    P, _, _ = workloads.solar_small(n_lines=400, cntm="callables")
    assert P.cntm_mode == _lib.CNTM_AT_LINES
    ref = check_fp64(ctx, P)
    check_fp32(ctx, P, ref)


def test_solar_wide_subrange(ctx):
    """The 300,000-line wide workload's parameters, restricted to 5000-5020 A."""
    P, _, _ = workloads.solar_wide_subset(sub_range=(5000, 5020))
    ref = check_fp64(ctx, P)
    check_fp32(ctx, P, ref)


# ----------------------------------------------------------------------------- semantics
def test_accumulates_into_alpha(ctx):
    """Lines are added into what alpha holds.  The Float64 form sums a wavelength point's lines
    from zero and adds once (bitwise alpha_in + lines); the Float32 form adds its far-wing and
    Voigt-core sums in two steps, equal to rounding."""
    P, _, _ = workloads.solar_small(n_lines=200)
    base = np.random.default_rng(1).random((P.n_layers, P.n_wl)) * 1e-6
    z = gpu_alpha(ctx, P, "fp64")
    assert np.array_equal(gpu_alpha(ctx, P, "fp64", base=base), base + z)
    z = gpu_alpha(ctx, P, "fp32")
    np.testing.assert_allclose(gpu_alpha(ctx, P, "fp32", base=base), base + z, rtol=1e-15,
                               atol=0)


def test_empty_linelist_leaves_alpha(ctx):
    P, _, _ = workloads.solar_small(n_lines=10)
    E = Problem(P.all_wls, P.lines.take(np.zeros(0, dtype=int)), P.species_mass,
                P.species_is_mol, P.n_div_U, P.temps, P.n_e, P.n_HI, P.n_HeI, P.n_H2, P.xi,
                P.cntm_mode, P.cntm_values, P.cntm_knots, P.cutoff_threshold)
    base = np.asfortranarray(np.random.default_rng(2).random((P.n_layers, P.n_wl)))
    a = base.copy(order="F")
    ctx.line_absorption(a, E, "fp32")
    assert np.array_equal(a, base)


def test_deterministic_repeat(ctx):
    P, _, _ = workloads.solar_small(n_lines=500)
    for prec in ("fp64", "fp32"):
        assert np.array_equal(gpu_alpha(ctx, P, prec), gpu_alpha(ctx, P, prec))


def test_float32_alpha_storage(ctx):
    P, _, _ = workloads.solar_small(n_lines=300)
    ref = oracle_alpha(P)
    got = gpu_alpha(ctx, P, "fp32", dtype=np.float32)
    assert got.dtype == np.float32 and rel_err(got, ref) <= FP32_RTOL
    got64 = gpu_alpha(ctx, P, "fp64", dtype=np.float32)
    assert rel_err(got64, ref) <= 1e-6


def test_unsorted_lines(ctx):
    """Device entry point rejects an unsorted list and leaves alpha alone; the Python mirror sorts
    first, as synthesize does."""
    P, _, _ = workloads.solar_small(n_lines=100)
    perm = np.random.default_rng(4).permutation(len(P.lines))
    U = Problem(P.all_wls, P.lines.take(perm), P.species_mass, P.species_is_mol, P.n_div_U,
                P.temps, P.n_e, P.n_HI, P.n_HeI, P.n_H2, P.xi, P.cntm_mode, P.cntm_values,
                P.cntm_knots, P.cutoff_threshold)
    D = DeviceProblem(U)
    a = D.new_alpha("float32")
    with pytest.raises(_lib.KorgError) as e:
        ctx.line_absorption_device(a, D, "fp32")
    assert e.value.code == _lib.KORG_ERR_INVALID_ARGUMENT
    assert not a.numpy().any()


# ----------------------------------------------------------------------------- shapes
def _atm_subset(n_layers):
    atm = Atmosphere.from_model("sun")
    idx = np.linspace(0, atm.n_layers - 1, min(n_layers, atm.n_layers)).astype(int)
    if n_layers > atm.n_layers:  # repeat layers to make a deeper stand-in
        idx = np.resize(np.arange(atm.n_layers), n_layers)
    return atm, idx


@pytest.mark.parametrize("n_layers", [1, 5, 8, 17, 70])
def test_layer_counts(ctx, n_layers):
    atm, idx = _atm_subset(n_layers)
    wls = Wavelengths((5000, 5020))
    ll = workloads.synthetic_linelist(5, 150, 4995.0, 5025.0, (-4.0, 0.5), (0.0, 5.0))
    knots, vals = continuum_grid(atm, float(wls.all_wls[0]), float(wls.all_wls[-1]))
    nd = {k: v[idx] for k, v in atm.n_densities.items()}
    P = Problem.from_korg_args(ll, wls, atm.temps[idx], atm.n_e[idx], nd, atm.partition_fns,
                               1e5 * (1 + 0.1 * np.arange(len(idx))), ContinuumGrid(knots, vals[idx]))
    ref = check_fp64(ctx, P)
    check_fp32(ctx, P, ref)


def test_multi_range_grid(ctx):
    atm = Atmosphere.from_model("sun")
    wls = Wavelengths([(5000, 5005), (5006, 5010, 0.02), (5012.5, 5015, 0.005)])
    ll = workloads.synthetic_linelist(6, 300, 4998.0, 5017.0, (-4.0, 0.5), (0.0, 5.0))
    P = workloads.make_problem(atm, ll, wls)
    ref = check_fp64(ctx, P)
    check_fp32(ctx, P, ref)


def test_single_wavelength(ctx):
    """The reference-wavelength call of synthesize: a one-point grid at 5000 A
    (src/synthesize.jl:266-276)."""
    atm = Atmosphere.from_model("sun")
    wls = Wavelengths([5000.0])
    ll = workloads.synthetic_linelist(7, 200, 4990.0, 5010.0, (-4.0, 0.5), (0.0, 5.0))
    P = workloads.make_problem(atm, ll, wls)
    assert P.n_wl == 1
    ref = check_fp64(ctx, P)
    check_fp32(ctx, P, ref)


def test_molecules_abo_and_missing_perturbers(ctx):
    atm = Atmosphere.from_model("sun")
    wls = Wavelengths((5000, 5010))
    lines = [Line(5001.0, -1.0, "CO", 0.5, 1e8, 1e-4, -7.0),
             Line(5002.0, -0.5, "Fe I", 1.0, 1e8, 1e-5, (1.3917417470792187e-14, 0.227)),
             Line(5003.0, -0.5, "Fe I", 2.0, 1e8, 1e-5, 497.227),
             Line(5004.0, 0.0, "Ti I", 0.3),
             Line(5005.0, -2.0, "FeH", 0.2, 1e7, 0.0, 0.0)]
    ll = LineList.from_lines(lines)
    nd = dict(atm.n_densities)
    L = atm.n_layers
    nd[Species("CO")] = np.full(L, 1e12)
    nd[Species("FeH")] = np.full(L, 1e9)
    pf = dict(atm.partition_fns)
    pf[Species("CO")] = lambda x: 10.0 + 0 * x
    pf[Species("FeH")] = lambda x: 50.0 + 0 * x
    knots, vals = continuum_grid(atm, float(wls.all_wls[0]), float(wls.all_wls[-1]))
    P = Problem.from_korg_args(ll, wls, atm.temps, atm.n_e, nd, pf, 1e5, ContinuumGrid(knots, vals))
    ref = check_fp64(ctx, P)
    check_fp32(ctx, P, ref)


def test_extreme_damping_lines(ctx):
    """Radiative damping up to 1e18 s^-1 (a ~ 1e3 .. 1e7): past a ~ 1e4 the far-wing series
    coefficients leave Float32 range and the Float32 form evaluates the line exactly over its
    whole window.  Finite everywhere and within the Float32 tolerance of the oracle."""
    atm = Atmosphere.from_model("sun")
    wls = Wavelengths((5000, 5010))
    lines = [Line(5000.5 + 0.9 * k, 2.0, "Fe I", 1.0, g, 1e-5, 1e-7)
             for k, g in enumerate([1e12, 1e14, 1e15, 1e16, 1e17, 1e18])]
    lines += workloads.synthetic_linelist(31, 200, 4995.0, 5015.0, (-4.0, 0.5), (0.0, 5.0)).to_lines()
    ll = LineList.from_lines(sorted(lines, key=lambda x: x.wl))
    knots, vals = continuum_grid(atm, float(wls.all_wls[0]), float(wls.all_wls[-1]))
    P = Problem.from_korg_args(ll, wls, atm.temps, atm.n_e, atm.n_densities, atm.partition_fns,
                               1e5, ContinuumGrid(knots, vals))
    ref, lb, ub, _ = oracle_alpha(P, windows=True)
    broad = np.isin(ll.gamma_rad, [1e16, 1e17, 1e18])
    assert np.sum(lb[broad] <= ub[broad]) >= 2  # the extreme lines have windows
    got = gpu_alpha(ctx, P, "fp32")
    assert np.all(np.isfinite(got))
    assert rel_err(got, ref) <= FP32_RTOL, rel_err(got, ref)


# ----------------------------------------------------------------------------- full size
def test_solar_wide_full_size_properties(ctx):
    """300,000 lines x 300,001 points x 56 layers: windows identical to the oracle's window pass
    (index work), the same evaluation count, and the two device forms within the Float32
    tolerance of each other; the Float32 run is bitwise repeatable."""
    P, _, _ = workloads.solar_wide()
    lb, ub, nev = oracle.line_absorption(np.zeros((P.n_layers, P.n_wl), order="F"), P,
                                         mode=oracle.WINDOWS_ONLY, return_windows=True)
    D = DeviceProblem(P)
    a64 = D.new_alpha("float64")
    a32 = D.new_alpha("float32")
    b32 = D.new_alpha("float32")
    ctx.line_absorption_device(a64, D, "fp64")
    glb, gub = ctx.windows(len(P.lines))
    skip = lb > ub
    assert np.array_equal(skip, glb > gub)
    assert np.array_equal(lb[~skip], glb[~skip]) and np.array_equal(ub[~skip], gub[~skip])
    assert ctx.stats()["n_evals"] == nev
    ctx.line_absorption_device(a32, D, "fp32")
    ctx.line_absorption_device(b32, D, "fp32")
    assert np.array_equal(a32.numpy(), b32.numpy())
    err = rel_err(a32.numpy(), a64.numpy())
    assert err <= FP32_RTOL, err


def test_work_counts(ctx):
    """Near-zone, tail and far-wing counts of the Float32 form: the near zone is part of the
    windows, tail evaluations are part of the far zone.  With the far field off every far-zone
    evaluation is a term of the far-wing bfloat16 MFMAs (three split products per group of 4 lines,
    32-point block and 32-layer half: 1.5 NB per line-tile entry and 64-layer block, for tiles of
    32 NB points), each MFMA triple covering 4 lines x 32 points x 32 layers.  With it on (the
    default) part of the pairs become far-field list entries, far fewer MFMAs are issued, and the
    evaluation counts are the same.  The Float64 form reports none of them."""
    P, _, _ = workloads.solar_small()
    saved = ctx.tuning()
    try:
        ctx.set_tuning(**{**saved, "far_field": 1})
        a_off = gpu_alpha(ctx, P, "fp32")
        s = ctx.stats()
        far = s["n_evals"] - s["n_evals_near"]
        assert 0 < s["n_evals_near"] < s["n_evals"]
        assert s["n_evals_near"] % P.n_layers == 0
        assert 0 < s["n_evals_tail"] <= far and s["n_evals_tail"] % P.n_layers == 0
        assert s["n_far_entries"] == 0
        n_lblocks = -(-P.n_layers // 64)
        entries_nb = 2 * s["n_wing_mfma"] / (3 * n_lblocks)  # line-tile entries x NB
        assert entries_nb == round(entries_nb) and entries_nb > 0
        assert 4096 * s["n_wing_mfma"] / 3 >= far
        # a line with a W-point window meets at most W/(32 NB) + 2 tiles, each costing 1.5 NB
        # instructions per 64-layer block; NB <= 4
        lb, ub = ctx.windows(len(P.lines))
        w = (ub - lb + 1)[ub >= lb]
        assert s["n_wing_mfma"] <= 1.5 * n_lblocks * np.sum(w / 32 + 8)
        ctx.set_tuning(**{**saved, "far_field": 2})
        a_on = gpu_alpha(ctx, P, "fp32")
        s2 = ctx.stats()
        assert s2["n_far_entries"] > 0 and s2["n_wing_mfma"] < s["n_wing_mfma"]
        for k in ("n_evals", "n_evals_near", "n_evals_tail"):
            assert s2[k] == s[k], k
        assert rel_err(a_on, a_off) <= FP32_RTOL
    finally:
        ctx.set_tuning(**saved)
    gpu_alpha(ctx, P, "fp64")
    s = ctx.stats()
    assert s["n_evals_near"] == 0 and s["n_wing_mfma"] == 0 and s["n_evals_tail"] == 0
    assert s["n_far_entries"] == 0


def test_invalid_species_index_rejected(ctx):
    """A species index outside the species table is reported, and alpha is left alone (the
    device checks the index before any read through it)."""
    P, _, _ = workloads.solar_small(n_lines=50)
    bad = P.lines.take(np.arange(len(P.lines)))
    bad.species_index = bad.species_index.copy()
    bad.species_index[7] = len(P.species_mass)
    B = Problem(P.all_wls, bad, P.species_mass, P.species_is_mol, P.n_div_U, P.temps, P.n_e,
                P.n_HI, P.n_HeI, P.n_H2, P.xi, P.cntm_mode, P.cntm_values, P.cntm_knots,
                P.cutoff_threshold)
    a = np.zeros((P.n_layers, P.n_wl), order="F")
    with pytest.raises(_lib.KorgError) as e:
        ctx.line_absorption(a, B, "fp32")
    assert e.value.code == _lib.KORG_ERR_INVALID_ARGUMENT
    assert not a.any()
