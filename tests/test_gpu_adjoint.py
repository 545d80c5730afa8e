"""The device adjoint (korg_line_absorption_vjp / _device, through the C ABI) against the Float64
oracle's adjoint, which tests/test_adjoint.py pins against central differences.  Needs a GPU.

Tolerance: the per-layer gradients agree with the oracle to 1e-9 relative to the largest
|gradient| of the quantity (both are Float64; they differ in transcendental last bits and in
summation order), and the end-to-end chain of line_absorption_vjp_ matches central differences
of the device's Float64 forward to 1e-5 relative.
"""
import numpy as np
import pytest

from oracle import oracle
from korg_line import (Context, DeviceProblem, Line, LineList, Problem, Species, Wavelengths,
                       line_absorption_vjp_, workloads)
from korg_line.atmosphere import Atmosphere, continuum_grid
from korg_line.hipmem import DeviceArray
from korg_line.problem import ContinuumGrid

pytestmark = pytest.mark.gpu

NAMES = ("T", "n_e", "xi", "n_HI", "n_HeI", "n_H2", "n_div_U")
RTOL = 1e-9


def random_G(P, seed=3, dtype=np.float64):
    rng = np.random.default_rng(seed)
    return np.asfortranarray(rng.standard_normal((P.n_layers, P.n_wl)).astype(dtype))


def check_vjp(ctx, P, G):
    got = ctx.line_absorption_vjp(G, P)
    ref = oracle.line_absorption_vjp(np.asfortranarray(G.astype(np.float64)), P)
    for k in NAMES:
        scale = np.max(np.abs(ref[k])) if np.any(ref[k]) else 1.0
        np.testing.assert_allclose(got[k], ref[k], rtol=RTOL, atol=RTOL * scale, err_msg=k)
    return got


def test_vjp_solar_small(ctx):
    P, _, _ = workloads.solar_small()
    check_vjp(ctx, P, random_G(P))


def test_vjp_float32_G(ctx):
    P, _, _ = workloads.solar_small(n_lines=500)
    check_vjp(ctx, P, random_G(P, seed=4, dtype=np.float32))


def test_vjp_device_equals_host_and_repeats(ctx):
    P, _, _ = workloads.solar_small(n_lines=800)
    G = random_G(P, seed=5)
    host = ctx.line_absorption_vjp(G, P)
    D = DeviceProblem(P)
    Gd = DeviceArray.from_numpy(np.ascontiguousarray(G.T))  # (n_wl, n_layers): alpha's layout
    for _ in range(2):
        out = {k: DeviceArray.zeros(v.shape, np.float64) for k, v in host.items()}
        ctx.line_absorption_vjp_device(Gd, D, out)
        for k in NAMES:
            assert np.array_equal(out[k].numpy(), host[k]), k


def test_vjp_empty_linelist(ctx):
    P, _, _ = workloads.solar_small(n_lines=50)
    E = Problem(P.all_wls, P.lines.take(np.arange(0)), P.species_mass, P.species_is_mol,
                P.n_div_U, P.temps, P.n_e, P.n_HI, P.n_HeI, P.n_H2, P.xi, P.cntm_mode,
                P.cntm_values, P.cntm_knots, P.cutoff_threshold)
    got = ctx.line_absorption_vjp(random_G(E), E)
    assert all(not np.any(v) for v in got.values())


@pytest.mark.parametrize("n_layers", [1, 17, 70])
def test_vjp_layer_counts(ctx, n_layers):
    atm = Atmosphere.from_model("sun")
    idx = np.resize(np.arange(atm.n_layers), n_layers) if n_layers > atm.n_layers else \
        np.linspace(0, atm.n_layers - 1, n_layers).astype(int)
    wls = Wavelengths((5000, 5020))
    ll = workloads.synthetic_linelist(5, 150, 4995.0, 5025.0, (-4.0, 0.5), (0.0, 5.0))
    knots, vals = continuum_grid(atm, float(wls.all_wls[0]), float(wls.all_wls[-1]))
    nd = {k: v[idx] for k, v in atm.n_densities.items()}
    P = Problem.from_korg_args(ll, wls, atm.temps[idx], atm.n_e[idx], nd, atm.partition_fns,
                               1e5 * (1 + 0.1 * np.arange(len(idx))),
                               ContinuumGrid(knots, vals[idx]))
    check_vjp(ctx, P, random_G(P, seed=n_layers))


def test_vjp_molecules_and_abo(ctx):
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
    check_vjp(ctx, P, random_G(P, seed=9))


def test_vjp_wide_window_subset(ctx):
    """Lines from the 300,000-line list over the full 4000-7000 A grid: windows of thousands of
    points, split over several work items per line."""
    P, _, _ = workloads.solar_wide()
    sub = np.arange(0, len(P.lines), 300)  # 1,000 lines
    S = Problem(P.all_wls, P.lines.take(sub), P.species_mass, P.species_is_mol, P.n_div_U,
                P.temps, P.n_e, P.n_HI, P.n_HeI, P.n_H2, P.xi, P.cntm_mode, P.cntm_values,
                P.cntm_knots, P.cutoff_threshold)
    check_vjp(ctx, S, random_G(S, seed=11, dtype=np.float32))


def test_vjp_chain_matches_central_differences(ctx):
    """The mirror line_absorption_vjp_ (with n/U through the partition functions) against central
    differences of the device's Float64 forward, directional derivative in T, n_e and xi."""
    atm = Atmosphere.from_model("sun")
    wls = Wavelengths((5000, 5030))
    ll = workloads.synthetic_linelist(12, 120, 4995.0, 5035.0, (-4.0, 0.5), (0.0, 5.0))
    knots, vals = continuum_grid(atm, float(wls.all_wls[0]), float(wls.all_wls[-1]))
    cg = ContinuumGrid(knots, vals)
    xi = 1e5 * np.ones(atm.n_layers)
    rng = np.random.default_rng(13)
    G = np.asfortranarray(rng.standard_normal((atm.n_layers, wls.all_wls.size)))
    grads = line_absorption_vjp_(G, ll, wls, atm.temps, atm.n_e, atm.n_densities,
                                 atm.partition_fns, xi, cg, ctx=ctx)

    def objective(temps, n_e, xi_):
        P = Problem.from_korg_args(ll, wls, temps, n_e, atm.n_densities, atm.partition_fns, xi_, cg)
        a = np.zeros((P.n_layers, P.n_wl), order="F")
        ctx.line_absorption(a, P, "fp64")
        lb, ub = ctx.windows(len(P.lines))
        return float(np.sum(G * a)), lb, ub

    for name, base, g in (("temps", atm.temps, grads["temps"]), ("n_e", atm.n_e, grads["n_e"]),
                          ("xi", xi, grads["xi"])):
        d = rng.uniform(0.5, 1.5, base.shape) * base
        for h in (1e-6, 1e-7, 1e-8):
            args = {"temps": atm.temps, "n_e": atm.n_e, "xi_": xi}
            key = "xi_" if name == "xi" else name
            fp, lbp, ubp = objective(**{**args, key: base + h * d})
            fm, lbm, ubm = objective(**{**args, key: base - h * d})
            if np.array_equal(lbp, lbm) and np.array_equal(ubp, ubm):
                break
        else:
            pytest.fail("every step moved a window")
        fd = (fp - fm) / (2 * h)
        pred = float(np.sum(g * d))
        assert abs(pred - fd) <= 1e-5 * abs(fd) + 1e-12 * float(np.sum(np.abs(g * d))), (name, pred, fd)
