"""The Float32 far wing's far field (korg_tuning.far_field): where a 64-point sub-tile lies
inside a line's window, outside its near zone and more than 8 half-widths from its centre, the
line's series is summed as a 12-term expansion about the sub-tile's centre.  Checked through the C
ABI against the Float64 oracle and against the same call with every (line, tile) pair term by
term (far_field = 1), on workloads whose far fields are large (wide windows, lines outside the
grid), on an unevenly spaced grid (which takes no far field), and for every far-wing tile shape.
Needs a GPU.
"""
import numpy as np
import pytest

from oracle import oracle
from korg_line import Wavelengths, workloads
from korg_line.atmosphere import Atmosphere

pytestmark = pytest.mark.gpu

FP32_RTOL = 1e-4


def rel_err(a, ref, floor_frac=1e-8):
    ref = np.asarray(ref, dtype=np.float64)
    a = np.asarray(a, dtype=np.float64)
    m = np.abs(ref) >= floor_frac * np.abs(ref).max()
    return float(np.max(np.abs(a[m] - ref[m]) / np.abs(ref[m])))


def run(ctx, P, far_field, **tuning):
    saved = ctx.tuning()
    try:
        ctx.set_tuning(**{**saved, **tuning, "far_field": far_field})
        a = np.zeros((P.n_layers, P.n_wl), order="F")
        ctx.line_absorption(a, P, "fp32")
        return a, ctx.stats()
    finally:
        ctx.set_tuning(**saved)


def oracle_alpha(P):
    a = np.zeros((P.n_layers, P.n_wl), order="F")
    oracle.line_absorption(a, P)
    return a


def strong_lines_problem():
    """60 A at 0.01 A with 6,000 lines from 40 A either side of it: strong lines (log gf up to
    0.5, broad Stark widths) whose windows span thousands of points, many centred outside the
    grid."""
    atm = Atmosphere.from_model("sun")
    wls = Wavelengths((5000, 5060))
    ll = workloads.synthetic_linelist(11, 6000, 4960.0, 5100.0, (-3.0, 0.5), (0.0, 3.0),
                                      log_gamma_stark=(-5.0, -3.5))
    return workloads.make_problem(atm, ll, wls)


@pytest.mark.parametrize("make", ["solar_small", "strong_lines"])
def test_far_field_matches_oracle_and_term_by_term(ctx, make):
    P = workloads.solar_small()[0] if make == "solar_small" else strong_lines_problem()
    ref = oracle_alpha(P)
    on, s_on = run(ctx, P, 2)
    off, s_off = run(ctx, P, 1)
    assert s_on["n_far_entries"] > 0 and s_off["n_far_entries"] == 0
    assert s_on["n_evals"] == s_off["n_evals"]
    e_on, e_off, e_pair = rel_err(on, ref), rel_err(off, ref), rel_err(on, off)
    print(f"{make}: far field vs oracle {e_on:.3e}, term by term {e_off:.3e}, between {e_pair:.3e}; "
          f"far entries {s_on['n_far_entries']:.0f}")
    assert e_on <= FP32_RTOL and e_off <= FP32_RTOL
    assert e_pair <= 5e-5


def test_far_field_repeatable(ctx):
    P = strong_lines_problem()
    a, _ = run(ctx, P, 2)
    b, _ = run(ctx, P, 2)
    assert np.array_equal(a, b)


def test_uneven_grid_takes_no_far_field(ctx):
    """A grid of three ranges with gaps and different steps: no star-wide even spacing, so no line
    takes a far field and the result is bitwise the term-by-term one."""
    atm = Atmosphere.from_model("sun")
    wls = Wavelengths([(5000, 5020), (5021, 5040, 0.02), (5042.5, 5060, 0.005)])
    ll = workloads.synthetic_linelist(6, 3000, 4990.0, 5070.0, (-3.0, 0.5), (0.0, 3.0))
    P = workloads.make_problem(atm, ll, wls)
    on, s_on = run(ctx, P, 2)
    off, _ = run(ctx, P, 1)
    assert s_on["n_far_entries"] == 0
    assert np.array_equal(on, off)
    assert rel_err(on, oracle_alpha(P)) <= FP32_RTOL


@pytest.mark.parametrize("tile_points,waves_per_tile", [(32, 1), (32, 4), (64, 1), (64, 2),
                                                        (128, 1), (128, 2)])
def test_far_field_every_tile_shape(ctx, tile_points, waves_per_tile):
    """The far-wing tiles of 32, 64 and 128 points (one sub-tile spans two 32-point tiles; a
    128-point tile spans two sub-tiles, masked per 32-point block)."""
    P = strong_lines_problem()
    ref = oracle_alpha(P)
    a, s = run(ctx, P, 2, tile_points=tile_points, waves_per_tile=waves_per_tile)
    assert s["n_far_entries"] > 0
    assert rel_err(a, ref) <= FP32_RTOL
