"""Seeded synthetic workloads for the line-opacity path.

Random numbers come from a counter-based generator (splitmix64 of a per-stream key plus the
element index, top 53 bits as a uniform in [0, 1)), so a list of any length is reproducible
element by element in any language.

Workloads (lines sorted by wavelength, as ``synthesize`` sorts them,
``src/synthesize.jl:191-195``):

* ``solar_small`` -- solar atmosphere, 5000-5100 A at 0.01 A (10,001 points), 2,000 lines:
  species 60% Fe I / 15% Ti I / 10% Ni I / 10% Fe II / 5% Ca I, log gf ~ U[-4, 0.5],
  E_lower ~ U[0, 5] eV, all broadening from Korg's default recipes, vmic 1 km/s.
* ``solar_wide`` -- same atmosphere, 4000-7000 A at 0.01 A (300,001 points), 300,000 lines:
  log gf ~ U[-6, 0.5], E_lower ~ U[0, 8] eV, log10 gamma_rad ~ U[6, 9],
  log10 gamma_stark ~ U[-7, -4], vdW from the default recipe.
* ``batched_stars`` -- per-star atmospheres, wavelength windows and line subsets.
"""
import numpy as np

from .atmosphere import Atmosphere, continuum_grid
from .linelist import LineList, _pow10
from .problem import ContinuumGrid, Problem
from .species import Species
from .wavelengths import Wavelengths

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _splitmix64(x):
    x = np.asarray(x, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = x + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def uniform(seed, stream, n, offset=0):
    """n uniforms in [0, 1) for (seed, stream), elements offset..offset+n-1."""
    key = _splitmix64(_splitmix64(np.uint64(seed)) ^ np.uint64(stream))
    with np.errstate(over="ignore"):
        ctr = key + np.arange(offset, offset + n, dtype=np.uint64)
    return (_splitmix64(ctr) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


SPECIES_MIX = (("Fe I", 0.60), ("Ti I", 0.15), ("Ni I", 0.10), ("Fe II", 0.10), ("Ca I", 0.05))

# stream ids
_S_WL, _S_SPEC, _S_GF, _S_EL, _S_GRAD, _S_GSTARK = 1, 2, 3, 4, 5, 6


def synthetic_linelist(seed, n, wl_lo, wl_hi, log_gf_range, E_lower_range, log_gamma_rad=None,
                       log_gamma_stark=None, offset=0):
    """Seeded synthetic line list (wavelengths in A), sorted by wavelength."""
    species = [Species(s) for s, _ in SPECIES_MIX]
    cum = np.cumsum([p for _, p in SPECIES_MIX])
    wl = wl_lo + (wl_hi - wl_lo) * uniform(seed, _S_WL, n, offset)
    sidx = np.searchsorted(cum, uniform(seed, _S_SPEC, n, offset), side="right").astype(np.int32)
    sidx = np.minimum(sidx, len(species) - 1)
    log_gf = log_gf_range[0] + (log_gf_range[1] - log_gf_range[0]) * uniform(seed, _S_GF, n, offset)
    E_lower = E_lower_range[0] + (E_lower_range[1] - E_lower_range[0]) * \
        uniform(seed, _S_EL, n, offset)
    grad = None if log_gamma_rad is None else _pow10(
        log_gamma_rad[0] + (log_gamma_rad[1] - log_gamma_rad[0]) * uniform(seed, _S_GRAD, n, offset))
    gstark = None if log_gamma_stark is None else _pow10(
        log_gamma_stark[0] + (log_gamma_stark[1] - log_gamma_stark[0]) *
        uniform(seed, _S_GSTARK, n, offset))
    ll = LineList.from_arrays(wl, log_gf, species, sidx, E_lower, gamma_rad=grad,
                              gamma_stark=gstark)
    return ll.sorted()


def make_problem(atm, linelist, wls, vmic_kms=1.0, cutoff_threshold=3e-4, cntm="grid"):
    """Assemble the inputs of one ``line_absorption!`` call on a stand-in atmosphere."""
    knots, vals = continuum_grid(atm, float(wls.all_wls[0]), float(wls.all_wls[-1]))
    grid = ContinuumGrid(knots, vals)
    alpha_cntm = grid if cntm == "grid" else [grid.layer(l) for l in range(atm.n_layers)]
    return Problem.from_korg_args(linelist, wls, atm.temps, atm.n_e, atm.n_densities,
                                  atm.partition_fns, vmic_kms * 1e5, alpha_cntm,
                                  cutoff_threshold=cutoff_threshold)


def solar_small(seed=1, n_lines=2000, cntm="grid"):
    atm = Atmosphere.from_model("sun")
    wls = Wavelengths((5000, 5100))
    ll = synthetic_linelist(seed, n_lines, 5000.0, 5100.0, (-4.0, 0.5), (0.0, 5.0))
    return make_problem(atm, ll, wls, 1.0, cntm=cntm), atm, wls


def solar_wide(seed=2, n_lines=300_000, wl_range=(4000, 7000), cntm="grid"):
    atm = Atmosphere.from_model("sun")
    wls = Wavelengths(wl_range)
    ll = synthetic_linelist(seed, n_lines, float(wl_range[0]), float(wl_range[1]), (-6.0, 0.5),
                            (0.0, 8.0), log_gamma_rad=(6.0, 9.0), log_gamma_stark=(-7.0, -4.0))
    return make_problem(atm, ll, wls, 1.0, cntm=cntm), atm, wls


def solar_wide_subset(seed=2, n_lines=300_000, wl_range=(4000, 7000), sub_range=(5000, 5010)):
    """The ``solar_wide`` problem restricted to the lines and grid of a sub-range (tests)."""
    atm = Atmosphere.from_model("sun")
    ll = synthetic_linelist(seed, n_lines, float(wl_range[0]), float(wl_range[1]), (-6.0, 0.5),
                            (0.0, 8.0), log_gamma_rad=(6.0, 9.0), log_gamma_stark=(-7.0, -4.0))
    lo, hi = sub_range[0] * 1e-8, sub_range[1] * 1e-8
    keep = np.nonzero((ll.wl >= lo) & (ll.wl <= hi))[0]
    wls = Wavelengths(sub_range)
    return make_problem(atm, ll.take(keep), wls, 1.0), atm, wls


# stream ids of the batched workload's per-star draws
_S_TEFF, _S_LOGG, _S_MH, _S_WIDTH, _S_START, _S_NLINES = 101, 102, 103, 104, 105, 106


def batched_stars(n_stars=256, seed=3, scale=1.0, wl_range=(4000.0, 7000.0)):
    """Many stars with uneven line counts and wavelength ranges.

    Star k: Teff ~ U[4000, 6500] K, log g ~ U[1, 5], [M/H] ~ U[-2.5, 0.5] (a scaled solar
    structure, ``Atmosphere.scaled``), a window of width ~ U[50, 1000] A * scale inside
    ``wl_range`` at 0.01 A, and its own seeded list of ~ U[5,000, 150,000] * scale lines over the
    window, drawn from the solar_small distribution (log gf ~ U[-4, 0.5], E_lower ~ U[0, 5] eV).
    Every list shares one species table.  Returns a list of (Problem, Atmosphere, Wavelengths).
    """
    u = lambda stream: uniform(seed, stream, n_stars)
    teff = 4000.0 + 2500.0 * u(_S_TEFF)
    logg = 1.0 + 4.0 * u(_S_LOGG)
    mh = -2.5 + 3.0 * u(_S_MH)
    width = (50.0 + 950.0 * u(_S_WIDTH)) * scale
    lo, hi = wl_range
    start = lo + (hi - lo - width) * u(_S_START)
    n_lines = np.maximum(1, np.round((5_000 + 145_000 * u(_S_NLINES)) * scale)).astype(int)
    out = []
    for k in range(n_stars):
        atm = Atmosphere.scaled(float(teff[k]), float(logg[k]), float(mh[k]))
        a = round(float(start[k]), 2)
        b = round(a + float(width[k]), 2)
        wls = Wavelengths((a, b))
        ll = synthetic_linelist(seed * 1_000_003 + k, int(n_lines[k]), a, b, (-4.0, 0.5), (0.0, 5.0))
        out.append((make_problem(atm, ll, wls, 1.0), atm, wls))
    return out



def fit_stars(n_stars=64, seed=7, n_lines=100_000, wl_range=(5000.0, 6000.0)):
    """The batched fit workload: ``n_stars`` stars drawn as ``batched_stars`` draws them (Teff
    ~ U[4000, 6500] K, log g ~ U[1, 5], [M/H] ~ U[-2.5, 0.5]), all on one grid (5000-6000 A at
    0.01 A: 100,001 points) with one shared seeded list of ``n_lines`` lines from the solar_small
    distribution.  Returns (problems, atmospheres, linelist, wavelengths)."""
    u = lambda stream: uniform(seed, stream, n_stars)
    teff = 4000.0 + 2500.0 * u(_S_TEFF)
    logg = 1.0 + 4.0 * u(_S_LOGG)
    mh = -2.5 + 3.0 * u(_S_MH)
    wls = Wavelengths(wl_range)
    ll = synthetic_linelist(seed, n_lines, float(wl_range[0]), float(wl_range[1]), (-4.0, 0.5),
                            (0.0, 5.0))
    atms = [Atmosphere.scaled(float(teff[k]), float(logg[k]), float(mh[k])) for k in range(n_stars)]
    probs = [make_problem(a, ll, wls, 1.0) for a in atms]
    return probs, atms, ll, wls


def fit_reference(atm):
    """The anchored scheme's reference optical depth and absorption (at 5000 A) of a stand-in
    atmosphere, per layer."""
    return 10.0 ** atm.lgTau5, atm.continuum(np.array([5000e-8]))[:, 0]
