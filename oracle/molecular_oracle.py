"""ORACLE -- test infrastructure only: numpy restatement of Korg's
``interpolate_molecular_cross_sections!`` (``src/molecular_cross_sections.jl:99-114``), the
interpolator of ``:76-78`` (``extrapolate(interpolate!((vmic, logT, wls), vals,
(Gridded(Linear()), ...)), 0.0)``): linear in each gridded dimension, zero outside the table.

Tables are ``korg_line.molecular.MolecularCrossSection``-like objects (``vmic_vals``,
``log_temp_vals``, ``wls.all_wls``, ``sigma[iv, j, iT]``, ``species``).  Only the tests use this
module, as the checker of the device library.
"""
import numpy as np


def _cells(knots, x):
    """Interval index i (0 .. n-2) and weight of each x in the ascending knots; -1 outside."""
    knots = np.asarray(knots, dtype=np.float64)
    x = np.asarray(x, dtype=np.float64)
    i = np.clip(np.searchsorted(knots, x, side="right") - 1, 0, knots.size - 2)
    w = (x - knots[i]) / (knots[i + 1] - knots[i])
    inside = (x >= knots[0]) & (x <= knots[-1])
    return np.where(inside, i, -1), w


def cross_section(table, vmic, logT, wls):
    """sigma(vmic, logT, wls) of one table for one (vmic, logT) and an array of wavelengths."""
    iv, wv = _cells(table.vmic_vals, vmic)
    iT, wT = _cells(table.log_temp_vals, logT)
    ij, wj = _cells(table.wls.all_wls, wls)
    out = np.zeros(np.shape(wls))
    if iv < 0 or iT < 0:
        return out
    ok = ij >= 0
    j = ij[ok]
    s = table.sigma
    val = 0.0
    for dv, fv in ((0, 1 - wv), (1, wv)):
        for dT, fT in ((0, 1 - wT), (1, wT)):
            val = val + fv * fT * ((1 - wj[ok]) * s[iv + dv, j, iT + dT] +
                                   wj[ok] * s[iv + dv, j + 1, iT + dT])
    out[ok] = val
    return out


def interpolate_molecular_cross_sections(alpha, tables, wls, temps, vmic, n_densities):
    """alpha[l, :] += sigma(vmic_l, log10 T_l, wls) * n_densities[species][l] for every table
    (src/molecular_cross_sections.jl:106-113); vmic in km/s, scalar or per layer."""
    wls = np.asarray(getattr(wls, "all_wls", wls), dtype=np.float64)
    for t in tables:
        n = np.broadcast_to(np.asarray(n_densities[t.species], dtype=np.float64), (len(temps),))
        for i, T in enumerate(temps):
            vm = vmic if np.ndim(vmic) == 0 else vmic[i]
            alpha[i, :] += cross_section(t, vm, np.log10(T), wls) * n[i]
    return alpha
