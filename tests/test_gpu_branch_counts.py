"""The count_branches diagnostic (evaluations per Hunger branch, the figures behind the
near-zone kernel's algorithmic flops) against a classification of every (line, layer, window
point) made here from the oracle's Doppler widths, Stark and van der Waals scalings and windows.

A layer's core points are those strictly within its half-width hw of the line centre: 5 Doppler
units (sigma sqrt 2) for a <= 0.2 (branch B, the rest of the window A) and (3.2 - a) units for
0.2 < a <= 1.4 (branch C, the rest D); for a > 1.4 every point is D
(src/line_absorption.jl:286-310).  Needs a GPU.
"""
import numpy as np
import pytest

from oracle import oracle
from korg_line import workloads
from korg_line.constants import c_cgs, amu_cgs

pytestmark = pytest.mark.gpu


def expected_counts(P):
    """(A, B, C, D) evaluation counts over the oracle's windows."""
    a_ = np.zeros((P.n_layers, P.n_wl), order="F")
    lb, ub, _ = oracle.line_absorption(a_, P, return_windows=True)
    ll = P.lines
    a_H, a_He, a_H2 = 4.50711, 1.38375, 5.503
    m_H, m_He = 1.008 * amu_cgs, 4.003 * amu_cgs
    c_He = (a_He / a_H) ** 0.4 * (m_He / m_H) ** -0.3
    c_H2 = (a_H2 / a_H) ** 0.4 * 2.0 ** -0.3
    n_eff = P.n_HI + c_He * P.n_HeI + c_H2 * P.n_H2
    counts = np.zeros(4, dtype=np.int64)
    wls = P.all_wls
    for i in range(len(ll)):
        if ub[i] < lb[i]:
            continue
        lam0 = ll.wl[i]
        s = ll.species_index[i]
        m, mol = P.species_mass[s], P.species_is_mol[s] != 0
        pts = wls[lb[i] - 1:ub[i]]
        for l in range(P.n_layers):
            T = P.temps[l]
            sigma = oracle.doppler_width(lam0, T, m, P.xi[l])
            G = ll.gamma_rad[i]
            if not mol:
                G += P.n_e[l] * oracle.scaled_stark(ll.gamma_stark[i], T)
                G += n_eff[l] * oracle.scaled_vdW((ll.vdW1[i], ll.vdW2[i]), m, T)
            gamma = G * (lam0 * lam0) / (c_cgs * (4.0 * np.pi))
            inv = 1.0 / (sigma * 1.4142135623730951)
            a = gamma * inv
            hw = 5.0 / inv if a <= 0.2 else ((3.2 - a) / inv if a <= 1.4 else -1.0)
            core = int(np.count_nonzero((pts > lam0 - hw) & (pts < lam0 + hw))) if hw >= 0 else 0
            wing = pts.size - core
            if a <= 0.2:
                counts[0] += wing
                counts[1] += core
            else:
                counts[2] += core
                counts[3] += wing
    return counts


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_branch_counts_match_classification(ctx, precision):
    P = workloads.solar_small(n_lines=300)[0]
    want = expected_counts(P)
    saved = ctx.tuning()
    try:
        ctx.set_tuning(**{**saved, "count_branches": 1})
        a = np.zeros((P.n_layers, P.n_wl), order="F")
        ctx.line_absorption(a, P, precision)
        st = ctx.stats()
    finally:
        ctx.set_tuning(**saved)
    got = np.array(st["n_evals_branch"], dtype=np.float64)
    print(f"{precision}: device {got.astype(np.int64)}, classified {want}")
    assert want.sum() > 0 and want[1] > 0
    assert got.sum() == st["n_evals"]
    # (a core edge that falls on a grid point within an ulp could flip one count)
    assert np.all(np.abs(got - want) <= 2), (got, want)
    if precision == "fp32":
        assert 0 <= st["n_evals_near_A"] <= st["n_evals_near"]
