"""Stand-in inputs for the line-opacity path: layers, number densities, partition functions and
continuum opacity.

The reference computes these upstream of ``line_absorption!`` (``chemical_equilibrium``,
``total_continuum_absorption`` and ``interpolate_marcs``: ``src/synthesize.jl:211-240``), which is
outside the accelerated path.  This module builds physically plausible values of the same shapes
from the MARCS model atmospheres that the reference ships as test data (``test/data/*.mod``):

* T, P_e, P_g and the H I / H2 partial pressures of each layer come from the model file;
* metals are split into I/II/III by the Saha equation with Barklem & Collet (2016) partition
  functions (linear in ln T between table points), H II likewise;
* the continuum is H- bound-free + free-free (Gray 2005, eqs 8.11-8.13) plus Thomson scattering,
  tabulated on the 1 A grid that ``synthesize`` uses (``src/synthesize.jl:184-188``).

It is a stand-in, not Korg's chemistry: both the device path and the oracle consume exactly the
same arrays, so parity does not depend on it.
"""
import math
import os

import numpy as np

from .constants import (electron_mass_cgs, hplanck_cgs, kboltz_cgs, kboltz_eV)
from .species import Species, atomic_numbers

_DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")


def _read_model(name):
    rows, abund = [], None
    with open(os.path.join(_DATA, f"{name}_layers.tsv")) as f:
        for line in f:
            if line.startswith("# abundances"):
                abund = np.array([float(x) for x in line.split("\t")[1].split()])
            elif line.startswith("#") or line.startswith("k\t"):
                continue
            else:
                rows.append([float(x) for x in line.split()])
    a = np.array(rows)
    return {"lgTau5": a[:, 1], "T": a[:, 2], "Pe": a[:, 3], "Pg": a[:, 4], "lgP_HI": a[:, 5],
            "lgP_H2": a[:, 6], "A_X": abund}


def _read_partition():
    table = {}
    with open(os.path.join(_DATA, "atomic_partition.tsv")) as f:
        for line in f:
            if line.startswith("#"):
                continue
            t = line.split()
            table[t[0]] = np.array([float(x) for x in t[1:]])
    T = table.pop("T_K")
    return T, table


_PF_T, _PF = _read_partition()
_PF_LNT = np.log(_PF_T)
_IONIZATION = {}
with open(os.path.join(_DATA, "ionization_energies.tsv")) as _f:
    for _l in _f:
        if not _l.startswith("#"):
            _t = _l.split()
            _IONIZATION[_t[1]] = (float(_t[2]), float(_t[3]), float(_t[4]))


def partition_function(name):
    """U(ln T) for a species key like ``"Fe_I"`` (linear in ln T, clamped at the table ends)."""
    vals = _PF[name]

    def U(lnT):
        return np.interp(lnT, _PF_LNT, vals)
    return U


METALS = ("Fe", "Ti", "Ni", "Ca")


class Atmosphere:
    """Layered stand-in atmosphere with the quantities ``line_absorption!`` consumes.

    Attributes (arrays of length ``n_layers``): ``temps``, ``n_e``, ``lgTau5`` and
    ``n_densities`` (dict ``Species -> array``) for H I, He I, H2 and the I/II/III stages of
    Fe, Ti, Ni, Ca.  ``partition_fns`` maps the same species to callables of ln T.
    """

    def __init__(self, temps, Pe, Pg, lgP_HI, lgP_H2, A_X, lgTau5):
        self.temps = np.asarray(temps, dtype=np.float64)
        self.lgTau5 = np.asarray(lgTau5, dtype=np.float64)
        kT = kboltz_cgs * self.temps
        self.n_e = np.asarray(Pe) / kT
        n_HI = np.power(10.0, lgP_HI) / kT
        n_H2 = np.power(10.0, lgP_H2) / kT
        theta_saha = (2 * math.pi * electron_mass_cgs * kT / hplanck_cgs ** 2) ** 1.5
        lnT = np.log(self.temps)
        # hydrogen ionisation from Saha with U(H I) = 2, U(H II) = 1
        r_H = 2 * 1.0 / 2.0 * theta_saha * np.exp(-13.5984 / (kboltz_eV * self.temps)) / self.n_e
        n_Htot = n_HI * (1 + r_H) + 2 * n_H2
        nd = {Species("H I"): n_HI, Species("H2"): n_H2,
              Species("He I"): np.power(10.0, A_X[1] - 12) * n_Htot}
        pf = {Species("H I"): (lambda lnT: 2.0 + 0 * lnT), Species("He I"): partition_function("He_I")}
        for el in METALS:
            Z = atomic_numbers[el]
            ntot = np.power(10.0, A_X[Z - 1] - 12) * n_Htot
            U1, U2, U3 = (partition_function(f"{el}_{s}")(lnT) for s in ("I", "II", "III"))
            chi1, chi2, _ = _IONIZATION[el]
            r1 = 2 * U2 / U1 * theta_saha * np.exp(-chi1 / (kboltz_eV * self.temps)) / self.n_e
            r2 = 2 * U3 / U2 * theta_saha * np.exp(-chi2 / (kboltz_eV * self.temps)) / self.n_e
            n1 = ntot / (1 + r1 * (1 + r2))
            for stage, n in (("I", n1), ("II", n1 * r1), ("III", n1 * r1 * r2)):
                sp = Species(f"{el} {stage}")
                nd[sp] = n
                pf[sp] = partition_function(f"{el}_{stage}")
        self.n_densities = nd
        self.partition_fns = pf
        self.Pe = np.asarray(Pe, dtype=np.float64)
        self.n_HI = n_HI

    @property
    def n_layers(self):
        return self.temps.size

    @classmethod
    def from_model(cls, name="sun"):
        m = _read_model(name)
        return cls(m["T"], m["Pe"], m["Pg"], m["lgP_HI"], m["lgP_H2"], m["A_X"], m["lgTau5"])

    @classmethod
    def scaled(cls, Teff, logg, M_H, base="sun", abundances=None):
        """Synthetic star: the solar model's structure scaled to (Teff, log g, [M/H]).

        T scales with Teff/5777, pressures with sqrt(g/g_sun), metal abundances by 10^[M/H], and
        ``abundances`` ({element symbol: [X/H] offset}) moves single elements further.  This is a
        stand-in for ``interpolate_marcs``, whose atmosphere grids are not available offline.
        """
        m = _read_model(base)
        tscale = Teff / 5777.0
        pscale = math.sqrt(10.0 ** (logg - 4.44))
        A = m["A_X"].copy()
        A[2:] += M_H
        for el, dx in (abundances or {}).items():
            A[atomic_numbers[el] - 1] += dx
        return cls(m["T"] * tscale, m["Pe"] * pscale * tscale ** 6, m["Pg"] * pscale,
                   m["lgP_HI"] + math.log10(pscale), m["lgP_H2"] + math.log10(pscale), A,
                   m["lgTau5"])

    def continuum(self, wl_cm):
        """Continuum opacity [cm^-1], shape (n_layers, len(wl_cm)): H- bf+ff and Thomson."""
        lam = np.asarray(wl_cm, dtype=np.float64) * 1e8  # A
        theta = (5040.0 / self.temps)[:, None]
        Pe = self.Pe[:, None]
        a = (1.99654, -1.18267e-5, 2.64243e-6, -4.40524e-10, 3.23992e-14, -1.39568e-18,
             2.78701e-23)
        alpha_bf = np.zeros_like(lam)
        for k in range(6, -1, -1):
            alpha_bf = alpha_bf * lam + a[k]
        alpha_bf = np.where(lam < 16300.0, np.maximum(alpha_bf, 0.0) * 1e-18, 0.0)
        chi_lam = 1.2398e4 / lam
        k_bf = 4.158e-10 * alpha_bf * Pe * theta ** 2.5 * 10.0 ** (0.754 * theta) * \
            (1 - 10.0 ** (-chi_lam * theta))
        L = np.log10(lam)
        f0 = -2.2763 - 1.6850 * L + 0.76661 * L ** 2 - 0.053346 * L ** 3
        f1 = 15.2827 - 9.2846 * L + 1.99381 * L ** 2 - 0.142631 * L ** 3
        f2 = -197.789 + 190.266 * L - 67.9775 * L ** 2 + 10.6913 * L ** 3 - 0.625151 * L ** 4
        lt = np.log10(theta)
        k_ff = 1e-26 * Pe * 10.0 ** (f0 + f1 * lt + f2 * lt ** 2)
        return self.n_HI[:, None] * (k_bf + k_ff) + 6.65e-25 * self.n_e[:, None]


def continuum_grid(atm, wl_lo_cm, wl_hi_cm, line_buffer=10e-8, cntm_step=1e-8):
    """Knots and values of the continuum on the coarse grid ``synthesize`` builds
    (``src/synthesize.jl:184-188``): (lo - buffer - step) : step : (hi + buffer + step), in cm."""
    lo = wl_lo_cm - line_buffer - cntm_step
    hi = wl_hi_cm + line_buffer + cntm_step
    n = int(math.floor((hi - lo) / cntm_step + 1e-9)) + 1
    knots = lo + cntm_step * np.arange(n)
    return knots, np.ascontiguousarray(atm.continuum(knots))
