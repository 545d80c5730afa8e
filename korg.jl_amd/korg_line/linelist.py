"""Spectral lines and Korg's default broadening recipes.

``Line`` restates the Korg ``Line`` constructor (``src/linelist.jl:56-100``): unit conversion of the
wavelength, imputation of missing broadening parameters by ``approximate_radiative_gamma``
(``src/linelist.jl:124-129``) and ``approximate_gammas`` (``src/linelist.jl:149-183``), and the
decoding of the van der Waals parameter (log gamma / Unsoeld fudge factor / packed ABO).

``LineList`` is the struct-of-arrays form handed to the C ABI; ``LineList.from_arrays`` applies the
same recipes element-wise with numpy (same IEEE operation order as the scalar code), which is how
the large synthetic lists are built.
"""
import math
import os
import warnings

import numpy as np

from .constants import (bohr_radius_cgs, c_cgs, electron_charge_cgs, electron_mass_cgs,
                        hplanck_eV, kboltz_cgs, RydbergH_eV, Rydberg_eV)
from .species import Species

_DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")


def _load_ionization_energies():
    table = {}
    with open(os.path.join(_DATA, "ionization_energies.tsv")) as f:
        for row in f:
            if row.startswith("#") or not row.strip():
                continue
            t = row.split()
            table[int(t[0])] = (float(t[2]), float(t[3]), float(t[4]))
    return table


ionization_energies = _load_ionization_energies()

_PI = math.pi
_8PI2 = 8 * (_PI * _PI)  # Julia's 8π^2


def approximate_radiative_gamma(wl, log_gf):
    """Classical radiative damping, ``src/linelist.jl:124-129`` (wl in cm)."""
    e, m, c = electron_charge_cgs, electron_mass_cgs, c_cgs
    return _8PI2 * (e * e) / (m * c * (wl * wl)) * 10.0 ** log_gf


def approximate_gammas(wl, species, E_lower, ionization_energies=ionization_energies):
    """Cowley (1971) Stark and Unsoeld vdW at 10^4 K, ``src/linelist.jl:149-183``.

    Returns ``(gamma_stark, log10(gamma_vdW))``; ``log10(gamma_vdW)`` is 0.0 for autoionizing lines
    (E_upper > chi), and both are 0 for molecules and for ionization stages above III.
    """
    Z = species.charge + 1
    if species.ismolecule() or Z > 3:
        return 0.0, 0.0
    chi = ionization_energies[species.get_atoms()[0]][Z - 1]
    c, h, k = c_cgs, hplanck_eV, kboltz_cgs
    E_upper = E_lower + (h * c / wl)
    nstar4_upper = (float(Z * Z) * RydbergH_eV / (chi - E_upper)) ** 2
    if Z == 1:
        gamma_stark = 2.25910152e-7 * nstar4_upper
    else:
        gamma_stark = 5.42184365e-7 * nstar4_upper / float((Z + 1) * (Z + 1))
    drbar2 = (5 / 2) * (Rydberg_eV * Rydberg_eV) * float(Z * Z) * \
        (1 / (chi - E_upper) ** 2 - 1 / (chi - E_lower) ** 2)
    if chi < E_upper:
        log_gamma_vdW = 0.0
    else:
        log_gamma_vdW = 6.33 + 0.4 * math.log10(drbar2) + 0.3 * math.log10(10_000) + math.log10(k)
    return gamma_stark, log_gamma_vdW


def _is_missing(x):
    return x is None or (isinstance(x, float) and math.isnan(x))


class Line:
    """One spectral line (``src/linelist.jl:5-100``).

    ``wl`` in cm (converted from Angstrom when >= 1), ``log_gf``, ``species`` (a ``Species`` or a
    code string), ``E_lower`` in eV, ``gamma_rad``/``gamma_stark`` in s^-1 (FWHM, per perturber
    for Stark) and ``vdW`` as log10(gamma_vdW) (< 0), 0 (none), an Unsoeld fudge factor (0-20),
    packed ABO (>= 20) or an ``(sigma, alpha)`` tuple.  Missing values (None/NaN) are imputed.
    """
    __slots__ = ("wl", "log_gf", "species", "E_lower", "gamma_rad", "gamma_stark", "vdW")

    def __init__(self, wl, log_gf, species, E_lower, gamma_rad=None, gamma_stark=None, vdW=None):
        if not isinstance(species, Species):
            species = Species(species)
        wl = float(wl)
        if wl >= 1:
            wl *= 1e-8
        if _is_missing(gamma_stark) or _is_missing(vdW):
            stark_approx, vdW_approx = approximate_gammas(wl, species, E_lower)
            if _is_missing(gamma_stark):
                gamma_stark = stark_approx
            if _is_missing(vdW):
                vdW = vdW_approx
        if _is_missing(gamma_rad):
            gamma_rad = approximate_radiative_gamma(wl, log_gf)
        if not isinstance(vdW, tuple):
            vdW = float(vdW)
            if vdW < 0:
                vdW = (10.0 ** vdW, -1.0)
            elif vdW == 0:
                vdW = (0.0, -1.0)
            elif 0 < vdW < 20:
                vdW = (vdW * 10.0 ** approximate_gammas(wl, species, E_lower)[1], -1.0)
            else:
                fl = math.floor(vdW)
                vdW = (fl * bohr_radius_cgs * bohr_radius_cgs, vdW - fl)
        self.wl = wl
        self.log_gf = float(log_gf)
        self.species = species
        self.E_lower = float(E_lower)
        self.gamma_rad = float(gamma_rad)
        self.gamma_stark = float(gamma_stark)
        self.vdW = (float(vdW[0]), float(vdW[1]))

    def __repr__(self):
        return (f"{self.species!r} {self.wl * 1e8:.6f} Å (log gf = {self.log_gf:.2f}, "
                f"χ = {self.E_lower:.2f} eV)")


class LineList:
    """Struct-of-arrays line list: the layout the C ABI consumes.

    Arrays (length ``n``): ``wl`` [cm], ``log_gf``, ``E_lower`` [eV], ``gamma_rad``,
    ``gamma_stark``, ``vdW1``, ``vdW2`` (``vdW2 == -1`` marks a plain gamma_vdW at 10^4 K in
    ``vdW1``; otherwise ``(vdW1, vdW2)`` are ABO (sigma, alpha)), ``species_index`` (int32 into
    ``species``).
    """

    def __init__(self, wl, log_gf, E_lower, gamma_rad, gamma_stark, vdW1, vdW2, species_index,
                 species):
        self.wl = np.ascontiguousarray(wl, dtype=np.float64)
        self.log_gf = np.ascontiguousarray(log_gf, dtype=np.float64)
        self.E_lower = np.ascontiguousarray(E_lower, dtype=np.float64)
        self.gamma_rad = np.ascontiguousarray(gamma_rad, dtype=np.float64)
        self.gamma_stark = np.ascontiguousarray(gamma_stark, dtype=np.float64)
        self.vdW1 = np.ascontiguousarray(vdW1, dtype=np.float64)
        self.vdW2 = np.ascontiguousarray(vdW2, dtype=np.float64)
        self.species_index = np.ascontiguousarray(species_index, dtype=np.int32)
        self.species = list(species)

    def __len__(self):
        return self.wl.size

    @classmethod
    def from_lines(cls, lines):
        species, index = [], {}
        idx = []
        for l in lines:
            if l.species not in index:
                index[l.species] = len(species)
                species.append(l.species)
            idx.append(index[l.species])
        return cls([l.wl for l in lines], [l.log_gf for l in lines], [l.E_lower for l in lines],
                   [l.gamma_rad for l in lines], [l.gamma_stark for l in lines],
                   [l.vdW[0] for l in lines], [l.vdW[1] for l in lines], idx, species)

    def to_lines(self):
        out = []
        for i in range(len(self)):
            l = Line.__new__(Line)
            l.wl, l.log_gf = float(self.wl[i]), float(self.log_gf[i])
            l.species = self.species[self.species_index[i]]
            l.E_lower = float(self.E_lower[i])
            l.gamma_rad, l.gamma_stark = float(self.gamma_rad[i]), float(self.gamma_stark[i])
            l.vdW = (float(self.vdW1[i]), float(self.vdW2[i]))
            out.append(l)
        return out

    def take(self, idx):
        idx = np.asarray(idx)
        return LineList(self.wl[idx], self.log_gf[idx], self.E_lower[idx], self.gamma_rad[idx],
                        self.gamma_stark[idx], self.vdW1[idx], self.vdW2[idx],
                        self.species_index[idx], self.species)

    def is_sorted(self):
        return bool(np.all(self.wl[1:] >= self.wl[:-1]))

    def sorted(self):
        """Stable sort by wavelength, as ``synthesize`` does (``src/synthesize.jl:191-195``)."""
        return self.take(np.argsort(self.wl, kind="stable"))

    @classmethod
    def from_arrays(cls, wl, log_gf, species, species_index, E_lower, gamma_rad=None,
                    gamma_stark=None, vdW=None):
        """Vectorised ``Line`` construction; ``None``/NaN entries are imputed like ``Line``."""
        wl = np.asarray(wl, dtype=np.float64).copy()
        n = wl.size
        wl = np.where(wl >= 1, wl * 1e-8, wl)
        log_gf = np.asarray(log_gf, dtype=np.float64)
        E_lower = np.asarray(E_lower, dtype=np.float64)
        species_index = np.asarray(species_index, dtype=np.int32)

        def _arr(x):
            return np.full(n, np.nan) if x is None else np.asarray(x, dtype=np.float64).copy()

        gamma_rad, gamma_stark, vdW = _arr(gamma_rad), _arr(gamma_stark), _arr(vdW)
        need = np.isnan(gamma_stark) | np.isnan(vdW)
        stark_a, vdw_a = _approximate_gammas_vec(wl, species, species_index, E_lower)
        gamma_stark = np.where(need & np.isnan(gamma_stark), stark_a, gamma_stark)
        vdW = np.where(need & np.isnan(vdW), vdw_a, vdW)
        gamma_rad = np.where(np.isnan(gamma_rad), _approximate_radiative_gamma_vec(wl, log_gf),
                             gamma_rad)
        vdW1 = np.empty(n)
        vdW2 = np.full(n, -1.0)
        neg, zero = vdW < 0, vdW == 0
        fudge = (vdW > 0) & (vdW < 20)
        abo = vdW >= 20
        vdW1[neg] = _pow10(vdW[neg])
        vdW1[zero] = 0.0
        vdW1[fudge] = vdW[fudge] * _pow10(vdw_a[fudge])
        fl = np.floor(vdW[abo])
        vdW1[abo] = fl * bohr_radius_cgs * bohr_radius_cgs
        vdW2[abo] = vdW[abo] - fl
        return cls(wl, log_gf, E_lower, gamma_rad, gamma_stark, vdW1, vdW2, species_index, species)


def _pow10(x):
    """Element-wise 10.0 ** x with the same libm pow as the scalar ``Line`` path."""
    return np.array([10.0 ** v for v in np.asarray(x, dtype=np.float64).tolist()])


def _log10(x):
    return np.array([math.log10(v) if v > 0 else (-math.inf if v == 0 else math.nan)
                     for v in np.asarray(x, dtype=np.float64).tolist()])


def _approximate_radiative_gamma_vec(wl, log_gf):
    e, m, c = electron_charge_cgs, electron_mass_cgs, c_cgs
    return _8PI2 * (e * e) / (m * c * (wl * wl)) * _pow10(log_gf)


def _approximate_gammas_vec(wl, species, species_index, E_lower):
    """Element-wise ``approximate_gammas`` with the scalar version's operation order."""
    n = wl.size
    stark = np.zeros(n)
    logvdw = np.zeros(n)
    c, h, k = c_cgs, hplanck_eV, kboltz_cgs
    log10_k = math.log10(k)
    c03 = 0.3 * math.log10(10_000)
    for si, sp in enumerate(species):
        sel = species_index == si
        if not np.any(sel):
            continue
        Z = sp.charge + 1
        if sp.ismolecule() or Z > 3:
            continue
        chi = ionization_energies[sp.get_atoms()[0]][Z - 1]
        El = E_lower[sel]
        Eu = El + (h * c / wl[sel])
        nstar4 = (float(Z * Z) * RydbergH_eV / (chi - Eu)) ** 2
        if Z == 1:
            stark[sel] = 2.25910152e-7 * nstar4
        else:
            stark[sel] = 5.42184365e-7 * nstar4 / float((Z + 1) * (Z + 1))
        with np.errstate(divide="ignore", invalid="ignore"):
            drbar2 = (5 / 2) * (Rydberg_eV * Rydberg_eV) * float(Z * Z) * \
                (1 / (chi - Eu) ** 2 - 1 / (chi - El) ** 2)
            lv = 6.33 + 0.4 * _log10(drbar2) + c03 + log10_k
        logvdw[sel] = np.where(chi < Eu, 0.0, lv)
    return stark, logvdw


# ---------------------------------------------------------------------------------------------
# Linelist ingest: Korg's HDF5 formats and filter_linelist
# ---------------------------------------------------------------------------------------------
def _species_table(formula_rows, charge):
    """Distinct (atoms, charge) pairs of per-line formula rows -> (species list, int32 index)."""
    keys = np.concatenate([formula_rows.astype(np.int64),
                           np.asarray(charge, dtype=np.int64)[:, None]], axis=1)
    uniq, inv = np.unique(keys, axis=0, return_inverse=True)
    species = []
    for row in uniq:
        atoms = tuple(int(z) for z in row[:-1] if z != 0)
        species.append(Species(atoms, charge=int(row[-1])))
    return species, inv.reshape(-1).astype(np.int32)


def _tentothe_or_missing(x):
    """``tentotheOrMissing`` (``src/linelist.jl:438``): 0 means missing, else 10^x."""
    x = np.asarray(x, dtype=np.float64)
    out = np.full(x.shape, np.nan)
    nz = x != 0
    out[nz] = _pow10(x[nz])
    return out


def _id_or_missing(x):
    """``idOrMissing`` (``src/linelist.jl:439``): 0 means missing."""
    x = np.asarray(x, dtype=np.float64)
    return np.where(x == 0, np.nan, x)


def galah_linelist_from_columns(formula, ionization, wl, log_gf, E_lo, gamma_rad, gamma_stark,
                                vdW):
    """Lines from the GALAH DR3 file's columns, as Korg builds them (``get_GALAH_DR3_linelist``,
    ``src/linelist.jl:894-922``): formula rows of up to 3 atomic numbers (a diatomic or triatomic
    molecule's sorted), charge = ionization - 1, wavelengths in (vacuum) Angstrom, gamma_rad and
    gamma_stark as log10 (0 = missing: Korg's defaults), vdW as stored (0 = missing), and the H I
    lines removed.  Returns a ``LineList`` in the columns' order.
    """
    formula = np.asarray(formula)
    ion = np.asarray(ionization).astype(np.int64)
    # single atom when the second entry is 0, else the sorted first two or all three
    rows = np.zeros_like(formula)
    single = formula[:, 1] == 0
    two = ~single & (formula[:, 2] == 0)
    three = ~single & ~two
    rows[single, 0] = formula[single, 0]
    rows[two, :2] = np.sort(formula[two, :2], axis=1)
    rows[three] = np.sort(formula[three], axis=1)
    keep = ~(single & (formula[:, 0] == 1) & (ion == 1))  # H I
    species, sidx = _species_table(rows[keep], ion[keep] - 1)
    col = lambda x: np.asarray(x).astype(np.float64)[keep]
    return LineList.from_arrays(col(wl), col(log_gf), species, sidx, col(E_lo),
                                gamma_rad=_tentothe_or_missing(col(gamma_rad)),
                                gamma_stark=_tentothe_or_missing(col(gamma_stark)),
                                vdW=_id_or_missing(col(vdW)))


GALAH_COLUMNS = ("formula", "ionization", "wl", "log_gf", "E_lo", "gamma_rad", "gamma_stark", "vdW")


def read_galah_dr3_linelist(path):
    """The GALAH DR3 linelist file (``galah_dr3_linelist.h5``), read with the HDF5 reader and
    converted by ``galah_linelist_from_columns``."""
    from .hdf5 import H5File
    f = H5File(path)
    return galah_linelist_from_columns(*(f.read(k) for k in GALAH_COLUMNS))


def read_korg_linelist(path):
    """A linelist written by Korg's ``save_linelist`` (``src/linelist.jl:785-818``), read as
    ``read_korg_linelist`` does (``src/linelist.jl:820-839``): wavelengths in cm, the formula as
    Korg's 6-slot atomic-number array (zeros first), charge, E_lower, gamma_rad, gamma_stark and
    the (vdW_1, vdW_2) pair, all stored as given (no defaults are imputed).
    """
    from .hdf5 import H5File
    f = H5File(path)
    formula = f.read("formula")  # (n, 6)
    if np.any(formula.max(axis=1) == 0):
        raise ValueError(f"{path}: a line has an empty formula")
    species, sidx = _species_table(formula, f.read("charge"))
    return LineList(f.read("wl"), f.read("log_gf"), f.read("E_lower"), f.read("gamma_rad"),
                    f.read("gamma_stark"), f.read("vdW_1"), f.read("vdW_2"), sidx, species)


def filter_linelist(linelist, wls, line_buffer):
    """``filter_linelist`` (``src/synthesize.jl:331-350``): the lines (of a wavelength-sorted
    list) within ``line_buffer`` cm of each of the grid's ranges, each line at most once, in order.
    """
    wl = linelist.wl
    idx, last = [], 0
    for lo, hi in wls.eachwindow():
        first = int(np.searchsorted(wl, lo - line_buffer, side="left")) + 1  # 1-based
        first = max(first, last + 1)
        last = int(np.searchsorted(wl, hi + line_buffer, side="right"))
        if last >= first:
            idx.append(np.arange(first - 1, last))
    sel = np.concatenate(idx) if idx else np.zeros(0, dtype=np.int64)
    return linelist.take(sel)


def air_to_vacuum(lam):
    """Air to vacuum wavelength by Birch & Downs (1994), the formula VALD publishes
    (``src/utils.jl:240-251``): ``lam`` in Angstrom if >= 1, else in cm."""
    cgs = lam < 1
    if cgs:
        lam = lam * 1e8
    s = 1e4 / lam
    n = 1 + 0.00008336624212083 + 0.02408926869968 / (130.1065924522 - s ** 2) + \
        0.0001599740894897 / (38.92568793293 - s ** 2)
    if cgs:
        lam = lam * 1e-8
    return lam * n


_VALD_COLUMNS = {
    (True, True): ["species", "wl", "E_low", "loggf", "gamma_rad", "gamma_stark", "gamma_vdW",
                   "lande", "reference"],
    (True, False): ["species", "wl", "E_low", "Vmic", "loggf", "gamma_rad", "gamma_stark",
                    "gamma_vdW", "lande", "depth", "reference"],
    (False, True): ["species", "wl", "loggf", "E_low", "J_lo", "E_up", "J_up", "lower_lande",
                    "upper_lande", "mean_lande", "gamma_rad", "gamma_stark", "gamma_vdW"],
}
_VALD_COLUMNS[(False, False)] = _VALD_COLUMNS[(False, True)]


def read_vald_linelist(path):
    """A VALD linelist (short or long format, "extract all" or "extract stellar"), read as
    ``parse_vald_linelist`` does (``src/linelist.jl:540-634``): energies in eV (converted from
    cm^-1 when the header says so), wavelengths in vacuum (converted from air when the header
    says so), gamma_rad 0 imputed by Korg's default, gamma_stark as log10 and vdW as given (0 =
    missing).  Files whose oscillator strengths VALD did not scale by isotopic abundance are read
    only if no isotope is named in their references (the isotopic-abundance table is not part of
    this library).  As ``read_linelist``, lines of charge outside 0..2 and H I lines are dropped
    and the list is sorted by wavelength.
    """
    import csv
    import re
    with open(path) as f:
        lines = [l.rstrip("\n") for l in f if len(l.rstrip("\n")) > 0 and not l.startswith("#")]
    if lines[0].startswith(" WARNING: Output was truncated to 100000 lines"):
        lines = lines[1:]
    lines = [l.replace("'", '"') for l in lines]
    extractall = not re.match(r"^\s+\d", lines[0])
    first = 2 if extractall else 3  # 0-based index of the first line of the body
    header = lines[first - 1]
    unscaled = any(l.startswith("* oscillator strengths were NOT scaled ") for l in lines)
    if not unscaled and not any(l.startswith("* oscillator strengths were scaled ") for l in lines):
        raise ValueError("can't detect whether the log(gf)s are scaled by isotopic abundance")
    short = not re.match(r'^"? ', lines[first + 1])
    body = lines[first::(1 if short else 4)]
    end = next((k for k, l in enumerate(body) if l[0] != '"' or not l[1].isupper()), len(body))
    body = body[:end]
    cols = _VALD_COLUMNS[(short, extractall)]
    rows = [dict(zip(cols, r)) for r in csv.reader(body, quotechar='"', skipinitialspace=True)]

    def num(r, k):
        return float(r[k])
    E_low = np.array([num(r, "E_low") for r in rows])
    if "cm" in header:
        E_low = E_low * c_cgs * hplanck_eV
    elif "eV" not in header:
        raise ValueError(f"can't determine the energy units from the header: {header!r}")
    wl = np.array([num(r, "wl") for r in rows])
    if "air" in header:
        wl = np.array([air_to_vacuum(x) for x in wl])
    elif "vac" not in header:
        raise ValueError(f"can't determine vacuum or air wavelengths from the header: {header!r}")
    wl = 1e-8 * wl
    if unscaled:
        refs = [lines[first + 3 + 4 * k] for k in range(len(rows))] if not short else \
            [r.get("reference", "") for r in rows]
        if any(re.search(r"\((\d\d?\d?)\)([A-Z][a-z]?)", ref) for ref in refs):
            raise NotImplementedError("isotopic scaling of unscaled VALD log(gf)s is not supported")
    loggf = np.array([num(r, "loggf") for r in rows])
    g_rad = np.array([num(r, "gamma_rad") for r in rows])
    gamma_rad = np.where(g_rad == 0, np.nan, _pow10(np.where(g_rad == 0, 0.0, g_rad)))
    names = [r["species"].strip() for r in rows]
    species, index = [], {}
    sidx = np.empty(len(rows), dtype=np.int32)
    for k, nm in enumerate(names):
        sp = Species(nm)
        if sp not in index:
            index[sp] = len(species)
            species.append(sp)
        sidx[k] = index[sp]
    ll = LineList.from_arrays(wl, loggf, species, sidx, E_low, gamma_rad=gamma_rad,
                              gamma_stark=_tentothe_or_missing([num(r, "gamma_stark") for r in rows]),
                              vdW=_id_or_missing([num(r, "gamma_vdW") for r in rows]))
    return _finish_text_linelist(ll)


def _finish_text_linelist(ll):
    """``read_linelist``'s last steps for every text format (``src/linelist.jl:416-424``): lines of
    charge outside 0..2 and H I lines are dropped, and the list is sorted by wavelength (stable)."""
    ok = np.array([0 <= sp.charge <= 2 and sp != Species("H I") for sp in ll.species], dtype=bool)
    ll = ll.take(np.flatnonzero(ok[ll.species_index])) if ll.species else ll
    return ll if ll.is_sorted() else ll.sorted()


def _or_zero(s):
    """``parse_or_zero`` (``src/linelist.jl:442-449``): a blank field reads as 0."""
    s = s.strip()
    return float(s) if s else 0.0


_KURUCZ_COLUMN = object()  # the default of read_kurucz_linelist's isotopic_abundances


def read_kurucz_linelist(path, vacuum=False, isotopic_abundances=_KURUCZ_COLUMN):
    """An atomic Kurucz linelist (160-column rows) as ``read_linelist(...; format="kurucz")`` reads
    it (``parse_kurucz_linelist``, ``src/linelist.jl:451-511``, then the common filter and sort):
    wavelengths in nm (air, converted to vacuum unless ``vacuum``), E_lower the smaller of the two
    levels' |wavenumber|, log gf plus the hyperfine column, gamma_rad and gamma_stark as log10 and
    vdW as given (0 = missing).

    Isotopes, as Korg's ``isotopic_abundances`` argument: ``None`` adds Kurucz's own isotope
    column to log gf (Korg's ``isotopic_abundances = nothing``); a table ``{Z: {A: fraction}}``
    adds log10 of the fraction for rows that name an isotope (columns 107-109, else 116-118) found
    in it and nothing for the others (Korg's rule with its table, whose NIST values this library
    does not ship).  Left at its default the reader adds Kurucz's column and warns when a row names
    an isotope, since Korg's default read differs there.
    Molecular Kurucz files (first row of 100 columns or fewer) are not supported.
    """
    with open(path) as f:
        rows = [r.rstrip("\n") for r in f]
    first = next((r for r in rows if r.strip()), "")
    if len(first) <= 100:
        raise ValueError(f"{path}: molecular Kurucz linelists are not supported")
    wl, loggf, E, grad, gstark, vdw, names = [], [], [], [], [], [], []
    n_tagged = 0
    for row in rows:
        if row == "":
            continue
        if len(row) == 159:  # a missing column in the wavelength field
            row = " " + row
        row = row.ljust(160)
        e1 = abs(float(row[24:36])) * c_cgs * hplanck_eV
        e2 = abs(float(row[52:64])) * c_cgs * hplanck_eV
        E.append(min(e1, e2))
        names.append(row[18:24])
        gf = float(row[11:18]) + _or_zero(row[109:115])
        iso = int(_or_zero(row[106:109])) or int(_or_zero(row[115:118]))
        if isotopic_abundances is None or isotopic_abundances is _KURUCZ_COLUMN:
            gf += _or_zero(row[118:124])
            n_tagged += iso != 0
        elif iso != 0:
            frac = isotopic_abundances.get(_atomic_number(row[18:24]), {}).get(iso)
            if frac is not None:
                gf += math.log10(frac)
        loggf.append(gf)
        lam = float(row[0:11]) * 1e-7
        wl.append(lam if vacuum else air_to_vacuum(lam))
        grad.append(_or_zero(row[80:86]))
        gstark.append(_or_zero(row[86:92]))
        vdw.append(_or_zero(row[92:98]))
    species, index = [], {}
    sidx = np.empty(len(names), dtype=np.int32)
    for k, nm in enumerate(names):
        sp = Species(nm)
        if sp not in index:
            index[sp] = len(species)
            species.append(sp)
        sidx[k] = index[sp]
    ll = LineList.from_arrays(np.array(wl), np.array(loggf), species, sidx, np.array(E),
                              gamma_rad=_tentothe_or_missing(grad),
                              gamma_stark=_tentothe_or_missing(gstark), vdW=_id_or_missing(vdw))
    if n_tagged and isotopic_abundances is _KURUCZ_COLUMN:
        warnings.warn(f"{path}: {n_tagged} rows name an isotope; their log gf includes Kurucz's "
                      "isotope column (Korg's isotopic_abundances = nothing), where Korg's default "
                      "read uses its isotopic-abundance table.  Pass isotopic_abundances=None to "
                      "accept this, or a table {Z: {A: fraction}}.", stacklevel=2)
    return _finish_text_linelist(ll)


def _atomic_number(code):
    """The atomic number of an atomic species code such as ``"26.01"``."""
    atoms = Species(code).atoms
    if len(atoms) != 1:
        raise ValueError(f"{code!r} is not an atomic species")
    return int(atoms[0])
