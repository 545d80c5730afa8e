"""Linelist ingest: the HDF5 reader (korg_line/hdf5.py), Korg's GALAH DR3 linelist file
(src/linelist.jl:894-922) and filter_linelist (src/synthesize.jl:331-350).  CPU only.

The reader is pinned by files written by libhdf5 itself: the Lyman-profile test file of the
reference (tests/golden/lyman_absorption.h5, copied from test/data/), whose attributes fix the
grid its dataset must have, and -- when the reference tree is present -- the GALAH DR3 linelist
(307,002 lines in 64 deflate-compressed chunks per column), which Korg's own test requires to be
sorted by wavelength (test/linelist.jl:28).  filter_linelist is checked as the reference checks it
(test/synthesize.jl:89-110): against a naive filter.
"""
import os

import numpy as np
import pytest

from korg_line.hdf5 import H5File, H5Error
from korg_line.linelist import (Line, LineList, filter_linelist, read_galah_dr3_linelist,
                                read_korg_linelist)
from korg_line.species import Species
from korg_line.wavelengths import Wavelengths

HERE = os.path.dirname(os.path.abspath(__file__))
LYMAN = os.path.join(HERE, "golden", "lyman_absorption.h5")
GALAH = "/root/reference/data/linelists/GALAH_DR3/galah_dr3_linelist.h5"
needs_galah = pytest.mark.skipif(not os.path.exists(GALAH), reason="reference data tree absent")


def test_hdf5_lyman_profile_and_attributes():
    f = H5File(LYMAN)
    assert f.keys() == ["profile"]
    a = f.attrs("profile")
    assert a["T"] == 9000.0 and a["ne"] == 1e11 and a["nH_I"] == 1e13
    prof = f.read("profile")
    # the dataset has exactly the grid the attributes describe
    n = int(round((a["stop_wl"] - a["start_wl"]) / a["wl_step"])) + 1
    assert prof.shape == (n,) == (4001,) and prof.dtype == np.float64
    assert np.all(np.isfinite(prof)) and np.all(prof > 0)
    assert f.shape("profile") == (4001,)


def test_hdf5_shuffle_inverse():
    """The shuffle filter stores byte k of every element together; undoing it restores them."""
    f = H5File(LYMAN)
    x = np.arange(1000, dtype="<f8") * np.pi
    raw = x.tobytes()
    shuffled = np.frombuffer(raw, dtype=np.uint8).reshape(-1, 8).T.tobytes()
    back = f._unfilter(shuffled, [(2, [8])], 0, 8)
    assert back == raw
    # a filter skipped by the chunk's mask is not applied
    assert f._unfilter(shuffled, [(2, [8])], 1, 8) == shuffled


def test_hdf5_rejects_non_hdf5(tmp_path):
    p = tmp_path / "x.h5"
    p.write_bytes(b"not an hdf5 file" * 100)
    with pytest.raises(H5Error):
        H5File(str(p))


@needs_galah
def test_hdf5_galah_columns():
    f = H5File(GALAH)
    assert f.keys() == ["E_lo", "formula", "gamma_rad", "gamma_stark", "ionization", "log_gf",
                        "vdW", "wl"]
    n = 307002
    for k in ("E_lo", "gamma_rad", "gamma_stark", "ionization", "log_gf", "vdW", "wl"):
        assert f.shape(k) == (n,), k
    wl = f.read("wl")
    assert wl.dtype == np.float32 and np.all(np.diff(wl) >= 0)  # every chunk in place
    formula = f.read("formula")
    assert formula.shape == (n, 3) and formula.dtype == np.uint8
    assert np.all(formula[:, 0] > 0) and formula.max() <= 92
    assert set(np.unique(f.read("ionization"))) == {1, 2, 3}
    E = f.read("E_lo")
    assert np.all(np.isfinite(E)) and E.min() >= 0 and E.max() < 20


@needs_galah
def test_galah_linelist_as_korg_builds_it():
    raw = H5File(GALAH)
    ll = read_galah_dr3_linelist(GALAH)
    # H I lines are removed (two in the file); the list is sorted (test/linelist.jl:28)
    assert len(ll) == 307000 and ll.is_sorted()
    assert Species("H I") not in ll.species
    # each kept line against the scalar Line constructor on the file's values, with Korg's
    # tentotheOrMissing / idOrMissing for the broadening columns (src/linelist.jl:438-439)
    formula, ion = raw.read("formula"), raw.read("ionization")
    keep = np.flatnonzero(~((formula[:, 0] == 1) & (formula[:, 1] == 0) & (ion == 1)))
    cols = {k: raw.read(k).astype(np.float64) for k in
            ("wl", "log_gf", "E_lo", "gamma_rad", "gamma_stark", "vdW")}
    rng = np.random.default_rng(5)
    for i in rng.choice(len(ll), 300, replace=False):
        r = keep[i]
        atoms = [int(z) for z in formula[r] if z != 0]
        sp = Species(tuple(sorted(atoms)), charge=int(ion[r]) - 1)
        g = lambda x: None if x == 0 else 10.0 ** x
        ref = Line(cols["wl"][r], cols["log_gf"][r], sp, cols["E_lo"][r],
                   gamma_rad=g(cols["gamma_rad"][r]), gamma_stark=g(cols["gamma_stark"][r]),
                   vdW=None if cols["vdW"][r] == 0 else cols["vdW"][r])
        assert ll.species[ll.species_index[i]] == sp
        assert ll.wl[i] == ref.wl and ll.log_gf[i] == ref.log_gf and ll.E_lower[i] == ref.E_lower
        assert ll.gamma_rad[i] == ref.gamma_rad and ll.gamma_stark[i] == ref.gamma_stark
        assert (ll.vdW1[i], ll.vdW2[i]) == ref.vdW, (r, (ll.vdW1[i], ll.vdW2[i]), ref.vdW)
    # molecules: TiO appears as (O, Ti), neutral
    assert Species((8, 22), charge=0) in ll.species


@needs_galah
def test_galah_fixture_is_the_files_data():
    """tests/golden/galah_4800_fe_ti_ni_ca.npz (used on the device) holds the file's rows."""
    from korg_line.linelist import GALAH_COLUMNS, galah_linelist_from_columns
    z = np.load(os.path.join(HERE, "golden", "galah_4800_fe_ti_ni_ca.npz"))
    f = H5File(GALAH)
    for k in GALAH_COLUMNS:
        assert np.array_equal(z[k], f.read(k)[z["file_index"]]), k
    sub = galah_linelist_from_columns(*(z[k] for k in GALAH_COLUMNS))
    full = read_galah_dr3_linelist(GALAH)
    fo, ion = f.read("formula"), f.read("ionization")
    kept = ~((fo[:, 0] == 1) & (fo[:, 1] == 0) & (ion == 1))
    i = np.cumsum(kept)[z["file_index"]] - 1  # the rows' places in the full list (H I dropped)
    for name in ("wl", "log_gf", "E_lower", "gamma_rad", "gamma_stark", "vdW1", "vdW2"):
        assert np.array_equal(getattr(sub, name), getattr(full, name)[i]), name
    assert np.count_nonzero(sub.vdW2 != -1.0) >= 20  # packed ABO parameters among them


def _naive_filter(ll, wls, buf):
    """test/synthesize.jl:90-96: lines within the buffer of any of the grid's ranges."""
    m = np.zeros(len(ll), dtype=bool)
    s = 0
    for _, _, n in wls.wl_ranges:
        lo, hi = wls.all_wls[s], wls.all_wls[s + n - 1]
        s += n
        m |= (ll.wl >= lo - buf) & (ll.wl <= hi + buf)
    return np.flatnonzero(m)


@needs_galah
@pytest.mark.parametrize("ranges", [
    (4750, 4850),
    (6500, 6600),
    [(4750, 4850), (5700, 5750), (7600, 7700)],
    [(5750, 5751), (5752, 5753)],  # buffers overlap: no line twice
])
def test_filter_linelist_against_naive(ranges):
    ll = read_galah_dr3_linelist(GALAH)
    wls = Wavelengths(ranges)
    b = 2.0 * 1e-8
    got = filter_linelist(ll, wls, b)
    ref = _naive_filter(ll, wls, b)
    assert len(got) == ref.size
    assert np.array_equal(got.wl, ll.wl[ref])


def test_filter_linelist_small():
    """Hand-made list: windows, buffer edges (inclusive) and an empty result."""
    wl = np.array([4990.0, 4999.0, 5000.5, 5003.0, 5010.0, 5012.0, 6000.0]) * 1e-8
    ll = LineList.from_arrays(wl, np.zeros(7), [Species("Fe I")], np.zeros(7, dtype=np.int32),
                              np.ones(7))
    wls = Wavelengths([(5000, 5005), (5008, 5010)])
    got = filter_linelist(ll, wls, 1e-8)
    assert np.allclose(got.wl * 1e8, [4999.0, 5000.5, 5003.0, 5010.0])
    assert len(filter_linelist(ll, Wavelengths((7000, 7001)), 1e-8)) == 0


def test_read_korg_linelist_requires_its_columns():
    """The Korg-format reader needs Korg's column names; the Lyman file has none of them."""
    with pytest.raises(KeyError):
        read_korg_linelist(LYMAN)


def test_hdf5_truncated_file_is_reported(tmp_path):
    """A file cut short fails with H5Error, not with garbage values."""
    raw = open(LYMAN, "rb").read()
    p = tmp_path / "cut.h5"
    p.write_bytes(raw[:len(raw) - 20000])  # the profile's data (32 KB) runs past the new end
    with pytest.raises(H5Error):
        H5File(str(p)).read("profile")


# -------------------------------------------------------------------------------------- VALD
VALD = os.path.join(HERE, "golden", "vald")  # the reference's test files (test/data/linelists)


def test_vald_short_format_abo_and_missing_params():
    """test/linelist.jl:192-223, value by value."""
    from korg_line.linelist import read_vald_linelist
    ll = read_vald_linelist(os.path.join(VALD, "linelist.vald"))
    L = ll.to_lines()
    assert len(L) == 6
    assert L[0].wl == pytest.approx(3000.0414 * 1e-8, rel=1e-14)
    assert L[0].log_gf == -2.957
    assert L[0].species == Species("Fe_I")
    assert L[0].E_lower == pytest.approx(3.3014, rel=1e-14)
    assert L[0].gamma_rad == pytest.approx(1.905460717963248e7, rel=1e-14)
    assert L[0].gamma_stark == pytest.approx(0.0001230268770812381, rel=1e-14)
    assert L[0].vdW[0] == pytest.approx(4.6773514128719815e-8, rel=1e-14)
    # imputation of missing broadening parameters
    assert L[1].gamma_rad == pytest.approx(818252.5391161365, rel=1e-14)
    assert L[1].gamma_stark == L[0].gamma_stark and L[1].vdW == L[0].vdW
    assert L[2].gamma_rad == L[1].gamma_rad
    assert L[2].gamma_stark == pytest.approx(0.00019044182974029873, rel=1e-14)
    assert L[2].vdW == L[0].vdW
    assert L[3].gamma_rad == L[0].gamma_rad and L[3].gamma_stark == L[2].gamma_stark
    assert L[3].vdW[0] == pytest.approx(9.953360714197118e-8, rel=1e-14)
    assert L[4].gamma_rad == L[0].gamma_rad and L[4].gamma_stark == L[0].gamma_stark
    assert L[4].vdW == L[3].vdW
    # ABO parameters
    assert L[5].vdW[0] == pytest.approx(1.3917417470792187e-14, rel=1e-14)
    assert L[5].vdW[1] == pytest.approx(0.227, rel=1e-12)


def test_vald_various_formats():
    """test/linelist.jl:225-247: short/long, extract all/stellar, air wavenumbers, quotes."""
    from korg_line.linelist import read_vald_linelist
    rd = lambda f: read_vald_linelist(os.path.join(VALD, f)).to_lines()
    short_all = rd("short-extract-all.vald")
    long_air = rd("long-extract-all-air-wavenumber.vald")
    long_air_nq = rd("long-extract-all-air-wavenumber-noquotes.vald")
    short_stellar = rd("short-extract-stellar.vald")
    long_stellar = rd("long-extract-stellar.vald")
    assert len(short_all) == len(short_stellar) == len(long_air) == len(long_air_nq) == \
        len(long_stellar) == 2
    key = lambda l: (l.wl, l.log_gf, l.species, l.E_lower, l.gamma_rad, l.gamma_stark, l.vdW)
    assert key(long_air[0]) == key(long_air_nq[0])
    assert key(short_all[0]) == key(short_stellar[0]) == key(long_stellar[0])
    # with a unit conversion (air wavelengths, cm^-1 energies): approximately equal
    assert short_all[0].wl == pytest.approx(long_air[0].wl, rel=1e-7)
    assert short_all[0].log_gf == long_air[0].log_gf and short_all[0].species == long_air[0].species
    assert short_all[0].E_lower == pytest.approx(long_air[0].E_lower, abs=1e-3)
    assert short_all[0].gamma_rad == long_air[0].gamma_rad
    assert short_all[0].gamma_stark == long_air[0].gamma_stark and short_all[0].vdW == long_air[0].vdW


def test_air_to_vacuum_against_the_published_table_value():
    """Birch & Downs: n(air) ~ 1.000278 near 5000 A; vacuum_to_air(air_to_vacuum) needs none."""
    from korg_line.linelist import air_to_vacuum
    v = air_to_vacuum(5000.0)
    assert 1.00027 < v / 5000.0 < 1.00029
    assert air_to_vacuum(5000e-8) == pytest.approx(v * 1e-8, rel=1e-15)


# ------------------------------------------------------------------------------------ Kurucz
KURUCZ = os.path.join(HERE, "golden", "kurucz")  # the reference's test files (test/data/linelists)


@pytest.mark.parametrize("fname", ["gfallvac08oct17.stub.dat", "gfallvac08oct17-missing-col.stub.dat"])
def test_kurucz_linelist_parsing(fname):
    """test/linelist.jl:154-166, value by value (the file's isotope-tagged rows draw the
    default read's warning; its first line names no isotope)."""
    from korg_line.linelist import read_kurucz_linelist
    with pytest.warns(UserWarning, match="name an isotope"):
        ll = read_kurucz_linelist(os.path.join(KURUCZ, fname))
    assert ll.is_sorted() and len(ll) == 987
    L0 = ll.to_lines()[0]
    assert L0.wl == pytest.approx(0.0007234041763337705, rel=1e-14)
    assert L0.log_gf == -0.826 and L0.species == Species("Be_II")
    assert L0.E_lower == pytest.approx(17.360339371573698, rel=1e-14)
    assert L0.gamma_rad == pytest.approx(8.511380382023759e7, rel=1e-14)
    assert L0.gamma_stark == pytest.approx(0.003890451449942805, rel=1e-14)
    assert L0.vdW[0] == pytest.approx(1.2302687708123812e-7, rel=1e-14)


def test_kurucz_short_lines_and_molecular():
    """test/linelist.jl:168-189: stripped rows parse; a molecular list is refused."""
    from korg_line.linelist import read_kurucz_linelist
    ll = read_kurucz_linelist(os.path.join(KURUCZ, "gfallvac08oct17-short-lines.stub.dat"),
                              isotopic_abundances=None)
    assert len(ll) >= 2
    L0 = ll.to_lines()[0]
    assert L0.wl == pytest.approx(0.0007234041763337705, rel=1e-14)
    assert L0.log_gf == -0.826 and L0.species == Species("Be_II")
    with pytest.raises(ValueError):
        read_kurucz_linelist(os.path.join(KURUCZ, "kurucz_cn.txt"))


def test_kurucz_isotope_column_and_table(tmp_path):
    """src/linelist.jl:478-495: with isotopic_abundances = nothing Kurucz's isotope column is added
    to log gf; with a table, log10 of the named isotope's fraction (nothing when the table lacks
    it).  The default adds the column and warns when a row names an isotope."""
    import warnings
    from korg_line.linelist import read_kurucz_linelist
    row = open(os.path.join(KURUCZ, "gfallvac08oct17.stub.dat")).readline().rstrip("\n")
    base = float(row[11:18])  # no hyperfine adjustment in this row
    r = list(row.ljust(160))
    r[106:109] = " 9"[:3].rjust(3)   # isotope 9 of Be
    r[118:124] = "-0.100"           # Kurucz's isotope adjustment
    path = tmp_path / "iso.dat"
    path.write_text("".join(r) + "\n")
    with pytest.warns(UserWarning, match="name an isotope"):
        ll = read_kurucz_linelist(str(path))
    assert ll.to_lines()[0].log_gf == pytest.approx(base - 0.1, abs=1e-12)
    with warnings.catch_warnings():
        warnings.simplefilter("error")
        ll = read_kurucz_linelist(str(path), isotopic_abundances=None)
        assert ll.to_lines()[0].log_gf == pytest.approx(base - 0.1, abs=1e-12)
        ll = read_kurucz_linelist(str(path), isotopic_abundances={4: {9: 0.5}})
        assert ll.to_lines()[0].log_gf == pytest.approx(base + np.log10(0.5), abs=1e-12)
        ll = read_kurucz_linelist(str(path), isotopic_abundances={4: {7: 0.5}})
        assert ll.to_lines()[0].log_gf == pytest.approx(base, abs=1e-12)
