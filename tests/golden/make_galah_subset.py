"""Writes tests/golden/galah_4800_fe_ti_ni_ca.npz (run from the repository root, with the
reference tree present):

    python tests/golden/make_galah_subset.py

The raw columns of the GALAH DR3 linelist file (data/linelists/GALAH_DR3/galah_dr3_linelist.h5 of
the reference) for the Fe, Ti, Ni and Ca I-II lines between 4795 and 4825 A: real lines, with
Korg's packed ABO van der Waals parameters among them, for the species the stand-in atmosphere
provides, so that the device path can be run on them where the reference tree is absent.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "korg.jl_amd"))
from korg_line.hdf5 import H5File  # noqa: E402
from korg_line.linelist import GALAH_COLUMNS  # noqa: E402

SRC = "/root/reference/data/linelists/GALAH_DR3/galah_dr3_linelist.h5"
f = H5File(SRC)
cols = {k: f.read(k) for k in GALAH_COLUMNS}
fo, ion, wl = cols["formula"], cols["ionization"], cols["wl"]
sel = (fo[:, 1] == 0) & np.isin(fo[:, 0], [20, 22, 26, 28]) & np.isin(ion, [1, 2]) & \
    (wl >= 4795) & (wl <= 4825)
idx = np.flatnonzero(sel)
np.savez_compressed(os.path.join(HERE, "galah_4800_fe_ti_ni_ca.npz"), file_index=idx,
                    **{k: v[idx] for k, v in cols.items()})
print(f"wrote {idx.size} lines")
