"""Hydrogen lines on the device: synthesize's hydrogen-line loop (``src/synthesize.jl:293-301``)
over ``hydrogen_line_absorption!`` (``src/hydrogen_line_absorption.jl:75-169``) -- the Stehle &
Hutcheon (1999) Stark profiles of the Lyman, Balmer and Paschen series to n = 30, the Barklem et al.
(2000) resonance profiles of H alpha, beta and gamma, the MHD occupation probabilities, and the
Brackett series to the tables' highest upper level (Griem's impact and Holtsmark's quasistatic
profiles convolved, with the Doppler profile where it matters).

``StarkTables`` packs the tables of ``Stehle-Hutchson-hydrogen-profiles.h5`` (shipped in
``korg_line/data``) for ``korg_set_hydrogen_tables`` as Korg loads them
(``src/hydrogen_line_absorption.jl:6-39``): the transitions in the file's (name) order, log of
each profile with log 0 clipped to -700, log(delta_nu / F0) knots (the first, for delta_nu = 0,
stands for -floatmax and is not read), line-centre tables in cm.
"""
import ctypes
import os

import numpy as np

from . import _lib
from .hdf5 import H5File

DEFAULT_TABLES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data",
                              "Stehle-Hutchson-hydrogen-profiles.h5")


class StarkTables:
    def __init__(self, path=DEFAULT_TABLES):
        f = H5File(path)
        temps, nes, dnus, profs, lams, recs = [], [], [], [], [], []
        off = dict(T=0, ne=0, dnu=0, P=0, L=0)
        self.names = f.keys()  # Korg maps over keys(fid): the transitions in name order
        for name in self.names:
            a = f.attrs(name)
            T = f.read(name + "/temps")
            ne = f.read(name + "/electron_number_densities")
            d = f.read(name + "/delta_nu_over_F0")
            P = f.read(name + "/profile")
            L0 = f.read(name + "/lambda0")
            if P.shape != (d.size, ne.size, T.size) or L0.shape != (ne.size, T.size):
                raise ValueError(f"{path}: table {name} has inconsistent axes")
            with np.errstate(divide="ignore"):
                logP = np.log(P)
            logP[logP == -np.inf] = -700.0
            logd = np.empty_like(d)
            logd[0] = 0.0  # stands for -floatmax (delta_nu = 0); not read
            logd[1:] = np.log(d[1:])
            recs.append(_lib.KorgHTransition(int(a["lower"]), int(a["upper"]), T.size, ne.size,
                                             d.size, 0, float(a["log_gf"]), off["T"], off["ne"],
                                             off["dnu"], off["P"], off["L"]))
            for key, arr, lst in (("T", T, temps), ("ne", ne, nes), ("dnu", logd, dnus),
                                  ("P", logP.ravel(), profs), ("L", L0.ravel() * 1e-8, lams)):
                lst.append(np.asarray(arr, dtype=np.float64))
                off[key] += arr.size
        self.transitions = (_lib.KorgHTransition * len(recs))(*recs)
        self.temps, self.n_e, self.log_dnu, self.log_profile, self.lambda0 = (
            np.ascontiguousarray(np.concatenate(x)) for x in (temps, nes, dnus, profs, lams))
        p = lambda x: x.ctypes.data
        self.struct = _lib.KorgHTables(len(recs), self.transitions, p(self.temps), p(self.n_e),
                                       p(self.log_dnu), p(self.log_profile), p(self.lambda0),
                                       self.temps.size, self.n_e.size, self.log_dnu.size,
                                       self.log_profile.size, self.lambda0.size)

    def __len__(self):
        return len(self.names)
