"""Writes the small fixtures of tests/golden/ (run from the repository root):

    python tests/golden/make_golden.py

* rng_pins.npy -- the first values of the counter-based generator (seed 1, stream 1), so the
  synthetic workloads stay reproducible across edits;
* solar_small_lines.npz -- the 2,000-line solar_small line list (struct of arrays), the seeded
  input of the 5000-5100 A solar workload.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "korg.jl_amd"))
from korg_line import workloads  # noqa: E402

np.save(os.path.join(HERE, "rng_pins.npy"), workloads.uniform(1, 1, 4))
P, _, _ = workloads.solar_small()
ll = P.lines
np.savez_compressed(os.path.join(HERE, "solar_small_lines.npz"), wl=ll.wl, log_gf=ll.log_gf,
                    E_lower=ll.E_lower, gamma_rad=ll.gamma_rad, gamma_stark=ll.gamma_stark,
                    vdW1=ll.vdW1, vdW2=ll.vdW2, species_index=ll.species_index)
print("wrote", os.listdir(HERE))
