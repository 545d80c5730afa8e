"""Extract the plain-text physical data this package needs from a Korg.jl checkout.

Run once, in a container that has the reference checkout, to (re)generate the
``*.tsv`` files of ``korg.jl_amd/korg_line/data``:

    python tests/golden/extract_reference_data.py /root/reference

Only data is extracted (numbers from Korg's data files), never code:

* ``ionization_energies.tsv`` from
  ``data/barklem_collet_2016/BarklemCollet2016-ionization_energies.dat``
  (used by the default Unsoeld/Cowley broadening, ``src/linelist.jl:149-183``).
* ``atomic_partition.tsv`` -- rows of ``test/data/BarklemCollet2016-atomic_partition.dat``
  for the species of the synthetic workloads (stand-in partition functions).
* ``<model>_layers.tsv`` -- the depth structure and partial pressures of the
  MARCS model atmospheres held in ``test/data/*.mod`` (stand-in atmospheres).
"""
import os
import sys

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "..", "korg.jl_amd",
                    "korg_line", "data")

PARTITION_SPECIES = ["H_I", "H_II", "He_I", "He_II", "Ca_I", "Ca_II", "Ca_III", "Ti_I", "Ti_II",
                     "Ti_III", "Fe_I", "Fe_II", "Fe_III", "Ni_I", "Ni_II", "Ni_III"]

MODELS = {
    "sun": "test/data/sun.mod",
    "s5000_g3.0_m1.0": "test/data/s5000_g+3.0_m1.0_t02_st_z+0.00_a+0.00_c+0.00_n+0.00_o+0.00_r+0.00_s+0.00.mod",
    "s6000_g1.0_m0.5": "test/data/s6000_g+1.0_m0.5_t05_st_z+0.00_a+0.00_c+0.00_n+0.00_o+0.00_r+0.00_s+0.00.mod",
}


def extract_ionization(ref):
    src = os.path.join(ref, "data/barklem_collet_2016/BarklemCollet2016-ionization_energies.dat")
    out = ["# Z\tsymbol\tIE1_eV\tIE2_eV\tIE3_eV  (Barklem & Collet 2016; -1 = not available)"]
    for line in open(src):
        if line.startswith("#") or not line.strip():
            continue
        t = line.split()
        out.append("\t".join(t[:5]))
    with open(os.path.join(HERE, "ionization_energies.tsv"), "w") as f:
        f.write("\n".join(out) + "\n")


def extract_partition(ref):
    src = os.path.join(ref, "test/data/BarklemCollet2016-atomic_partition.dat")
    temps = None
    rows = {}
    for line in open(src):
        if line.startswith("#   T [K]"):
            temps = line.split()[3:]
        elif not line.startswith("#") and line.strip():
            t = line.split()
            if t[0] in PARTITION_SPECIES:
                rows[t[0]] = t[1:]
    out = ["# species\tU(T) at T = columns of the first data row (Barklem & Collet 2016 table 8)",
           "T_K\t" + "\t".join(temps)]
    for s in PARTITION_SPECIES:
        out.append(s + "\t" + "\t".join(rows[s]))
    with open(os.path.join(HERE, "atomic_partition.tsv"), "w") as f:
        f.write("\n".join(out) + "\n")


def extract_model(ref, name, rel):
    lines = open(os.path.join(ref, rel)).read().splitlines()
    teff = float(lines[1].split()[0])
    g = float(lines[3].split()[0])
    vmic = float(lines[4].split()[0])
    mh = float(lines[6].split()[0])
    i = next(k for k, l in enumerate(lines) if l.startswith("Logarithmic chemical number abundances"))
    abund = []
    k = i + 1
    while len(abund) < 92:
        abund += [float(x) for x in lines[k].split()]
        k += 1
    nlay = int(lines[k].split()[0])
    s1 = next(j for j, l in enumerate(lines) if l.startswith(" k lgTauR  lgTau5"))
    s3 = next(j for j, l in enumerate(lines) if l.startswith("Assorted logarithmic partial"))
    head3 = lines[s3 + 1].split()
    i_hi, i_h2 = head3.index("I") , head3.index("H2")  # columns: k lgPgas H I H- H2 ...
    out = [f"# MARCS model {os.path.basename(rel)}: Teff={teff} g={g} vmic={vmic} [M/H]={mh}",
           "# abundances(A_X, Z=1..92)\t" + " ".join(f"{a:.2f}" for a in abund),
           "k\tlgTau5\tT_K\tPe\tPg\tlgP_HI\tlgP_H2"]
    for j in range(nlay):
        a = lines[s1 + 1 + j].split()
        p = lines[s3 + 2 + j].split()
        # partial-pressure rows: k lgPgas "H I" H- H2 ...; the header splits "H I" in two tokens
        lgp_hi = p[2]
        lgp_h2 = p[4]
        out.append("\t".join([a[0], a[2], a[4], a[5], a[6], lgp_hi, lgp_h2]))
    del i_hi, i_h2
    with open(os.path.join(HERE, f"{name}_layers.tsv"), "w") as f:
        f.write("\n".join(out) + "\n")


if __name__ == "__main__":
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    extract_ionization(ref)
    extract_partition(ref)
    for name, rel in MODELS.items():
        extract_model(ref, name, rel)
    print("wrote data files to", HERE)
