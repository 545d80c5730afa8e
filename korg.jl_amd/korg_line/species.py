"""Chemical species: formula, charge, mass and the molecule test used on the line-opacity path.

Follows ``src/species.jl:8-275`` (``Formula``, ``Species``, ``get_mass``, ``ismolecule``) and the
atomic-mass table of ``src/atomic_data.jl:16-26``.  Only what the opacity path and its input
preparation need is provided.
"""
from .constants import amu_cgs

atomic_symbols = [
    "H", "He", "Li", "Be", "B", "C", "N", "O", "F", "Ne", "Na", "Mg", "Al", "Si", "P", "S", "Cl",
    "Ar", "K", "Ca", "Sc", "Ti", "V", "Cr", "Mn", "Fe", "Co", "Ni", "Cu", "Zn", "Ga", "Ge", "As",
    "Se", "Br", "Kr", "Rb", "Sr", "Y", "Zr", "Nb", "Mo", "Tc", "Ru", "Rh", "Pd", "Ag", "Cd", "In",
    "Sn", "Sb", "Te", "I", "Xe", "Cs", "Ba", "La", "Ce", "Pr", "Nd", "Pm", "Sm", "Eu", "Gd", "Tb",
    "Dy", "Ho", "Er", "Tm", "Yb", "Lu", "Hf", "Ta", "W", "Re", "Os", "Ir", "Pt", "Au", "Hg", "Tl",
    "Pb", "Bi", "Po", "At", "Rn", "Fr", "Ra", "Ac", "Th", "Pa", "U"]
atomic_numbers = {s: i + 1 for i, s in enumerate(atomic_symbols)}

# Standard atomic weights in amu (src/atomic_data.jl:16-26); each is scaled by amu_cgs exactly as
# the reference does (one IEEE product per element).
_atomic_weights = [
    1.008, 4.003, 6.941, 9.012, 10.81, 12.01, 14.01, 16.00, 19.00, 20.18,
    22.99, 24.31, 26.98, 28.08, 30.97, 32.06, 35.45, 39.95, 39.10, 40.08,
    44.96, 47.90, 50.94, 52.00, 54.94, 55.85, 58.93, 58.71, 63.55, 65.37,
    69.72, 72.59, 74.92, 78.96, 79.90, 83.80, 85.47, 87.62, 88.91, 91.22,
    92.91, 95.94, 98.91, 101.1, 102.9, 106.4, 107.9, 112.4, 114.8, 118.7,
    121.8, 127.6, 126.9, 131.3, 132.9, 137.3, 138.9, 140.1, 140.9, 144.2,
    145.0, 150.4, 152.0, 157.3, 158.9, 162.5, 164.9, 167.3, 168.9, 173.0,
    175.0, 178.5, 181.0, 183.9, 186.2, 190.2, 192.2, 195.1, 197.0, 200.6,
    204.4, 207.2, 209.0, 210.0, 210.0, 222.0, 223.0, 226.0, 227.0, 232.0,
    231.0, 238.0]
atomic_masses = [w * amu_cgs for w in _atomic_weights]

MAX_ATOMS_PER_MOLECULE = 6
roman_numerals = ["I", "II", "III", "IV", "V", "VI", "VII", "VIII", "IX", "X"]


def _parse_formula(code):
    """Atomic numbers (ascending) of a formula code such as "Fe", "H2", "FeH", "0801", "26"."""
    if code in atomic_numbers:
        return (atomic_numbers[code],)
    if code.isdigit():
        if len(code) <= 2:
            return (int(code),)
        if len(code) <= 4:
            a, b = int(code[:-2]), int(code[-2:])
            return (min(a, b), max(a, b))
        if len(code) % 2 == 1:
            code = "0" + code
        els = sorted(int(code[i:i + 2]) for i in range(0, len(code), 2))
        if len(els) > MAX_ATOMS_PER_MOLECULE:
            raise ValueError(f"at most {MAX_ATOMS_PER_MOLECULE} nuclei are supported ({code})")
        return tuple(els)
    starts = [i for i, c in enumerate(code) if c.isdigit() or c.isupper()] + [len(code)]
    atoms = []
    for a, b in zip(starts[:-1], starts[1:]):
        tok = code[a:b]
        if tok.isdigit():
            atoms += [atoms[-1]] * (int(tok) - 1)
        else:
            if tok not in atomic_numbers:
                raise ValueError(f"unknown element {tok!r} in {code!r}")
            atoms.append(atomic_numbers[tok])
    if not atoms or len(atoms) > MAX_ATOMS_PER_MOLECULE:
        raise ValueError(f"can't parse formula {code!r}")
    return tuple(sorted(atoms))


class Species:
    """An atom or molecule with a charge (``src/species.jl:181-262``).

    ``Species("Fe I")``, ``Species("Fe_II")``, ``Species("H2")``, ``Species("26.01")`` all parse as
    in Korg; ``charge`` is 0 for neutrals.
    """
    __slots__ = ("atoms", "charge")

    def __init__(self, code, charge=None):
        if charge is not None:
            self.atoms = tuple(sorted(code)) if not isinstance(code, str) else _parse_formula(code)
            self.charge = int(charge)
        else:
            self.atoms, self.charge = self._parse(str(code))
        if self.charge < -1:
            raise ValueError("charge < -1 is not supported")

    @staticmethod
    def _parse(code):
        code = code.strip("0 ")
        if code.endswith("+"):
            code = code[:-1] + " 2"
        elif code.endswith("-"):
            code = code[:-1] + " 0"
        toks = [t for t in code.replace(".", " ").replace("_", " ").split(" ") if t]
        if len(toks) > 2 or not toks:
            raise ValueError(f"{code} isn't a valid species code")
        atoms = _parse_formula(toks[0])
        if len(toks) == 1:
            return atoms, 0
        charge = roman_numerals.index(toks[1]) + 1 if toks[1] in roman_numerals else int(toks[1])
        try:
            float(code)
            numeric = True
        except ValueError:
            numeric = False
        if not numeric:
            charge -= 1
        return atoms, charge

    def ismolecule(self):
        """``src/species.jl:168``: more than one nucleus."""
        return len(self.atoms) > 1

    def get_mass(self):
        """``src/species.jl:175``: sum of the nuclei's masses in g (ascending-Z order)."""
        m = 0.0
        for z in self.atoms:
            m = m + atomic_masses[z - 1]
        return m

    def get_atoms(self):
        return self.atoms

    def __eq__(self, other):
        return isinstance(other, Species) and self.atoms == other.atoms and \
            self.charge == other.charge

    def __hash__(self):
        return hash((self.atoms, self.charge))

    def __repr__(self):
        if len(self.atoms) == 1:
            name = atomic_symbols[self.atoms[0] - 1]
        else:
            name = "".join(atomic_symbols[z - 1] for z in self.atoms)
        if 0 <= self.charge < len(roman_numerals):
            return f"{name} {roman_numerals[self.charge]}"
        return f"{name} {self.charge}"


def species(code):
    """Shorthand mirroring Korg's ``species"..."`` string macro."""
    return Species(code)
