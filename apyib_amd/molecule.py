"""Molecules and minimal basis sets without psi4 (host side).

Parses the subset of the psi4 geometry format used by the reference's
apyib/data/molecules.py (symbol x y z lines, `units bohr|angstrom`,
`symmetry c1`, `no_com`/`nocom`, `no_reorient`/`noreorient`, optional
charge/multiplicity line), and builds STO-3G (H, C, N, O) and STO-6G (H)
basis sets from the published Hehre-Stewart-Pople contraction tables, and
6-31G / 6-31G* (H, O; Cartesian d functions) from the Pople tables.
Coordinates are used as given (the reference's molecules all disable
re-centering and re-orientation).
"""
import numpy as np

from .gaussian import ANG, BasisFunction

BOHR_PER_ANGSTROM = 1.0 / 0.52917721067

Z = {'H': 1, 'HE': 2, 'LI': 3, 'BE': 4, 'B': 5, 'C': 6, 'N': 7, 'O': 8, 'F': 9, 'NE': 10}
N_CORE = {1: 0, 2: 0, 3: 1, 4: 1, 5: 1, 6: 1, 7: 1, 8: 1, 9: 1, 10: 1}

_S3 = (0.15432897, 0.53532814, 0.44463454)
_SP3_S = (-0.09996723, 0.39951283, 0.70011547)
_SP3_P = (0.15591627, 0.60768372, 0.39195739)

# 6-31G (Hehre-Ditchfield-Pople; split valence, sp shells sharing exponents) for H and O, and the d
# polarisation exponent 0.8 of 6-31G* on O (Hariharan-Pople), to the digits of the basis-set library files
# the reference's psi4 reads (the RHF/6-31G(d) golden of test_002_RHF.py:73 fixes them to 1e-12 Eh).
_H_631G = [((0,), (18.7311370, 2.8253937, 0.6401217), ((0.03349460, 0.23472695, 0.81375733),)),
           ((0,), (0.1612778,), ((1.0,),))]
_O_631G = [((0,), (5484.6717, 825.23495, 188.04696, 52.9645, 16.89757, 5.7996353),
            ((0.0018311, 0.0139501, 0.0684451, 0.2327143, 0.4701930, 0.3585209),)),
           ((0, 1), (15.539616, 3.5999336, 1.0137618),
            ((-0.1107775, -0.1480263, 1.1307670), (0.0708743, 0.3397528, 0.7271586))),
           ((0, 1), (0.2700058,), ((1.0,), (1.0,)))]

# element -> list of (angular momentum list, exponents, coefficients per l)
BASIS = {
    'STO-3G': {
        1: [((0,), (3.42525091, 0.62391373, 0.16885540), (_S3,))],
        6: [((0,), (71.6168370, 13.0450960, 3.5305122), (_S3,)),
            ((0, 1), (2.9412494, 0.6834831, 0.2222899), (_SP3_S, _SP3_P))],
        7: [((0,), (99.1061690, 18.0523120, 4.8856602), (_S3,)),
            ((0, 1), (3.7804559, 0.8784966, 0.2857144), (_SP3_S, _SP3_P))],
        8: [((0,), (130.7093200, 23.8088610, 6.4436083), (_S3,)),
            ((0, 1), (5.0331513, 1.1695961, 0.3803890), (_SP3_S, _SP3_P))],
    },
    'STO-6G': {
        1: [((0,), (35.52322122, 6.513143725, 1.822142904, 0.6259552659, 0.2430767471, 0.1001124280),
             ((0.00916359628, 0.04936149294, 0.16853830490, 0.37056279970, 0.41649152980, 0.13033408410),))],
    },
    '6-31G': {1: _H_631G, 8: _O_631G},
    '6-31G*': {1: _H_631G, 8: _O_631G + [((2,), (0.8,), ((1.0,),))]},   # Cartesian d (six functions)
}
BASIS['6-31G(D)'] = BASIS['6-31G*']      # psi4 reads both names as the same basis


class Molecule:
    def __init__(self, symbols, coords_bohr, charge=0, multiplicity=1):
        self.symbols = [s.upper() for s in symbols]
        self.Z = np.array([Z[s] for s in self.symbols], dtype=float)
        self.coords = np.asarray(coords_bohr, dtype=float).reshape(-1, 3)
        self.charge = charge
        self.multiplicity = multiplicity

    @classmethod
    def from_string(cls, text):
        units = 'angstrom'
        symbols, xyz = [], []
        charge, mult = 0, 1
        for raw in text.strip().splitlines():
            line = raw.split('#')[0].strip()
            if not line:
                continue
            tok = line.split()
            key = tok[0].lower()
            if key == 'units':
                units = tok[1].lower()
                continue
            if key in ('symmetry', 'no_com', 'nocom', 'no_reorient', 'noreorient'):
                continue
            if len(tok) == 2 and all(t.lstrip('-').isdigit() for t in tok):
                charge, mult = int(tok[0]), int(tok[1])
                continue
            symbols.append(tok[0])
            xyz.append([float(t) for t in tok[1:4]])
        xyz = np.array(xyz, dtype=float)
        if units in ('angstrom', 'ang', 'a'):
            xyz = xyz * BOHR_PER_ANGSTROM
        elif units not in ('bohr', 'au', 'a.u.'):
            raise ValueError(f"units {units!r}")
        return cls(symbols, xyz, charge, mult)

    def natom(self):
        return len(self.symbols)

    def true_atomic_number(self, i):
        return int(self.Z[i])

    def nuclear_repulsion_energy(self, field=(0.0, 0.0, 0.0)):
        """sum_{A<B} Z_A Z_B / R_AB, plus the nuclear dipole in a uniform
        electric field: -sum_A Z_A (F . R_A) with F as passed."""
        E = 0.0
        n = self.natom()
        for a in range(n):
            for b in range(a):
                E += self.Z[a] * self.Z[b] / np.linalg.norm(self.coords[a] - self.coords[b])
        F = np.asarray(field, dtype=float)
        E += float(np.sum(self.Z * (self.coords @ F)))
        return E

    def with_coords(self, coords_bohr):
        return Molecule(self.symbols, coords_bohr, self.charge, self.multiplicity)

    def geometry_string(self):
        lines = [f"{s} {x:.15f} {y:.15f} {z:.15f}" for s, (x, y, z) in zip(self.symbols, self.coords)]
        return "\n".join(lines + ["symmetry c1", "no_com", "no_reorient", "units bohr"])


class BasisSet:
    """Contracted Cartesian basis on a molecule (shell order as listed)."""

    def __init__(self, molecule, name):
        key = name.upper()
        if key not in BASIS:
            raise ValueError(f"basis {name!r} not available (have {sorted(BASIS)})")
        table = BASIS[key]
        self.name = key
        self.molecule = molecule
        self.functions = []
        self.atom_of = []
        for a, (z, R) in enumerate(zip(molecule.Z, molecule.coords)):
            z = int(z)
            if z not in table:
                raise ValueError(f"basis {name} has no entry for Z={z}")
            for ls, exps, coefs in table[z]:
                for l, c in zip(ls, coefs):
                    for lmn in ANG[l]:
                        self.functions.append(BasisFunction(R, lmn, exps, c))
                        self.atom_of.append(a)

    def nbf(self):
        return len(self.functions)

    def n_frozen_core(self):
        return int(sum(N_CORE[int(z)] for z in self.molecule.Z))

    def __eq__(self, other):
        return (isinstance(other, BasisSet) and self.name == other.name
                and np.array_equal(self.molecule.coords, other.molecule.coords))

    def __ne__(self, other):
        return not self.__eq__(other)
