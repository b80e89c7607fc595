"""Molecules and minimal basis sets without psi4 (host side).

Parses the subset of the psi4 geometry format used by the reference's
apyib/data/molecules.py (symbol x y z lines, `units bohr|angstrom`,
`symmetry c1`, `no_com`/`nocom`, `no_reorient`/`noreorient`, optional
charge/multiplicity line), and builds STO-3G (H, C, N, O) and STO-6G (H, O)
basis sets from the Hehre-Stewart-Pople contraction tables, and
6-31G / 6-31G* (H, O; Cartesian d functions) from the Pople tables.
Coordinates are used as given (the reference's molecules all disable
re-centering and re-orientation).
"""
import numpy as np

from .gaussian import ANG, BasisFunction, RawGaussian

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

# STO-6G for oxygen: the Hehre-Stewart-Pople least-squares Gaussian expansions of unit-exponent Slater
# functions (1s; 2s and 2p sharing exponents), scaled by zeta^2 with the standard molecular exponents
# zeta_1s = 7.66, zeta_2sp = 2.25 (the scaling that gives STO-3G's O exponents above).  The 1s expansion is
# the published one (H above is the same with zeta = 1.24); the 2sp expansion is the minimiser of
# |2s - sum c g_1s|^2 + |2p - sum d g_2p|^2 over the six shared exponents, the published criterion
# (tests/test_sto6g.py checks stationarity and that the same fit reproduces the 1s table to 5e-9).
_STO6G_1S = (23.10303149, 4.235915534, 1.185056519, 0.4070988982, 0.1580884151, 0.06510953954)
_STO6G_1S_C = (0.00916359628, 0.04936149294, 0.16853830490, 0.37056279970, 0.41649152980, 0.13033408410)
_STO6G_2SP = (10.308693172995362, 2.0403594543359214, 0.6341422081020324, 0.24397736732149317,
              0.10595953752866882, 0.04856900874863939)
_STO6G_2S_C = (-0.013252775446846133, -0.04699166331605596, -0.03378533564215647, 0.2502415347244665,
               0.5951166505895096, 0.2407059357651443)
_STO6G_2P_C = (0.003759694627789826, 0.03767934829828925, 0.17389663899514973, 0.4180361780835387,
               0.42585928682096597, 0.10170823445499358)


def _scaled(exps, zeta):
    return tuple(a * zeta * zeta for a in exps)


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
        8: [((0,), _scaled(_STO6G_1S, 7.66), (_STO6G_1S_C,)),
            ((0, 1), _scaled(_STO6G_2SP, 2.25), (_STO6G_2S_C, _STO6G_2P_C))],
    },
    '6-31G': {1: _H_631G, 8: _O_631G},
    '6-31G*': {1: _H_631G, 8: _O_631G + [((2,), (0.8,), ((1.0,),))]},   # Cartesian d (six functions)
}
BASIS['6-31G(D)'] = BASIS['6-31G*']      # psi4 reads both names as the same basis

# cc-pVDZ (Dunning, J. Chem. Phys. 90, 1007 (1989)) for H and O: generally contracted s and p shells
# ([9s4p1d/3s2p1d] on O, [4s1p/2s1p] on H) with spherical d functions, as psi4's library declares it
# (puream).  The first contractions carry every primitive of their shell; the span, and so every energy,
# is that of psi4's segmented copy, whose contractions differ by multiples of the uncontracted functions.
_O_S = (11720.0, 1759.0, 400.8, 113.7, 37.03, 13.27, 5.025, 1.013, 0.3023)
_O_P = (17.70, 3.854, 1.046, 0.2753)
_H_S = (13.01, 1.962, 0.4446, 0.1220)
BASIS['CC-PVDZ'] = {
    1: [((0,), _H_S, ((0.019685, 0.137977, 0.478148, 0.501240),)),
        ((0,), (0.1220,), ((1.0,),)),
        ((1,), (0.727,), ((1.0,),))],
    8: [((0,), _O_S, ((0.000710, 0.005470, 0.027837, 0.104800, 0.283062, 0.448719, 0.270952, 0.015458,
                        -0.002585),)),
        ((0,), _O_S, ((-0.000160, -0.001263, -0.006267, -0.025716, -0.070924, -0.165411, -0.116955, 0.557368,
                        0.572759),)),
        ((0,), (0.3023,), ((1.0,),)),
        ((1,), _O_P[:3], ((0.043018, 0.228913, 0.508728),)),
        ((1,), (0.2753,), ((1.0,),)),
        ((2,), (1.185,), ((1.0,),))],
}
# bases whose d shells are spherical (five real solid harmonics) rather than the six Cartesian functions
PURE_D = {'CC-PVDZ'}
# primitives per function the integral kernels take (APY_GTO_MAXPRIM): longer contractions are split into
# pieces with the contraction's own normalised coefficients, summed back by BasisSet.transform
MAX_PRIM = 6

# spherical d as combinations of the Cartesian (xx, xy, xz, yy, yz, zz) functions of one shell (equal
# normalisation for xx, yy, zz): xy, xz, yz, xx - yy, 2 zz - xx - yy; they span the d space without its
# r^2 (s-type) component
_PURE_D = np.array([[0.0, 0.0, 0.0, 1.0, -1.0],
                    [1.0, 0.0, 0.0, 0.0, 0.0],
                    [0.0, 1.0, 0.0, 0.0, 0.0],
                    [0.0, 0.0, 0.0, -1.0, -1.0],
                    [0.0, 0.0, 1.0, 0.0, 0.0],
                    [0.0, 0.0, 0.0, 0.0, 2.0]])


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
        """sum_{A<B} Z_A Z_B / R_AB + sum_A Z_A (R_A . field): psi4's
        nuclear_repulsion_energy(dipole_field), to which hamiltonian.py:60-63
        passes -F_el, giving the nuclear dipole term -F_el . sum_A Z_A R_A."""
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
    """Contracted Gaussian basis on a molecule (shell order as listed).

    `functions` (with `atom_of`) are the Cartesian functions the integral
    engines evaluate.  When the basis has spherical d shells or contractions
    longer than MAX_PRIM primitives, `transform` [len(functions), nbf()] maps
    them to the basis functions (integrals X_basis = T^T X T, per index);
    otherwise it is None and the functions are the basis."""

    def __init__(self, molecule, name):
        key = name.upper()
        if key not in BASIS:
            raise ValueError(f"basis {name!r} not available (have {sorted(BASIS)})")
        table = BASIS[key]
        self.name = key
        self.molecule = molecule
        self.functions = []
        self.atom_of = []
        cols = []                                   # per basis function: [(function index, weight)]
        pure = key in PURE_D
        for a, (z, R) in enumerate(zip(molecule.Z, molecule.coords)):
            z = int(z)
            if z not in table:
                raise ValueError(f"basis {name} has no entry for Z={z}")
            for ls, exps, coefs in table[z]:
                for l, c in zip(ls, coefs):
                    shell = []                      # per Cartesian component: its pieces
                    for lmn in ANG[l]:
                        f = BasisFunction(R, lmn, exps, c)
                        idx = []
                        for p0 in range(0, len(f.exps), MAX_PRIM):
                            self.functions.append(f if len(f.exps) <= MAX_PRIM else
                                                  RawGaussian(f.center, f.lmn, f.exps[p0:p0 + MAX_PRIM],
                                                              f.coefs[p0:p0 + MAX_PRIM]))
                            self.atom_of.append(a)
                            idx.append(len(self.functions) - 1)
                        shell.append(idx)
                    if l == 2 and pure:
                        for k in range(5):
                            cols.append([(q, _PURE_D[m, k]) for m in range(6) for q in shell[m] if _PURE_D[m, k]])
                    else:
                        cols.extend([[(q, 1.0) for q in idx] for idx in shell])
        self.transform = None
        if len(cols) != len(self.functions) or any(len(c) != 1 for c in cols):
            T = np.zeros((len(self.functions), len(cols)))
            for k, col in enumerate(cols):
                for q, w in col:
                    T[q, k] = w
            self.transform = T

    def nbf(self):
        return len(self.functions) if self.transform is None else self.transform.shape[1]

    def to_basis(self, X):
        """One-electron matrices [..., n_fn, n_fn] over `functions` -> over the basis."""
        if self.transform is None:
            return X
        T = self.transform
        return np.einsum('mp,...mn,nq->...pq', T, X, T, optimize=True)

    def eri_to_basis(self, G):
        """(mn|ls) [n_fn]^4 over `functions` (host) -> over the basis."""
        if self.transform is None:
            return G
        T = self.transform
        return np.einsum('mp,nq,lr,gs,mnlg->pqrs', T, T, T, T, G, optimize=True)

    def n_frozen_core(self):
        return int(sum(N_CORE[int(z)] for z in self.molecule.Z))

    def __eq__(self, other):
        return (isinstance(other, BasisSet) and self.name == other.name
                and np.array_equal(self.molecule.coords, other.molecule.coords))

    def __ne__(self, other):
        return not self.__eq__(other)
