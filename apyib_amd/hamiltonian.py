"""Atomic-orbital Hamiltonian with apyib's interface (hamiltonian.py:8-70),
built from this package's own integral engine instead of psi4.

Attributes match the reference: T, V, ERI, S (NumPy, AO basis), E_nuc,
basis_set (nbf(), n_frozen_core()), molecule (natom(),
true_atomic_number(i), nuclear_repulsion_energy(field)).  Uniform electric
and magnetic fields enter V as in hamiltonian.py:41-70:
    V -= F_el . mu_el,   mu_el = -<m| r |n>           (electron charge -1)
    V -= F_mag . mu_mag, mu_mag = -1/2 i <m| L |n>     with L the angular
         momentum integrals in the convention of hamiltonian.py:55-57.
The phase-space extensions (hamiltonian.py:72-538) are not part of this path.
"""

import numpy as np

from . import gaussian
from .molecule import BasisSet, Molecule

# <m| L |n> as passed to -1/2 i (.) in hamiltonian.py:57: the real
# antisymmetric matrix of r x nabla times this sign.  -1 reproduces the
# reference's finite-difference AAT arrays (test_011_AAT.py,
# test_013_AAT_parallel.py); +1 flips the sign of every element
# (tests/test_oracle_aat.py).
ANGULAR_MOMENTUM_SIGN = -1.0


class Hamiltonian(object):
    def __init__(self, parameters, integrals=None):
        """integrals: None (built here by the host engine) or the AO arrays
        (S, T, V, D[3], L[3]) of this molecule and basis, L = <m|r x nabla|n>
        (as build_hamiltonians passes them from the GPU)."""
        self.molecule = Molecule.from_string(parameters['geom'])
        self.basis_set = BasisSet(self.molecule, parameters['basis'])
        fns = self.basis_set.functions
        bs = self.basis_set
        if integrals is None:
            charges = list(zip(self.molecule.Z, self.molecule.coords))
            self.S, self.T, self.V = (bs.to_basis(x) for x in gaussian.build_one_electron(fns, charges))
            D = L = None
        else:
            self.S, self.T, self.V, D, L = (None if x is None else bs.to_basis(x) for x in integrals)
        self._ERI = None

        F_el = list(parameters.get('F_el', [0.0, 0.0, 0.0]))
        F_mag = list(parameters.get('F_mag', [0.0, 0.0, 0.0]))
        E_field = any(f != 0.0 for f in F_el)
        M_field = any(f != 0.0 for f in F_mag)
        if E_field:
            D = bs.to_basis(gaussian.build_dipole(fns)) if D is None else D
            self.mu_el = [-D[a] for a in range(3)]
        if M_field:
            L = ANGULAR_MOMENTUM_SIGN * (bs.to_basis(gaussian.build_angular_momentum(fns)) if L is None else L)
            self.mu_mag = [-0.5j * L[a] for a in range(3)]
        self.E_nuc = self.molecule.nuclear_repulsion_energy([-f for f in F_el])
        for a in range(3):
            if E_field:
                self.V = self.V - F_el[a] * self.mu_el[a]
            if M_field:
                self.V = self.V - F_mag[a] * self.mu_mag[a]
        self._device_cache = {}

    @property
    def ERI(self):
        """(mn|ls) on the host (NumPy), built on first use; the GPU path
        reads device_eri instead."""
        if self._ERI is None:
            self._ERI = self.basis_set.eri_to_basis(gaussian.build_eri(self.basis_set.functions))
        return self._ERI

    def device_eri(self, device):
        """The AO integrals as a complex128 device tensor, computed on the
        device (apy_eri_gto) once per device."""
        from .integrals import ao_eri, eri_to_basis
        key = str(device)
        t = self._device_cache.get(key)
        if t is None:
            t = eri_to_basis(self.basis_set, ao_eri(self.basis_set, self.molecule.coords, device))
            self._device_cache[key] = t
        return t


def build_hamiltonians(parameter_list, device):
    """[Hamiltonian(p) for p in parameter_list] with the one-electron
    integrals (S, T, V, dipole, angular momentum) of all of them computed on
    the device, one apy_oei_gto launch per basis (the same functions on the
    same atoms at different geometries), and one copy to the host."""
    from .integrals import hamiltonian_integrals
    mols = [Molecule.from_string(p['geom']) for p in parameter_list]
    groups = {}
    for q, (p, m) in enumerate(zip(parameter_list, mols)):
        groups.setdefault((str(p['basis']).upper(), tuple(int(z) for z in m.Z)), []).append(q)
    ints = [None] * len(parameter_list)
    for qs in groups.values():
        bs = BasisSet(mols[qs[0]], parameter_list[qs[0]]['basis'])
        coords = np.stack([np.asarray(mols[q].coords, dtype=np.float64) for q in qs])
        out = hamiltonian_integrals(bs, coords, mols[qs[0]].Z, device).cpu().numpy()
        for k, q in enumerate(qs):
            o = out[k]
            sym = lambda x: 0.5 * (x + x.swapaxes(-1, -2))
            ints[q] = (sym(o[0]), sym(o[1]), sym(o[2]), sym(o[3:6]), 0.5 * (o[6:9] - o[6:9].swapaxes(-1, -2)))
    return [Hamiltonian(p, integrals=x) for p, x in zip(parameter_list, ints)]
