"""Finite-difference driver with apyib's interface (fin_diff.py:11-372).

`finite_difference(parameters, unperturbed_basis, unperturbed_C)
.compute_AAT(nuc_pert_strength, mag_pert_strength)` returns the reference's
12-tuple (fin_diff.py:267-372): +dR then -dR on every Cartesian coordinate
(alpha = 3 * atom + xyz, bohr), then +dH then -dH on every component of the
magnetic field, each a phase_corrected_energy solve on the GPU (the
Hartree-Fock parts of all of them batched).
Other finite-difference properties of the reference (Hessian, APT, nuclear
and field gradients, phase-space AATs) are not part of this path.
"""
import copy

import numpy as np

from .energy import phase_corrected_energies, phase_corrected_energy
from .molecule import Molecule


class finite_difference(object):
    def __init__(self, parameters, unperturbed_basis, unperturbed_C):
        self.parameters = parameters
        self.molecule = Molecule.from_string(parameters['geom'])
        self.geom = self.molecule.coords.copy()
        self.natom = self.molecule.natom()
        self.unperturbed_basis = unperturbed_basis
        self.unperturbed_C = unperturbed_C

    def _solve(self):
        E_list, T_list, C, basis = phase_corrected_energy(self.parameters, self.unperturbed_basis,
                                                          self.unperturbed_C)
        return C, T_list, basis

    def compute_AAT(self, nuc_pert_strength, mag_pert_strength):
        """fin_diff.py:267-372.  The 6N+6 displaced parameter sets are solved
        as one set (energy.phase_corrected_energies: integrals and SCFs
        batched on the GPU); every solve equals the reference's one-by-one
        phase_corrected_energy."""
        self.parameters.setdefault('F_mag', [0.0, 0.0, 0.0])
        plist = []
        for sign in (+1, -1):
            for alpha in range(3 * self.natom):
                g = np.copy(self.geom)
                g[alpha // 3][alpha % 3] += sign * nuc_pert_strength
                q = copy.deepcopy(self.parameters)
                q['geom'] = self.molecule.with_coords(g).geometry_string()
                plist.append(q)
        for sign in (+1, -1):
            for beta in range(3):
                q = copy.deepcopy(self.parameters)
                q['F_mag'][beta] += sign * mag_pert_strength
                plist.append(q)
        res = phase_corrected_energies(plist, self.unperturbed_basis, self.unperturbed_C)
        nn = 3 * self.natom
        parts = [res[:nn], res[nn:2 * nn], res[2 * nn:2 * nn + 3], res[2 * nn + 3:]]
        out = []
        for part in parts:
            out.append(([r[2] for r in part], [r[3] for r in part], [r[1] for r in part]))
        (npC, npB, npT), (nnC, nnB, nnT), (mpC, mpB, mpT), (mnC, mnB, mnT) = out
        return npC, nnC, npB, nnB, npT, nnT, mpC, mnC, mpB, mnB, mpT, mnT
