"""Finite-difference driver with apyib's interface (fin_diff.py:11-372).

`finite_difference(parameters, unperturbed_basis, unperturbed_C)
.compute_AAT(nuc_pert_strength, mag_pert_strength)` returns the reference's
12-tuple (fin_diff.py:267-372): +dR then -dR on every Cartesian coordinate
(alpha = 3 * atom + xyz, bohr), then +dH then -dH on every component of the
magnetic field, each a phase_corrected_energy solve on the GPU.
Other finite-difference properties of the reference (Hessian, APT, nuclear
and field gradients, phase-space AATs) are not part of this path.
"""
import numpy as np

from .energy import phase_corrected_energy
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
        """fin_diff.py:267-372."""
        geom0 = self.parameters['geom']
        self.parameters.setdefault('F_mag', [0.0, 0.0, 0.0])
        out = {}
        for sign, key in ((+1, 'nuc_pos'), (-1, 'nuc_neg')):
            C_l, B_l, T_l = [], [], []
            for alpha in range(3 * self.natom):
                g = np.copy(self.geom)
                g[alpha // 3][alpha % 3] += sign * nuc_pert_strength
                self.parameters['geom'] = self.molecule.with_coords(g).geometry_string()
                C, T, basis = self._solve()
                C_l.append(C)
                T_l.append(T)
                B_l.append(basis)
                self.parameters['geom'] = geom0
            out[key] = (C_l, B_l, T_l)
        for sign, key in ((+1, 'mag_pos'), (-1, 'mag_neg')):
            C_l, B_l, T_l = [], [], []
            for beta in range(3):
                self.parameters['F_mag'][beta] += sign * mag_pert_strength
                C, T, basis = self._solve()
                C_l.append(C)
                T_l.append(T)
                B_l.append(basis)
                self.parameters['F_mag'][beta] -= sign * mag_pert_strength
            out[key] = (C_l, B_l, T_l)
        npC, npB, npT = out['nuc_pos']
        nnC, nnB, nnT = out['nuc_neg']
        mpC, mpB, mpT = out['mag_pos']
        mnC, mnB, mnT = out['mag_neg']
        return npC, nnC, npB, nnB, npT, nnT, mpC, mnC, mpB, mnB, mpT, mnT
