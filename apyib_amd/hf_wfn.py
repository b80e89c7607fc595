"""Complex restricted Hartree-Fock with apyib's interface (hf_wfn.py:9-170).

The upstream producer of C and eps for the correlated path.  The O(nbf^4)
Fock build runs on the GPU (apy_fock_build); the nbf x nbf algebra
(orthogonalisation, eigh, DIIS on Fock matrices) is the reference's small
host-side linear algebra.
"""
import numpy as np
import scipy.linalg as la
import torch

from . import transforms as tr
from .utils import device, solve_general_DIIS


class hf_wfn(object):
    def __init__(self, H, charge=0):
        self.H = H
        self.nelec = self.nelectron(charge)
        self.ndocc = self.nelec // 2
        self.nbf = H.basis_set.nbf()
        self.C = np.zeros((self.nbf, self.nbf))
        self.eps = np.zeros((self.nbf))
        self.E_SCF = 0

    def nelectron(self, charge):
        nelec = -charge
        for atom in range(self.H.molecule.natom()):
            nelec += self.H.molecule.true_atomic_number(atom)
        return nelec

    def solve_SCF(self, parameters, print_level=0):
        """hf_wfn.py:36-170.  Returns (E_SCF, C)."""
        H = self.H
        dev = device()
        H_core = H.T + H.V
        X = np.linalg.inv(la.sqrtm(H.S))
        e, C_p = np.linalg.eigh(X @ H_core @ X)
        C = X @ C_p
        D = 2 * C[:, :self.ndocc] @ np.conjugate(C[:, :self.ndocc]).T
        E_SCF = np.einsum('mn,mn->', D, H_core)
        eri = H.device_eri(dev)
        hcore_d = tr.to_dev(H_core, dev)
        if print_level > 0:
            print("\n Iter      E_elec(real)       E_elec(imaginary)        E(tot)           Delta_E(real)       "
                  "Delta_E(imaginary)      RMS_D(real)      RMS_D(imaginary)")
            print(" %02d %20.12f %20.12f %20.12f" % (0, E_SCF.real, E_SCF.imag, E_SCF.real + H.E_nuc))
        F_iter = e_iter = None
        for i in range(1, parameters['max_iterations'] + 1):
            E_old, D_old = E_SCF, D
            F = tr.fock(hcore_d, eri, tr.to_dev(D, dev), 0.5).cpu().numpy()           # hf_wfn.py:81
            if not np.iscomplexobj(H_core) and not np.iscomplexobj(D):
                F = F.real
            if parameters['DIIS']:
                SDF = H.S @ D @ F
                res_vec = np.reshape(X @ (SDF - np.conjugate(SDF.T)) @ X, (-1))
                F_vec = np.reshape(F, (-1))
                if i == 1:
                    F_iter = np.atleast_2d(F_vec).T
                    e_iter = np.atleast_2d(res_vec).T
                F_vec, e_iter, F_iter = solve_general_DIIS(parameters, res_vec, F_vec, e_iter, F_iter, i)
                F = np.reshape(F_vec, (self.nbf, self.nbf))
            self.eps, C_p = np.linalg.eigh(X @ F @ X)
            C = self.C = X @ C_p
            D = 2 * C[:, :self.ndocc] @ np.conjugate(C[:, :self.ndocc]).T
            E_SCF = 0.5 * np.einsum('nm,mn->', D, H_core + F)
            delta_E = E_SCF - E_old
            rms_D = np.sqrt(np.sum((D_old - D) ** 2))
            if print_level > 0:
                print(" %02d %20.12f %20.12f %20.12f %20.12f %20.12f %20.12f %20.12f" % (
                    i, E_SCF.real, E_SCF.imag, (E_SCF + H.E_nuc).real, delta_E.real, delta_E.imag,
                    rms_D.real, rms_D.imag))
            if i > 1 and abs(delta_E) < parameters['e_convergence'] and rms_D < parameters['d_convergence']:
                break
            if i == parameters['max_iterations']:
                if abs(delta_E) > parameters['e_convergence'] or rms_D > parameters['d_convergence']:
                    if print_level > 0:
                        print("Not converged.")
        self.E_SCF = E_SCF
        return E_SCF, self.C
