"""Analytic atomic axial tensors with apyib's interface (analytic_aats.py).

`analytic_derivative(parameters).compute_RHF_AATs(orbitals='non-canonical')`
as in the reference (analytic_aats.py:14-212): the Hartree-Fock AAT from
coupled-perturbed Hartree-Fock coefficients for every nuclear displacement
and magnetic-field component, with no displaced wavefunctions.

Skeleton first-derivative integrals (the reference's psi4 mo_oei_deriv1,
mo_tei_deriv1 and mo_overlap_half_deriv1) come from the package's host
integral engine (gaussian.build_deriv1); their AO -> MO transforms, the MO
two-electron integrals and the one-electron transforms run on the GPU
(apy_eri_mo, complex GEMMs).  The (nv*no)^2 CPHF matrices are inverted on
the host, as the reference does.
"""
import numpy as np
import torch

from . import gaussian
from . import transforms as tr
from .hamiltonian import ANGULAR_MOMENTUM_SIGN, Hamiltonian
from .hf_wfn import hf_wfn
from .utils import device, get_slices


def _c(expr, *ops):
    return np.einsum(expr, *ops, optimize=True)


class analytic_derivative(object):
    """analytic_aats.py:14-27."""

    def __init__(self, parameters):
        self.parameters = parameters
        self.H = Hamiltonian(self.parameters)
        self.wfn = hf_wfn(self.H)
        E_SCF, self.C = self.wfn.solve_SCF(self.parameters)
        self.dev = device()

    # ------------------------------------------------------------ building blocks
    def _mo1(self, M, Cd):
        return tr.sandwich(Cd, tr.to_dev(M, self.dev), Cd).cpu().numpy()

    def _cphf_setup(self, C, Cd):
        nbf, no = self.wfn.nbf, self.wfn.ndocc
        nv = nbf - no
        o, v = slice(0, no), slice(no, nbf)
        h = self._mo1(self.H.T + self.H.V, Cd)
        ERI = tr.eri_mo(self.H.device_eri(self.dev), Cd).cpu().numpy().swapaxes(1, 2)   # <pq|rs>
        F = h + _c('piqi->pq', 2 * ERI[:, o, :, o] - ERI.swapaxes(2, 3)[:, o, :, o])
        X = 2 * ERI - ERI.swapaxes(2, 3)
        A = (X + X.swapaxes(1, 3)).swapaxes(1, 2)
        Amag = (-X + X.swapaxes(1, 3)).swapaxes(1, 2)
        diag = _c('ab,ij,aibj->aibj', np.eye(nv), np.eye(no),
                  F[v, v].reshape(nv, 1, nv, 1) - F[o, o].reshape(1, no, 1, no))
        G = np.linalg.inv((diag + A[v, o, v, o]).reshape(nv * no, nv * no))
        Gm = np.linalg.inv((diag + Amag[v, o, v, o]).reshape(nv * no, nv * no))
        return F, A, G, Gm

    def _nuclear_U(self, C, Cd, F, A, G):
        """U (vo and ov blocks) and the MO half-derivative overlaps per
        (atom, axis), analytic_aats.py:84-141."""
        nbf, no = self.wfn.nbf, self.wfn.ndocc
        nv = nbf - no
        o, v = slice(0, no), slice(no, nbf)
        charges = list(zip(self.H.molecule.Z, self.H.molecule.coords))
        deriv = gaussian.build_deriv1(self.H.basis_set.functions, self.H.basis_set.atom_of, charges)
        U_R, half = [], []
        for N1 in range(len(charges)):
            for a in range(3):
                T1, V1, S1 = (self._mo1(deriv[k][N1][a], Cd) for k in ('T', 'V', 'S'))
                E1 = tr.eri_mo(tr.to_dev(deriv['ERI'][N1][a], self.dev), Cd).cpu().numpy().swapaxes(1, 2)
                F1 = T1 + V1 + _c('piqi->pq', 2 * E1[:, o, :, o] - E1.swapaxes(2, 3)[:, o, :, o])
                B = (-F1[v, o] + _c('ai,ii->ai', S1[v, o], F[o, o])
                     + 0.5 * _c('mn,amin->ai', S1[o, o], A.swapaxes(1, 2)[v, o, o, o]))
                U = np.zeros((nbf, nbf), dtype=np.complex128)
                U[v, o] += (G @ B.reshape(nv * no)).reshape(nv, no)
                U[o, v] -= U[v, o].T + S1[o, v]
                U_R.append(U)
                half.append(self._mo1(deriv['half_S'][N1][a], Cd))
        return U_R, half

    def _magnetic_U(self, Cd, Gm):
        """U for the three magnetic-field components (analytic_aats.py:143-186),
        operator -1/2 L."""
        nbf, no = self.wfn.nbf, self.wfn.ndocc
        nv = nbf - no
        o, v = slice(0, no), slice(no, nbf)
        L = ANGULAR_MOMENTUM_SIGN * gaussian.build_angular_momentum(self.H.basis_set.functions)
        out = []
        for a in range(3):
            mu = self._mo1(-0.5 * L[a], Cd)
            U = np.zeros((nbf, nbf), dtype=np.complex128)
            U[v, o] += (Gm @ mu[v, o].reshape(nv * no)).reshape(nv, no)
            U[o, v] += U[v, o].T
            out.append(U)
        return out

    # ------------------------------------------------------------------- AATs
    def compute_RHF_AATs(self, orbitals='non-canonical'):
        """analytic_aats.py:30-212.  Returns the [3N, 3] Hartree-Fock AAT.
        Only the virtual-occupied blocks of the CPHF coefficients enter, so
        `orbitals` (which fixes the occupied-occupied and virtual-virtual
        blocks) does not change the result."""
        if orbitals not in ('canonical', 'non-canonical'):
            raise ValueError(f"orbitals {orbitals!r}")
        C = np.asarray(self.C)
        Cd = tr.to_dev(C.astype(np.complex128), self.dev)
        F, A, G, Gm = self._cphf_setup(C, Cd)
        U_R, half = self._nuclear_U(C, Cd, F, A, G)
        U_H = self._magnetic_U(Cd, Gm)
        C_list, _ = get_slices(self.parameters, self.wfn)
        v_ = C_list[2]
        o = slice(0, self.wfn.ndocc)
        natom = self.H.molecule.natom()
        AAT = np.zeros((natom * 3, 3))
        for la in range(3 * natom):
            for b in range(3):
                AAT[la][b] += np.real(2 * _c('em,em', U_H[b][v_, o], U_R[la][v_, o] + half[la][o, v_].T))
        return AAT
