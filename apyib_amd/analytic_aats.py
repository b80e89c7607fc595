"""Analytic atomic axial tensors with apyib's interface (analytic_aats.py).

`analytic_derivative(parameters)` with `compute_RHF_AATs(orbitals)` and
`compute_MP2_AATs(normalization, orbitals)` as in the reference
(analytic_aats.py:14-550): the AAT from coupled-perturbed Hartree-Fock
coefficients for every nuclear displacement and magnetic-field component
and, for MP2, the first-order amplitudes, with no displaced wavefunctions.

Skeleton first-derivative integrals (the reference's psi4 mo_oei_deriv1,
mo_tei_deriv1 and mo_overlap_half_deriv1) come from the package's host
integral engine (gaussian.build_deriv1).  On the GPU: the AO -> MO
transforms of the two-electron integrals and of every derivative integral
set (apy_eri_mo), the one-electron transforms, the MP2 amplitudes, and the
four U-contractions of the first-order integrals per perturbation (n^5
each, complex GEMMs).  The (nv*no)^2 CPHF matrices are inverted and the
O(o^2 v^2 n) amplitude-level contractions done on the host, as the
reference does.
"""
import numpy as np
import torch

from . import gaussian
from . import transforms as tr
from .hamiltonian import ANGULAR_MOMENTUM_SIGN, Hamiltonian
from .hf_wfn import hf_wfn
from .kernels import zgemm_raw
from .mp2_wfn import mp2_wfn
from .utils import device, get_slices
from . import _lib

CDT = torch.complex128


def _c(expr, *ops):
    return np.einsum(expr, *ops, optimize=True)


def _dERI(U, E, signs):
    """sum_t of U contracted into each index of <pq|rs> with the given signs
    (analytic_aats.py:344-347, :479-483), on the device:
        s0 U[t,p] E[t,q,r,s] + s1 U[t,q] E[p,t,r,s] + s2 U[t,r] E[p,q,t,s] + s3 U[t,s] E[p,q,r,t]."""
    n = U.shape[0]
    out = torch.empty_like(E)
    n2, n3 = n * n, n * n * n
    KM, KN, MK = _lib.APY_KM, _lib.APY_KN, _lib.APY_MK
    zgemm_raw(m=n, n=n3, k=n, a=U, a_layout=KM, lda=n, b=E, b_layout=KN, ldb=n3, c=out, ldc=n3,
              alpha=signs[0], beta=0.0)
    zgemm_raw(m=n, n=n2, k=n, a=U, a_layout=KM, lda=n, b=E, b_layout=KN, ldb=n2, c=out, ldc=n2,
              alpha=signs[1], beta=1.0, batch=n, stride_a=0, stride_b=n3, stride_c=n3)
    zgemm_raw(m=n, n=n, k=n, a=U, a_layout=KM, lda=n, b=E, b_layout=KN, ldb=n, c=out, ldc=n,
              alpha=signs[2], beta=1.0, batch=n2, stride_a=0, stride_b=n2, stride_c=n2)
    zgemm_raw(m=n3, n=n, k=n, a=E, a_layout=MK, lda=n, b=U, b_layout=KN, ldb=n, c=out, ldc=n,
              alpha=signs[3], beta=1.0)
    return out


class analytic_derivative(object):
    """analytic_aats.py:14-27."""

    def __init__(self, parameters):
        self.parameters = parameters
        self.H = Hamiltonian(self.parameters)
        self.wfn = hf_wfn(self.H)
        E_SCF, self.C = self.wfn.solve_SCF(self.parameters)
        self.dev = device()
        self._ready = False

    # ------------------------------------------------------------ building blocks
    def _setup(self):
        if self._ready:
            return
        H, wfn = self.H, self.wfn
        self.no = no = wfn.ndocc
        self.nbf = nbf = wfn.nbf
        self.nv = nv = nbf - no
        self.o, self.v = slice(0, no), slice(no, nbf)
        o, v = self.o, self.v
        C = np.real_if_close(np.asarray(self.C))
        self.Cr = C
        self.Cd = tr.to_dev(C.astype(np.complex128), self.dev)
        self.eps = np.real(np.asarray(wfn.eps))
        self.freeze_core = bool(self.parameters.get('freeze_core', False))
        self.C_list, _ = get_slices(self.parameters, wfn)
        self.Ed = tr.eri_mo(H.device_eri(self.dev), self.Cd).permute(0, 2, 1, 3).contiguous()   # <pq|rs>
        ERI = np.real_if_close(self.Ed.cpu().numpy())
        h = self._mo1(H.T + H.V)
        self.F = F = h + _c('piqi->pq', 2 * ERI[:, o, :, o] - ERI.swapaxes(2, 3)[:, o, :, o])
        X = 2 * ERI - ERI.swapaxes(2, 3)
        self.A = (X + X.swapaxes(1, 3)).swapaxes(1, 2)
        self.Amag = (-X + X.swapaxes(1, 3)).swapaxes(1, 2)
        diag = _c('ab,ij,aibj->aibj', np.eye(nv), np.eye(no),
                  F[v, v].reshape(nv, 1, nv, 1) - F[o, o].reshape(1, no, 1, no))
        self.G = np.linalg.inv((diag + self.A[v, o, v, o]).reshape(nv * no, nv * no))
        self.Gm = np.linalg.inv((diag + self.Amag[v, o, v, o]).reshape(nv * no, nv * no))
        charges = list(zip(H.molecule.Z, H.molecule.coords))
        self.natom = len(charges)
        self.deriv = gaussian.build_deriv1(H.basis_set.functions, H.basis_set.atom_of, charges)
        self.L = ANGULAR_MOMENTUM_SIGN * gaussian.build_angular_momentum(H.basis_set.functions)
        self._ready = True

    def _mo1(self, M):
        return np.real_if_close(tr.sandwich(self.Cd, tr.to_dev(M, self.dev), self.Cd).cpu().numpy())

    def _dep_denoms(self):
        o, v, eps = self.o, self.v, self.eps
        return (eps[o] - eps[o].reshape(-1, 1)) + np.eye(self.no), (eps[v] - eps[v].reshape(-1, 1)) + np.eye(self.nv)

    def _magnetic(self, b, orbitals):
        """(U_h, h_core) for field component b (analytic_aats.py:143-186, 278-313)."""
        o, v, no, nv, nbf = self.o, self.v, self.no, self.nv, self.nbf
        h_core = self._mo1(-0.5 * self.L[b])
        U = np.zeros((nbf, nbf))
        U[v, o] += np.real_if_close(self.Gm @ h_core[v, o].reshape(nv * no)).reshape(nv, no)
        U[o, v] += U[v, o].T
        Am = self.Amag.swapaxes(1, 2)
        if self.freeze_core or orbitals == 'canonical':
            Do, Dv = self._dep_denoms()
            U[o, o] += (-h_core[o, o] + _c('em,iejm->ij', U[v, o], Am[o, v, o, o])) / Do
            U[v, v] += (-h_core[v, v] + _c('em,aebm->ab', U[v, o], Am[v, v, v, o])) / Dv
            for j in range(nbf):
                U[j, j] = 0
        if orbitals == 'non-canonical':
            f_, o_, v_, _ = self.C_list
            U[f_, f_] = 0
            U[o_, o_] = 0
            U[v_, v_] = 0
        return U, h_core

    def _nuclear(self, N1, a, orbitals):
        """(U_R, F_core, S_core, ERI_core Dirac (device), half_S) for atom N1,
        axis a (analytic_aats.py:84-141, 412-470)."""
        o, v, no, nv, nbf = self.o, self.v, self.no, self.nv, self.nbf
        F, A, d = self.F, self.A.swapaxes(1, 2), self.deriv
        T1, V1, S1 = (self._mo1(d[k][N1][a]) for k in ('T', 'V', 'S'))
        E1d = tr.eri_mo(tr.to_dev(d['ERI'][N1][a], self.dev), self.Cd).permute(0, 2, 1, 3).contiguous()
        E1 = np.real_if_close(E1d.cpu().numpy())
        F1 = T1 + V1 + np.real_if_close(_c('piqi->pq', 2 * E1[:, o, :, o] - E1.swapaxes(2, 3)[:, o, :, o]))
        B = -F1[v, o] + _c('ai,ii->ai', S1[v, o], F[o, o]) + 0.5 * _c('mn,amin->ai', S1[o, o], A[v, o, o, o])
        U = np.zeros((nbf, nbf))
        U[v, o] += np.real_if_close(self.G @ B.reshape(nv * no)).reshape(nv, no)
        U[o, v] -= U[v, o].T + S1[o, v]
        if self.freeze_core or orbitals == 'canonical':
            Do, Dv = self._dep_denoms()
            Bo = (F1[o, o] - _c('ij,jj->ij', S1[o, o], F[o, o]) + _c('em,iejm->ij', U[v, o], A[o, v, o, o])
                  - 0.5 * _c('mn,imjn->ij', S1[o, o], A[o, o, o, o]))
            U[o, o] += np.real_if_close(Bo / Do)
            Bv = (F1[v, v] - _c('ab,bb->ab', S1[v, v], F[v, v]) + _c('em,aebm->ab', U[v, o], A[v, v, v, o])
                  - 0.5 * _c('mn,ambn->ab', S1[o, o], A[v, o, v, o]))
            U[v, v] += np.real_if_close(Bv / Dv)
            for j in range(nbf):
                U[j, j] = -0.5 * S1[j, j]
        if orbitals == 'non-canonical':
            f_, o_, v_, _ = self.C_list
            U[f_, f_] = -0.5 * S1[f_, f_]
            U[o_, o_] = -0.5 * S1[o_, o_]
            U[v_, v_] = -0.5 * S1[v_, v_]
        half = self._mo1(d['half_S'][N1][a])
        return U, F1, S1, E1d, half

    # ------------------------------------------------------------------- AATs
    def compute_RHF_AATs(self, orbitals='non-canonical'):
        """analytic_aats.py:30-212.  Returns the [3N, 3] Hartree-Fock AAT
        (only the virtual-occupied blocks of U enter, so `orbitals` does not
        change the result)."""
        if orbitals not in ('canonical', 'non-canonical'):
            raise ValueError(f"orbitals {orbitals!r}")
        self._setup()
        o, v_ = self.o, self.C_list[2]
        U_H = [self._magnetic(b, orbitals)[0] for b in range(3)]
        AAT = np.zeros((self.natom * 3, 3))
        for N1 in range(self.natom):
            for a in range(3):
                U_R, _, _, _, half = self._nuclear(N1, a, orbitals)
                for b in range(3):
                    AAT[3 * N1 + a][b] += np.real(2 * _c('em,em', U_H[b][v_, o], U_R[v_, o] + half[o, v_].T))
        return AAT

    def compute_MP2_AATs(self, normalization='full', orbitals='non-canonical'):
        """analytic_aats.py:215-550.  Returns the [3N, 3] MP2 AAT: HF term,
        the four doubles terms and (full normalisation) the normalisation
        term, with the reference's canonical / non-canonical term sets."""
        if orbitals not in ('canonical', 'non-canonical'):
            raise ValueError(f"orbitals {orbitals!r}")
        self._setup()
        E_MP2, t2 = mp2_wfn(self.parameters, self.wfn).solve_MP2()
        t2 = np.real_if_close(np.asarray(t2))
        o, v, eps = self.o, self.v, self.eps
        f_, o_, v_, t_ = self.C_list
        eo, ev = eps[o_], eps[v_]
        D = eo.reshape(-1, 1, 1, 1) + eo.reshape(-1, 1, 1) - ev.reshape(-1, 1) - ev
        tt = 2 * t2 - t2.swapaxes(2, 3)
        N = 1.0 if normalization == 'intermediate' else 1 / np.sqrt(1 + _c('ijab,ijab', t2, tt))
        Am, A = self.Amag.swapaxes(1, 2), self.A.swapaxes(1, 2)
        nbf = self.nbf
        U_H, dT2_dH = [], []
        for b in range(3):
            U, h_core = self._magnetic(b, orbitals)
            df = np.zeros((nbf, nbf))
            df[o, o] += (-h_core[o, o] + U[o, o] * eps[o].reshape(-1, 1) - U[o, o].T * eps[o]
                         + _c('em,iejm->ij', U[v, o], Am[o, v, o, o]))
            df[v, v] += (-h_core[v, v] + U[v, v] * eps[v].reshape(-1, 1) - U[v, v].T * eps[v]
                         + _c('em,aebm->ab', U[v, o], Am[v, v, v, o]))
            dE = np.real_if_close(_dERI(tr.to_dev(U.astype(np.complex128), self.dev), self.Ed,
                                        (-1.0, -1.0, 1.0, 1.0)).cpu().numpy())
            dt = dE.swapaxes(0, 2).swapaxes(1, 3)[o_, o_, v_, v_].astype(np.result_type(dE, t2))
            dt += _c('ac,ijcb->ijab', df[v_, v_], t2) + _c('bc,ijac->ijab', df[v_, v_], t2)
            dt -= _c('ki,kjab->ijab', df[o_, o_], t2) + _c('kj,ikab->ijab', df[o_, o_], t2)
            U_H.append(U)
            dT2_dH.append(dt / D)
        AAT = np.zeros((self.natom * 3, 3))
        for N1 in range(self.natom):
            for a in range(3):
                U_R, F1, S1, E1d, half = self._nuclear(N1, a, orbitals)
                df = np.zeros((nbf, nbf))
                df[o, o] += (F1[o, o] + U_R[o, o] * eps[o].reshape(-1, 1) + U_R[o, o].T * eps[o]
                             + _c('em,iejm->ij', U_R[v, o], A[o, v, o, o])
                             - 0.5 * _c('mn,imjn->ij', S1[o, o], A[o, o, o, o]))
                df[v, v] += (F1[v, v] + U_R[v, v] * eps[v].reshape(-1, 1) + U_R[v, v].T * eps[v]
                             + _c('em,aebm->ab', U_R[v, o], A[v, v, v, o])
                             - 0.5 * _c('mn,ambn->ab', S1[o, o], A[v, o, v, o]))
                dE = np.real_if_close((E1d + _dERI(tr.to_dev(U_R.astype(np.complex128), self.dev), self.Ed,
                                                   (1.0, 1.0, 1.0, 1.0))).cpu().numpy())
                dt = dE[o_, o_, v_, v_].astype(np.result_type(dE, t2))
                dt -= _c('kjab,ik->ijab', t2, df[o_, o_]) + _c('ikab,kj->ijab', t2, df[o_, o_])
                dt += _c('ijcb,ac->ijab', t2, df[v_, v_]) + _c('ijac,cb->ijab', t2, df[v_, v_])
                dt = dt / D
                dtt = 2 * dt - dt.swapaxes(2, 3)
                N_R = -(1 / np.sqrt((1 + _c('ijab,ijab', np.conj(t2), tt)) ** 3))
                N_R *= 0.5 * (_c('ijab,ijab', np.conj(dt), tt) + _c('ijab,ijab', dt, np.conj(tt)))
                R = U_R + half.T
                for b in range(3):
                    UH, dH = U_H[b], dT2_dH[b]
                    dHH = 2 * dH - dH.swapaxes(2, 3)
                    val = N ** 2 * 2 * _c('em,em', UH[v_, o], R[v_, o])
                    val += N ** 2 * _c('ijab,ijab', dtt, dH)
                    if orbitals == 'canonical':
                        val -= N ** 2 * 2.0 * _c('ijab,kjab,ki', dtt, t2, UH[o_, o_])
                        val += N ** 2 * 2.0 * _c('ijab,ijcb,ac', dtt, t2, UH[v_, v_])
                    val -= N ** 2 * 2.0 * _c('ijab,kjab,ki', dHH, t2, R[o_, o_])
                    val += N ** 2 * 2.0 * _c('ijab,ijcb,ac', dHH, t2, R[v_, v_])
                    val += N ** 2 * 2.0 * _c('ijab,kjab,km,im', t2, tt, UH[o_, o], R[o_, o])
                    if orbitals == 'canonical':
                        val += N ** 2 * 2.0 * _c('ijab,ijcb,ec,ea', t2, tt, UH[v_, v_], R[v_, v_])
                    val += N ** 2 * 2.0 * _c('ijab,ijab,em,em', t2, tt, UH[v_, o], R[v_, o])
                    val -= N ** 2 * 2.0 * _c('ijab,imab,ej,em', t2, tt, UH[v_, o_], R[v_, o_])
                    val -= N ** 2 * 2.0 * _c('ijab,ijae,bm,em', t2, tt, UH[v_, o], R[v_, o])
                    if normalization == 'full':
                        if orbitals == 'canonical':
                            val -= N * N_R * 2.0 * _c('ijab,kjab,ki', tt, t2, UH[o_, o_])
                            val += N * N_R * 2.0 * _c('ijab,ijcb,ac', tt, t2, UH[v_, v_])
                        val += N * N_R * _c('ijab,ijab', tt, dH)
                    AAT[3 * N1 + a][b] += np.real(val)
        return AAT
