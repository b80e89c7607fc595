"""Analytic atomic axial tensors with apyib's interface (analytic_aats.py).

`analytic_derivative(parameters)` with `compute_RHF_AATs(orbitals)`,
`compute_MP2_AATs`, `compute_CISD_AATs` and `compute_CID_AATs`
(normalization, orbitals) as in the reference (analytic_aats.py:14-1927):
the AAT from coupled-perturbed Hartree-Fock coefficients for every nuclear
displacement and magnetic-field component and, for the correlated methods,
the first-order amplitudes, with no displaced wavefunctions.

Skeleton first-derivative integrals (the reference's psi4 mo_oei_deriv1,
mo_tei_deriv1 and mo_overlap_half_deriv1) on the GPU: the one-electron
ones from apy_oei_gto over the derivative functions
(integrals.deriv1_one_electron), the two-electron ones from apy_eri_gto over
them (integrals.EriDerivatives, apy_eri_deriv_assemble per atom and axis).  On the GPU: the AO -> MO
transforms of the two-electron integrals and of every derivative integral
set (apy_eri_mo), the one-electron transforms, the MP2 amplitudes, the
four U-contractions of the first-order integrals per perturbation (n^5
each, complex GEMMs), the CISD/CID amplitudes and, per perturbation, the
first-order CISD/CID amplitude equations of all 3N+3 perturbations as one
set (spatial_ci.ResponseSet: the batched spin-adapted step with its
linear-response term, Jacobi + DIIS per perturbation).  The
(nv*no)^2 CPHF matrices are inverted and the O(o^3 v^2) amplitude-level
AAT contractions done on the host, as the reference does.
"""
import numpy as np
import torch

from . import transforms as tr
from .hamiltonian import ANGULAR_MOMENTUM_SIGN, build_hamiltonians
from .integrals import EriDerivatives, deriv1_one_electron, hamiltonian_integrals
from .hf_wfn import hf_wfn
from .ci_wfn import ci_wfn
from .kernels import zgemm_raw
from .mp2_wfn import mp2_wfn
from . import mp2_grad
from .spatial_ci import ResponseSet, SPIntegrals
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
        self.dev = device()
        self.H = build_hamiltonians([self.parameters], self.dev)[0]
        self.wfn = hf_wfn(self.H)
        E_SCF, self.C = self.wfn.solve_SCF(self.parameters)
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
        self.natom = H.molecule.natom()
        if H.basis_set.transform is not None:
            raise NotImplementedError(f"analytic derivatives for the {H.basis_set.name} basis (spherical d or split "
                                      "contractions) are not built")
        self.deriv = deriv1_one_electron(H.basis_set, H.molecule.coords, H.molecule.Z, self.dev)
        self.dERI = EriDerivatives(H.basis_set, H.molecule.coords, self.dev)
        L = hamiltonian_integrals(H.basis_set, np.asarray(H.molecule.coords)[None], H.molecule.Z, self.dev,
                                  which=_lib.APY_OEI_ANGMOM)[0].cpu().numpy()
        self.L = ANGULAR_MOMENTUM_SIGN * 0.5 * (L - L.swapaxes(1, 2))
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
        E1d = tr.eri_mo(self.dERI(N1, a), self.Cd).permute(0, 2, 1, 3).contiguous()
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

    def _df_magnetic(self, U, h_core):
        """Field derivative of the Fock matrix, oo and vv blocks
        (analytic_aats.py:300-308, :720-728)."""
        o, v, eps = self.o, self.v, self.eps
        Am = self.Amag.swapaxes(1, 2)
        df = np.zeros((self.nbf, self.nbf))
        df[o, o] += (-h_core[o, o] + U[o, o] * eps[o].reshape(-1, 1) - U[o, o].T * eps[o]
                     + _c('em,iejm->ij', U[v, o], Am[o, v, o, o]))
        df[v, v] += (-h_core[v, v] + U[v, v] * eps[v].reshape(-1, 1) - U[v, v].T * eps[v]
                     + _c('em,aebm->ab', U[v, o], Am[v, v, v, o]))
        return df

    def _df_nuclear(self, U_R, F1, S1):
        """Nuclear derivative of the Fock matrix, oo and vv blocks
        (analytic_aats.py:457-469, :964-974)."""
        o, v, eps = self.o, self.v, self.eps
        A = self.A.swapaxes(1, 2)
        df = np.zeros((self.nbf, self.nbf))
        df[o, o] += (F1[o, o] + U_R[o, o] * eps[o].reshape(-1, 1) + U_R[o, o].T * eps[o]
                     + _c('em,iejm->ij', U_R[v, o], A[o, v, o, o])
                     - 0.5 * _c('mn,imjn->ij', S1[o, o], A[o, o, o, o]))
        df[v, v] += (F1[v, v] + U_R[v, v] * eps[v].reshape(-1, 1) + U_R[v, v].T * eps[v]
                     + _c('em,aebm->ab', U_R[v, o], A[v, v, v, o])
                     - 0.5 * _c('mn,ambn->ab', S1[o, o], A[v, o, v, o]))
        return df

    def _doubles_part(self, orbitals, normalization, N, N_R, t2, dt2_R, dt2_H, UH, R):
        """Doubles/doubles and doubles normalisation terms of one AAT element
        (MP2 analytic_aats.py:491-528, CID :1865-1907, CISD :1226-1249,
        :1290-1308); R[p,q] = U_R[p,q] + half_S[q,p]."""
        o = self.o
        _, o_, v_, _ = self.C_list
        tt = 2 * t2 - t2.swapaxes(2, 3)
        dtt = 2 * dt2_R - dt2_R.swapaxes(2, 3)
        dHH = 2 * dt2_H - dt2_H.swapaxes(2, 3)
        can = orbitals == 'canonical'
        val = N ** 2 * _c('ijab,ijab', dtt, dt2_H)
        if can:
            val -= N ** 2 * 2.0 * _c('ijab,kjab,ki', dtt, t2, UH[o_, o_])
            val += N ** 2 * 2.0 * _c('ijab,ijcb,ac', dtt, t2, UH[v_, v_])
        val -= N ** 2 * 2.0 * _c('ijab,kjab,ki', dHH, t2, R[o_, o_])
        val += N ** 2 * 2.0 * _c('ijab,ijcb,ac', dHH, t2, R[v_, v_])
        val += N ** 2 * 2.0 * _c('ijab,kjab,km,im', t2, tt, UH[o_, o], R[o_, o])
        if can:
            val += N ** 2 * 2.0 * _c('ijab,ijcb,ec,ea', t2, tt, UH[v_, v_], R[v_, v_])
        val += N ** 2 * 2.0 * _c('ijab,ijab,em,em', t2, tt, UH[v_, o], R[v_, o])
        val -= N ** 2 * 2.0 * _c('ijab,imab,ej,em', t2, tt, UH[v_, o_], R[v_, o_])
        val -= N ** 2 * 2.0 * _c('ijab,ijae,bm,em', t2, tt, UH[v_, o], R[v_, o])
        if normalization == 'full':
            if can:
                val -= N * N_R * 2.0 * _c('ijab,kjab,ki', tt, t2, UH[o_, o_])
                val += N * N_R * 2.0 * _c('ijab,ijcb,ac', tt, t2, UH[v_, v_])
            val += N * N_R * _c('ijab,ijab', tt, dt2_H)
        return val

    def _singles_part(self, orbitals, normalization, N, N_R, t1, t2, dt1_R, dt2_R, dt1_H, dt2_H, UH, R):
        """Singles/reference, singles/singles, singles/doubles terms and the
        singles normalisation terms of one CISD AAT element
        (analytic_aats.py:1163-1224 and :1271-1288 canonical, :1252-1310
        non-canonical), as (coefficient, subscripts, operands) rows."""
        o = self.o
        _, o_, v_, _ = self.C_list
        tt = 2 * t2 - t2.swapaxes(2, 3)
        dtt = 2 * dt2_R - dt2_R.swapaxes(2, 3)
        dHH = 2 * dt2_H - dt2_H.swapaxes(2, 3)
        Hvo, Hvo_, Hov_, Hvv, Hoo, Hoo_ = UH[v_, o], UH[v_, o_], UH[o_, v_], UH[v_, v_], UH[o_, o], UH[o_, o_]
        Rvo, Rvo_, Rov_, Rvv, Roo, Roo_ = R[v_, o], R[v_, o_], R[o_, v_], R[v_, v_], R[o_, o], R[o_, o_]
        if orbitals == 'canonical':
            rows = [
                (2, 'ia,ai', dt1_R, Hvo_), (2, 'ia,ei,ea', t1, Hvo_, Rvv), (-2, 'ia,am,im', t1, Hvo, Roo),
                (2, 'kc,ck', dt1_H, Rvo_), (2, 'kc,fc,fk', t1, Hvv, Rvo_), (-2, 'kc,kn,cn', t1, Hoo, Rvo),
                (2, 'ia,ia', dt1_R, dt1_H), (2, 'kc,cf,kf', dt1_R, Hvv, t1), (-2, 'kc,nk,nc', dt1_R, Hoo_, t1),
                (2, 'ia,ae,ie', dt1_H, Rvv, t1), (-2, 'ia,mi,ma', dt1_H, Roo_, t1),
                (2, 'kc,fc,fa,ka', t1, Hvv, Rvv, t1), (-2, 'kc,fc,ik,if', t1, Hvv, Roo_, t1),
                (-2, 'kc,kn,ca,na', t1, Hoo_, Rvv, t1), (2, 'kc,kn,in,ic', t1, Hoo, Roo, t1),
                (4, 'kc,kc,ia,ia', t1, Hov_, Rov_, t1), (4, 'kc,fn,fn,kc', t1, Hvo, Rvo, t1),
                (-2, 'kc,fn,fk,nc', t1, Hvo_, Rvo_, t1), (-2, 'kc,fn,cn,kf', t1, Hvo, Rvo, t1),
                (4, 'kc,fn,ck,nf', t1, Hvo_, Rvo_, t1),
                (2, 'ijab,bj,ia', dtt, Hvo_, t1), (2, 'kc,ia,ikac', dt1_H, Rov_, tt),
                (2, 'kc,fc,ia,ikaf', t1, Hvv, Rov_, tt), (-2, 'kc,kn,ia,inac', t1, Hoo_, Rov_, tt),
                (-2, 'kc,fn,ik,incf', t1, Hvo_, Roo_, tt), (-2, 'kc,fn,in,ikfc', t1, Hvo, Roo, tt),
                (2, 'kc,fn,ca,knaf', t1, Hvo_, Rvv, tt), (2, 'kc,fn,fa,knca', t1, Hvo_, Rvv, tt),
                (2, 'ia,kc,ikac', dt1_R, Hov_, tt), (2, 'klcd,dl,kc', dHH, Rvo_, t1),
                (2, 'ia,kc,ea,kice', t1, Hov_, Rvv, tt), (-2, 'ia,kc,im,kmca', t1, Hov_, Roo_, tt),
                (2, 'ia,ac,em,imce', t1, Hvv, Rvo_, tt), (2, 'ia,ec,em,imac', t1, Hvv, Rvo_, tt),
                (-2, 'ia,ki,em,kmae', t1, Hoo_, Rvo_, tt), (-2, 'ia,km,em,kiea', t1, Hoo, Rvo, tt),
            ]
            norm = [(2, 'ia,ai', t1, Hvo_), (2, 'kc,kc', t1, Hov_), (2, 'ia,ia', t1, dt1_H),
                    (2, 'kc,cf,kf', t1, Hvv, t1), (-2, 'kc,nk,nc', t1, Hoo_, t1),
                    (2, 'ijab,bj,ia', tt, Hvo_, t1), (2, 'ia,kc,ikac', t1, Hov_, tt)]
        else:
            rows = [
                (2, 'ia,ai', dt1_R, Hvo_), (2, 'ia,ei,ea', t1, Hvo_, Rvv), (-2, 'ia,am,im', t1, Hvo, Roo),
                (2, 'ia,ai', dt1_H, Rvo_), (-2, 'ia,im,am', t1, Hoo, Rvo),
                (2, 'ia,ia', dt1_R, dt1_H), (2, 'ia,ae,ie', dt1_H, Rvv, t1), (-2, 'ia,mi,ma', dt1_H, Roo_, t1),
                (2, 'ia,im,km,ka', t1, Hoo, Roo, t1), (4, 'ia,ia,kc,kc', t1, Hov_, Rov_, t1),
                (4, 'ia,em,em,ia', t1, Hvo, Rvo, t1), (-2, 'ia,em,ei,ma', t1, Hvo_, Rvo_, t1),
                (-2, 'ia,em,am,ie', t1, Hvo, Rvo, t1), (4, 'ia,em,ai,me', t1, Hvo_, Rvo_, t1),
                (2, 'ijab,bj,ia', dtt, Hvo_, t1), (2, 'ia,kc,ikac', dt1_H, Rov_, tt),
                (-2, 'ia,em,km,kiea', t1, Hvo, Roo, tt), (2, 'ia,em,ec,imac', t1, Hvo_, Rvv, tt),
                (2, 'ia,kc,ikac', dt1_R, Hov_, tt), (2, 'ijab,bj,ia', dHH, Rvo_, t1),
                (-2, 'ia,km,em,kiea', t1, Hoo, Rvo, tt),
            ]
            norm = [(4, 'ia,ai', t1, Hvo_), (2, 'ia,ia', t1, dt1_H), (4, 'ijab,bj,ia', tt, Hvo_, t1)]
        val = N ** 2 * sum(c * _c(e, *ops) for c, e, *ops in rows)
        if normalization == 'full':
            val += N * N_R * sum(c * _c(e, *ops) for c, e, *ops in norm)
        return val

    def _ci_aats(self, singles, normalization, orbitals, print_level):
        """Shared body of compute_CISD_AATs / compute_CID_AATs."""
        if orbitals not in ('canonical', 'non-canonical'):
            raise ValueError(f"orbitals {orbitals!r}")
        self._setup()
        o = self.o
        f_, o_, v_, t_ = self.C_list
        wfn_ci = ci_wfn(self.parameters, self.wfn)
        solver = wfn_ci._spatial_solver(singles)
        E, t1d, t2d, _ = solver.solve(self.parameters, print_level, e_offset=wfn_ci._e_offset())
        t_ref = solver.t.clone()
        t1 = np.real_if_close(t1d.cpu().numpy()) if singles else None
        t2 = np.real_if_close(t2d.cpu().numpy())
        O = solver.O
        tt = 2 * t2 - t2.swapaxes(2, 3)
        norm2 = 1 + _c('ijab,ijab', np.conj(t2), tt) + (2 * _c('ia,ia', np.conj(t1), t1) if singles else 0)
        N = 1 / np.sqrt(norm2)          # the full factor in both normalisations (analytic_aats.py:628, :1462)

        # every field and nuclear perturbation's first-order amplitudes, iterated as one set
        rs = ResponseSet(solver.ints, solver.fock, solver.eps, singles, t_ref, E)

        def add(df, dEd):
            dints = SPIntegrals.from_mo(dEd[t_, t_, t_, t_].permute(0, 2, 1, 3), O, singles)   # chemists' order
            rs.add(dints, tr.to_dev(np.ascontiguousarray(df[t_, t_]).astype(np.complex128), self.dev))

        U_H, nuc = [], []
        for b in range(3):
            U, h_core = self._magnetic(b, orbitals)
            Ud = tr.to_dev(U.astype(np.complex128), self.dev)
            add(self._df_magnetic(U, h_core), _dERI(Ud, self.Ed, (-1.0, -1.0, 1.0, 1.0)))
            U_H.append(U)
        for N1 in range(self.natom):
            for a in range(3):
                U_R, F1, S1, E1d, half = self._nuclear(N1, a, orbitals)
                Ud = tr.to_dev(U_R.astype(np.complex128), self.dev)
                add(self._df_nuclear(U_R, F1, S1), E1d + _dERI(Ud, self.Ed, (1.0, 1.0, 1.0, 1.0)))
                nuc.append((N1, a, U_R, half))
        _, dt1s, dt2s, _ = rs.solve(self.parameters, print_level)
        dT1 = list(np.real_if_close(dt1s.cpu().numpy())) if singles else [None] * dt2s.shape[0]
        dT2 = list(np.real_if_close(dt2s.cpu().numpy()))
        AAT = np.zeros((self.natom * 3, 3))
        for k, (N1, a, U_R, half) in enumerate(nuc):
            dt1, dt2 = dT1[3 + k], dT2[3 + k]
            s = _c('ijab,ijab', np.conj(dt2), tt) + _c('ijab,ijab', dt2, np.conj(tt))
            if singles:
                s += 2 * _c('ia,ia', np.conj(dt1), t1) + 2 * _c('ia,ia', dt1, np.conj(t1))
            N_R = -(1 / np.sqrt(norm2 ** 3)) * 0.5 * s
            R = U_R + half.T
            for b in range(3):
                val = N ** 2 * 2 * _c('em,em', U_H[b][v_, o], R[v_, o])
                val += self._doubles_part(orbitals, normalization, N, N_R, t2, dt2, dT2[b], U_H[b], R)
                if singles:
                    val += self._singles_part(orbitals, normalization, N, N_R, t1, t2, dt1, dt2, dT1[b], dT2[b],
                                              U_H[b], R)
                AAT[3 * N1 + a][b] += np.real(val)
        return AAT

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
        tt = 2 * t2 - t2.swapaxes(2, 3)
        N = 1.0 if normalization == 'intermediate' else 1 / np.sqrt(1 + _c('ijab,ijab', t2, tt))
        # first-order amplitudes on the device (apy_mp2_sp_response): the Fock terms in the
        # magnetic order (analytic_aats.py:346-350, form 1) and the nuclear order (:446-451, form 0)
        act = np.r_[np.arange(eps.shape[0])[o_], np.arange(eps.shape[0])[v_]]
        eps_act = np.concatenate([eo, ev])
        t2_dev = tr.to_dev(np.asarray(t2, dtype=np.complex128), self.dev)
        act_dev = torch.as_tensor(act, device=self.dev)
        no = eo.shape[0]

        def first_order(dE_oovv, df, form):
            f_act = tr.to_dev(np.ascontiguousarray(df[np.ix_(act, act)]).astype(np.complex128), self.dev)
            dt = mp2_grad.t2_response(t2_dev, f_act, dE_oovv, eps_act, form=form)
            return np.real_if_close(dt.cpu().numpy())

        def oovv(X):   # X[p,q,r,s] over all orbitals -> [o_, o_, v_, v_] of the active ones
            Xa = X.index_select(0, act_dev).index_select(1, act_dev).index_select(2, act_dev)
            return Xa.index_select(3, act_dev)[:no, :no, no:, no:]

        U_H, dT2_dH = [], []
        for b in range(3):
            U, h_core = self._magnetic(b, orbitals)
            df = self._df_magnetic(U, h_core)
            dE = _dERI(tr.to_dev(U.astype(np.complex128), self.dev), self.Ed, (-1.0, -1.0, 1.0, 1.0))
            U_H.append(U)
            dT2_dH.append(first_order(oovv(dE.permute(2, 3, 0, 1)), df, 1))
        AAT = np.zeros((self.natom * 3, 3))
        for N1 in range(self.natom):
            for a in range(3):
                U_R, F1, S1, E1d, half = self._nuclear(N1, a, orbitals)
                df = self._df_nuclear(U_R, F1, S1)
                dE = E1d + _dERI(tr.to_dev(U_R.astype(np.complex128), self.dev), self.Ed, (1.0, 1.0, 1.0, 1.0))
                dt = first_order(oovv(dE), df, 0)
                N_R = -(1 / np.sqrt((1 + _c('ijab,ijab', np.conj(t2), tt)) ** 3))
                N_R *= 0.5 * (_c('ijab,ijab', np.conj(dt), tt) + _c('ijab,ijab', dt, np.conj(tt)))
                R = U_R + half.T
                for b in range(3):
                    val = N ** 2 * 2 * _c('em,em', U_H[b][v_, o], R[v_, o])
                    val += self._doubles_part(orbitals, normalization, N, N_R, t2, dt, dT2_dH[b], U_H[b], R)
                    AAT[3 * N1 + a][b] += np.real(val)
        return AAT

    def compute_CISD_AATs(self, normalization='full', orbitals='non-canonical', print_level=0):
        """analytic_aats.py:553-1391.  Returns the [3N, 3] CISD AAT: HF,
        singles/reference, reference/singles, singles/singles,
        doubles/singles, singles/doubles, doubles/doubles and normalisation
        terms, from first-order CISD amplitudes solved on the GPU for every
        field component and nuclear displacement."""
        return self._ci_aats(True, normalization, orbitals, print_level)

    def compute_CID_AATs(self, normalization='full', orbitals='non-canonical', print_level=0):
        """analytic_aats.py:1392-1927.  Returns the [3N, 3] CID AAT (HF,
        doubles/doubles and normalisation terms) from first-order CID
        amplitudes solved on the GPU."""
        return self._ci_aats(False, normalization, orbitals, print_level)
