"""Analytic atomic axial tensors, Hartree-Fock and MP2 levels (oracle only;
test infrastructure, never the product path).

Restates analytic_aats.py:30-212 (compute_RHF_AATs) and :215-550
(compute_MP2_AATs) in NumPy on the package's host integrals: the
coupled-perturbed Hartree-Fock (CPHF) coefficients U for each nuclear
displacement (skeleton derivative integrals, :84-141, :412-470) and each
magnetic-field component (the -1/2 L operator, :143-186, :278-330), with
the dependent (occupied-occupied, virtual-virtual) blocks fixed by the
'orbitals' option as the reference does; the first-order Fock matrices,
integrals and MP2 amplitudes (:332-358, :471-505); and the AAT terms
(HF, the four doubles terms and the normalisation term, :507-546).

Pinned by: the reference's analytic AAT arrays for H2O2/STO-3G
(test_018_RHF_Analytic_AAT.py:7, :62; test_019_MP2_Analytic_AAT.py:7, :62,
:117, :172; tests/golden/analytic_aat_goldens.json) and, for the derivative
integrals, central differences of the integrals themselves
(tests/test_oracle_analytic.py).
"""
import numpy as np

from apyib_amd import gaussian
from apyib_amd.hamiltonian import ANGULAR_MOMENTUM_SIGN, Hamiltonian

from . import scf, spatial, transforms


def contract(expr, *ops):
    return np.einsum(expr, *ops, optimize=True)


def _ndocc(H):
    return int(sum(H.molecule.true_atomic_number(a) for a in range(H.molecule.natom()))) // 2


def mo_tei(ERI, Cb, Ck):
    """(pq|rs) = sum C*_mp C_nq C*_lr C_gs (mn|lg) (analytic_aats.py:60-63)."""
    X = contract('mnlg,gs->mnls', ERI, Ck)
    X = contract('mnls,lr->mnrs', X, np.conj(Cb))
    X = contract('nq,mnrs->mqrs', Ck, X)
    return contract('mp,mqrs->pqrs', np.conj(Cb), X)


_DERIV_CACHE = {}


def _deriv1(H, parameters):
    key = (parameters['geom'], parameters['basis'])
    if key not in _DERIV_CACHE:
        charges = list(zip(H.molecule.Z, H.molecule.coords))
        _DERIV_CACHE[key] = gaussian.build_deriv1(H.basis_set.functions, H.basis_set.atom_of, charges)
    return _DERIV_CACHE[key]


class Context:
    """SCF solution, MO integrals and the CPHF matrices of one molecule."""

    def __init__(self, parameters):
        H = Hamiltonian(parameters)
        self.H = H
        self.no = no = _ndocc(H)
        self.nbf = nbf = H.basis_set.nbf()
        self.nv = nv = nbf - no
        _, C, eps, _ = scf.solve_scf(H.T, H.V, H.ERI, H.S, H.E_nuc, no, parameters)
        self.C = C = np.real_if_close(C)
        self.eps = np.real(eps)
        self.nfzc = H.basis_set.n_frozen_core() if parameters.get('freeze_core', False) else 0
        self.freeze_core = bool(parameters.get('freeze_core', False))
        self.C_list, _ = transforms.get_slices(nbf, no, self.nfzc, parameters['method'])
        self.o, self.v = slice(0, no), slice(no, nbf)
        o, v = self.o, self.v
        self.h = np.conj(C).T @ (H.T + H.V) @ C
        self.ERI = ERI = mo_tei(H.ERI, C, C).swapaxes(1, 2)                  # <pq|rs>
        self.F = self.h + contract('piqi->pq', 2 * ERI[:, o, :, o] - ERI.swapaxes(2, 3)[:, o, :, o])
        X = 2 * ERI - ERI.swapaxes(2, 3)
        self.A = (X + X.swapaxes(1, 3)).swapaxes(1, 2)
        self.Amag = (-X + X.swapaxes(1, 3)).swapaxes(1, 2)
        diag = contract('ab,ij,aibj->aibj', np.eye(nv), np.eye(no),
                        self.F[v, v].reshape(nv, 1, nv, 1) - self.F[o, o].reshape(1, no, 1, no))
        self.G = np.linalg.inv((diag + self.A[v, o, v, o]).reshape(nv * no, nv * no))
        self.Gm = np.linalg.inv((diag + self.Amag[v, o, v, o]).reshape(nv * no, nv * no))
        self.natom = H.molecule.natom()
        self.deriv = _deriv1(H, parameters)
        self.L = ANGULAR_MOMENTUM_SIGN * gaussian.build_angular_momentum(H.basis_set.functions)

    def _dep_denoms(self):
        o, v = self.o, self.v
        Do = (self.eps[o] - self.eps[o].reshape(-1, 1)) + np.eye(self.no)
        Dv = (self.eps[v] - self.eps[v].reshape(-1, 1)) + np.eye(self.nv)
        return Do, Dv

    def magnetic(self, b, orbitals):
        """(U_h, h_core) for field component b (analytic_aats.py:278-313)."""
        o, v, no, nv, nbf = self.o, self.v, self.no, self.nv, self.nbf
        C = self.C
        h_core = C.T @ (-0.5 * self.L[b]) @ C
        U = np.zeros((nbf, nbf))
        U[v, o] += (self.Gm @ h_core[v, o].reshape(nv * no)).reshape(nv, no)
        U[o, v] += U[v, o].T
        Am = self.Amag.swapaxes(1, 2)
        if self.freeze_core or orbitals == 'canonical':
            Do, Dv = self._dep_denoms()
            U[o, o] += (-h_core[o, o] + contract('em,iejm->ij', U[v, o], Am[o, v, o, o])) / Do
            U[v, v] += (-h_core[v, v] + contract('em,aebm->ab', U[v, o], Am[v, v, v, o])) / Dv
            for j in range(nbf):
                U[j, j] = 0
        if orbitals == 'non-canonical':
            f_, o_, v_, _ = self.C_list
            U[f_, f_] = 0
            U[o_, o_] = 0
            U[v_, v_] = 0
        return U, h_core

    def nuclear(self, N1, a, orbitals):
        """(U_R, F_core, S_core, ERI_core Dirac, half_S) for atom N1, axis a
        (analytic_aats.py:412-470)."""
        o, v, no, nv, nbf = self.o, self.v, self.no, self.nv, self.nbf
        C, F, A = self.C, self.F, self.A.swapaxes(1, 2)
        d = self.deriv
        T1, V1, S1 = (C.T @ d[k][N1][a] @ C for k in ('T', 'V', 'S'))
        E1 = contract('mp,nq,lr,gs,mnlg->pqrs', C, C, C, C, d['ERI'][N1][a]).swapaxes(1, 2)
        F1 = T1 + V1 + contract('piqi->pq', 2 * E1[:, o, :, o] - E1.swapaxes(2, 3)[:, o, :, o])
        B = -F1[v, o] + contract('ai,ii->ai', S1[v, o], F[o, o]) + 0.5 * contract('mn,amin->ai', S1[o, o], A[v, o, o, o])
        U = np.zeros((nbf, nbf))
        U[v, o] += (self.G @ B.reshape(nv * no)).reshape(nv, no)
        U[o, v] -= U[v, o].T + S1[o, v]
        if self.freeze_core or orbitals == 'canonical':
            Do, Dv = self._dep_denoms()
            Bo = (F1[o, o] - contract('ij,jj->ij', S1[o, o], F[o, o]) + contract('em,iejm->ij', U[v, o], A[o, v, o, o])
                  - 0.5 * contract('mn,imjn->ij', S1[o, o], A[o, o, o, o]))
            U[o, o] += Bo / Do
            Bv = (F1[v, v] - contract('ab,bb->ab', S1[v, v], F[v, v]) + contract('em,aebm->ab', U[v, o], A[v, v, v, o])
                  - 0.5 * contract('mn,ambn->ab', S1[o, o], A[v, o, v, o]))
            U[v, v] += Bv / Dv
            for j in range(nbf):
                U[j, j] = -0.5 * S1[j, j]
        if orbitals == 'non-canonical':
            f_, o_, v_, _ = self.C_list
            U[f_, f_] = -0.5 * S1[f_, f_]
            U[o_, o_] = -0.5 * S1[o_, o_]
            U[v_, v_] = -0.5 * S1[v_, v_]
        half = C.T @ d['half_S'][N1][a] @ C
        return U, F1, S1, E1, half


def rhf_aat(parameters, orbitals='non-canonical'):
    """[3N, 3] Hartree-Fock AAT (compute_RHF_AATs, analytic_aats.py:30-212):
    2 sum_em U_H[e,m] (U_R[e,m] + half_S[m,e]), e active virtual, m occupied."""
    ctx = Context(parameters)
    o, v_ = ctx.o, ctx.C_list[2]
    U_H = [ctx.magnetic(b, orbitals)[0] for b in range(3)]
    AAT = np.zeros((3 * ctx.natom, 3))
    for N1 in range(ctx.natom):
        for a in range(3):
            U_R, _, _, _, half = ctx.nuclear(N1, a, orbitals)
            for b in range(3):
                AAT[3 * N1 + a, b] += 2 * contract('em,em', U_H[b][v_, o], U_R[v_, o] + half[o, v_].T)
    return AAT


def _dERI(U, ERI, sgn):
    """sum over the four indices of U contracted into <pq|rs>
    (analytic_aats.py:344-347 with sgn = (-1, -1, +1, +1) pattern for the
    field, :479-483 all +1 for nuclei)."""
    out = sgn[0] * contract('tp,tqrs->pqrs', U, ERI)
    out = out + sgn[1] * contract('tq,ptrs->pqrs', U, ERI)
    out = out + sgn[2] * contract('tr,pqts->pqrs', U, ERI)
    out = out + sgn[3] * contract('ts,pqrt->pqrs', U, ERI)
    return out


def _df_magnetic(ctx, U, h_core):
    """Field derivative of the Fock matrix, oo and vv blocks
    (analytic_aats.py:300-308, :720-728)."""
    o, v, eps = ctx.o, ctx.v, ctx.eps
    Am = ctx.Amag.swapaxes(1, 2)
    df = np.zeros((ctx.nbf, ctx.nbf))
    df[o, o] += (-h_core[o, o] + U[o, o] * eps[o].reshape(-1, 1) - U[o, o].T * eps[o]
                 + contract('em,iejm->ij', U[v, o], Am[o, v, o, o]))
    df[v, v] += (-h_core[v, v] + U[v, v] * eps[v].reshape(-1, 1) - U[v, v].T * eps[v]
                 + contract('em,aebm->ab', U[v, o], Am[v, v, v, o]))
    return df


def _df_nuclear(ctx, U_R, F1, S1):
    """Nuclear derivative of the Fock matrix, oo and vv blocks
    (analytic_aats.py:457-469, :964-974)."""
    o, v, eps = ctx.o, ctx.v, ctx.eps
    A = ctx.A.swapaxes(1, 2)
    df = np.zeros((ctx.nbf, ctx.nbf))
    df[o, o] += (F1[o, o] + U_R[o, o] * eps[o].reshape(-1, 1) + U_R[o, o].T * eps[o]
                 + contract('em,iejm->ij', U_R[v, o], A[o, v, o, o])
                 - 0.5 * contract('mn,imjn->ij', S1[o, o], A[o, o, o, o]))
    df[v, v] += (F1[v, v] + U_R[v, v] * eps[v].reshape(-1, 1) + U_R[v, v].T * eps[v]
                 + contract('em,aebm->ab', U_R[v, o], A[v, v, v, o])
                 - 0.5 * contract('mn,ambn->ab', S1[o, o], A[v, o, v, o]))
    return df


def mp2_aat(parameters, normalization='full', orbitals='non-canonical'):
    """[3N, 3] MP2 AAT (compute_MP2_AATs, analytic_aats.py:215-550)."""
    ctx = Context(parameters)
    o, v = ctx.o, ctx.v
    f_, o_, v_, t_ = ctx.C_list
    eps = ctx.eps
    ERI = ctx.ERI
    # canonical MP2 amplitudes over the active orbitals (mp2_wfn.py:42-59)
    eri_act = ERI[t_, t_, t_, t_]
    O = ctx.no - ctx.nfzc
    _, t2 = spatial.solve_mp2(eri_act, eps[t_], O)
    eo, ev = eps[o_], eps[v_]
    D = eo.reshape(-1, 1, 1, 1) + eo.reshape(-1, 1, 1) - ev.reshape(-1, 1) - ev
    tt = 2 * t2 - t2.swapaxes(2, 3)
    N = 1.0 if normalization == 'intermediate' else 1 / np.sqrt(1 + contract('ijab,ijab', t2, tt))
    U_H, dT2_dH = [], []
    for b in range(3):
        U, h_core = ctx.magnetic(b, orbitals)
        df = _df_magnetic(ctx, U, h_core)
        dE = _dERI(U, ERI, (-1, -1, 1, 1))
        dt = dE.swapaxes(0, 2).swapaxes(1, 3)[o_, o_, v_, v_].copy()
        dt += contract('ac,ijcb->ijab', df[v_, v_], t2) + contract('bc,ijac->ijab', df[v_, v_], t2)
        dt -= contract('ki,kjab->ijab', df[o_, o_], t2) + contract('kj,ikab->ijab', df[o_, o_], t2)
        U_H.append(U)
        dT2_dH.append(dt / D)
    AAT = np.zeros((3 * ctx.natom, 3))
    for N1 in range(ctx.natom):
        for a in range(3):
            U_R, F1, S1, E1, half = ctx.nuclear(N1, a, orbitals)
            df = _df_nuclear(ctx, U_R, F1, S1)
            dE = E1 + _dERI(U_R, ERI, (1, 1, 1, 1))
            dt = dE[o_, o_, v_, v_].copy()
            dt -= contract('kjab,ik->ijab', t2, df[o_, o_]) + contract('ikab,kj->ijab', t2, df[o_, o_])
            dt += contract('ijcb,ac->ijab', t2, df[v_, v_]) + contract('ijac,cb->ijab', t2, df[v_, v_])
            dt = dt / D
            N_R = -(1 / np.sqrt((1 + contract('ijab,ijab', np.conj(t2), tt)) ** 3))
            N_R *= 0.5 * (contract('ijab,ijab', np.conj(dt), tt) + contract('ijab,ijab', dt, np.conj(tt)))
            R = U_R + half.T                      # U_R[p,q] + half_S[q,p]
            for b in range(3):
                val = N ** 2 * 2 * contract('em,em', U_H[b][v_, o], R[v_, o])
                val += _doubles_part(ctx, orbitals, normalization, N, N_R, t2, dt, dT2_dH[b], U_H[b], R)
                AAT[3 * N1 + a, b] += np.real(val)
    return AAT


def _doubles_part(ctx, orbitals, normalization, N, N_R, t2, dt2_R, dt2_H, UH, R):
    """Doubles/doubles and doubles normalisation terms of one AAT element,
    shared by MP2 (analytic_aats.py:491-528), CID (:1865-1907) and CISD
    (:1226-1249, :1290-1308).  R[p,q] = U_R[p,q] + half_S[q,p]."""
    o = ctx.o
    _, o_, v_, _ = ctx.C_list
    tt = 2 * t2 - t2.swapaxes(2, 3)
    dtt = 2 * dt2_R - dt2_R.swapaxes(2, 3)
    dHH = 2 * dt2_H - dt2_H.swapaxes(2, 3)
    can = orbitals == 'canonical'
    val = N ** 2 * contract('ijab,ijab', dtt, dt2_H)
    if can:
        val -= N ** 2 * 2.0 * contract('ijab,kjab,ki', dtt, t2, UH[o_, o_])
        val += N ** 2 * 2.0 * contract('ijab,ijcb,ac', dtt, t2, UH[v_, v_])
    val -= N ** 2 * 2.0 * contract('ijab,kjab,ki', dHH, t2, R[o_, o_])
    val += N ** 2 * 2.0 * contract('ijab,ijcb,ac', dHH, t2, R[v_, v_])
    val += N ** 2 * 2.0 * contract('ijab,kjab,km,im', t2, tt, UH[o_, o], R[o_, o])
    if can:
        val += N ** 2 * 2.0 * contract('ijab,ijcb,ec,ea', t2, tt, UH[v_, v_], R[v_, v_])
    val += N ** 2 * 2.0 * contract('ijab,ijab,em,em', t2, tt, UH[v_, o], R[v_, o])
    val -= N ** 2 * 2.0 * contract('ijab,imab,ej,em', t2, tt, UH[v_, o_], R[v_, o_])
    val -= N ** 2 * 2.0 * contract('ijab,ijae,bm,em', t2, tt, UH[v_, o], R[v_, o])
    if normalization == 'full':
        if can:
            val -= N * N_R * 2.0 * contract('ijab,kjab,ki', tt, t2, UH[o_, o_])
            val += N * N_R * 2.0 * contract('ijab,ijcb,ac', tt, t2, UH[v_, v_])
        val += N * N_R * contract('ijab,ijab', tt, dt2_H)
    return val


def _singles_part(ctx, orbitals, normalization, N, N_R, t1, t2, dt1_R, dt2_R, dt1_H, dt2_H, UH, R):
    """Reference/singles, singles/singles, singles/doubles terms and the
    singles normalisation terms of one CISD AAT element
    (analytic_aats.py:1163-1224, :1251-1288 canonical; :1252-1288 non-canonical
    branch, :1271-1310).  Each row: (coefficient, subscripts, operands)."""
    o = ctx.o
    _, o_, v_, _ = ctx.C_list
    tt = 2 * t2 - t2.swapaxes(2, 3)
    dtt = 2 * dt2_R - dt2_R.swapaxes(2, 3)
    dHH = 2 * dt2_H - dt2_H.swapaxes(2, 3)
    if orbitals == 'canonical':
        rows = [
            # singles / reference
            (2, 'ia,ai', dt1_R, UH[v_, o_]),
            (2, 'ia,ei,ea', t1, UH[v_, o_], R[v_, v_]),
            (-2, 'ia,am,im', t1, UH[v_, o], R[o_, o]),
            # reference / singles
            (2, 'kc,ck', dt1_H, R[v_, o_]),
            (2, 'kc,fc,fk', t1, UH[v_, v_], R[v_, o_]),
            (-2, 'kc,kn,cn', t1, UH[o_, o], R[v_, o]),
            # singles / singles
            (2, 'ia,ia', dt1_R, dt1_H),
            (2, 'kc,cf,kf', dt1_R, UH[v_, v_], t1),
            (-2, 'kc,nk,nc', dt1_R, UH[o_, o_], t1),
            (2, 'ia,ae,ie', dt1_H, R[v_, v_], t1),
            (-2, 'ia,mi,ma', dt1_H, R[o_, o_], t1),
            (2, 'kc,fc,fa,ka', t1, UH[v_, v_], R[v_, v_], t1),
            (-2, 'kc,fc,ik,if', t1, UH[v_, v_], R[o_, o_], t1),
            (-2, 'kc,kn,ca,na', t1, UH[o_, o_], R[v_, v_], t1),
            (2, 'kc,kn,in,ic', t1, UH[o_, o], R[o_, o], t1),
            (4, 'kc,kc,ia,ia', t1, UH[o_, v_], R[o_, v_], t1),
            (4, 'kc,fn,fn,kc', t1, UH[v_, o], R[v_, o], t1),
            (-2, 'kc,fn,fk,nc', t1, UH[v_, o_], R[v_, o_], t1),
            (-2, 'kc,fn,cn,kf', t1, UH[v_, o], R[v_, o], t1),
            (4, 'kc,fn,ck,nf', t1, UH[v_, o_], R[v_, o_], t1),
            # doubles / singles
            (2, 'ijab,bj,ia', dtt, UH[v_, o_], t1),
            (2, 'kc,ia,ikac', dt1_H, R[o_, v_], tt),
            (2, 'kc,fc,ia,ikaf', t1, UH[v_, v_], R[o_, v_], tt),
            (-2, 'kc,kn,ia,inac', t1, UH[o_, o_], R[o_, v_], tt),
            (-2, 'kc,fn,ik,incf', t1, UH[v_, o_], R[o_, o_], tt),
            (-2, 'kc,fn,in,ikfc', t1, UH[v_, o], R[o_, o], tt),
            (2, 'kc,fn,ca,knaf', t1, UH[v_, o_], R[v_, v_], tt),
            (2, 'kc,fn,fa,knca', t1, UH[v_, o_], R[v_, v_], tt),
            # singles / doubles
            (2, 'ia,kc,ikac', dt1_R, UH[o_, v_], tt),
            (2, 'klcd,dl,kc', dHH, R[v_, o_], t1),
            (2, 'ia,kc,ea,kice', t1, UH[o_, v_], R[v_, v_], tt),
            (-2, 'ia,kc,im,kmca', t1, UH[o_, v_], R[o_, o_], tt),
            (2, 'ia,ac,em,imce', t1, UH[v_, v_], R[v_, o_], tt),
            (2, 'ia,ec,em,imac', t1, UH[v_, v_], R[v_, o_], tt),
            (-2, 'ia,ki,em,kmae', t1, UH[o_, o_], R[v_, o_], tt),
            (-2, 'ia,km,em,kiea', t1, UH[o_, o], R[v_, o], tt),
        ]
        norm = [
            (2, 'ia,ai', t1, UH[v_, o_]),
            (2, 'kc,kc', t1, UH[o_, v_]),
            (2, 'ia,ia', t1, dt1_H),
            (2, 'kc,cf,kf', t1, UH[v_, v_], t1),
            (-2, 'kc,nk,nc', t1, UH[o_, o_], t1),
            (2, 'ijab,bj,ia', tt, UH[v_, o_], t1),
            (2, 'ia,kc,ikac', t1, UH[o_, v_], tt),
        ]
    else:
        rows = [
            (2, 'ia,ai', dt1_R, UH[v_, o_]),
            (2, 'ia,ei,ea', t1, UH[v_, o_], R[v_, v_]),
            (-2, 'ia,am,im', t1, UH[v_, o], R[o_, o]),
            (2, 'ia,ai', dt1_H, R[v_, o_]),
            (-2, 'ia,im,am', t1, UH[o_, o], R[v_, o]),
            (2, 'ia,ia', dt1_R, dt1_H),
            (2, 'ia,ae,ie', dt1_H, R[v_, v_], t1),
            (-2, 'ia,mi,ma', dt1_H, R[o_, o_], t1),
            (2, 'ia,im,km,ka', t1, UH[o_, o], R[o_, o], t1),
            (4, 'ia,ia,kc,kc', t1, UH[o_, v_], R[o_, v_], t1),
            (4, 'ia,em,em,ia', t1, UH[v_, o], R[v_, o], t1),
            (-2, 'ia,em,ei,ma', t1, UH[v_, o_], R[v_, o_], t1),
            (-2, 'ia,em,am,ie', t1, UH[v_, o], R[v_, o], t1),
            (4, 'ia,em,ai,me', t1, UH[v_, o_], R[v_, o_], t1),
            (2, 'ijab,bj,ia', dtt, UH[v_, o_], t1),
            (2, 'ia,kc,ikac', dt1_H, R[o_, v_], tt),
            (-2, 'ia,em,km,kiea', t1, UH[v_, o], R[o_, o], tt),
            (2, 'ia,em,ec,imac', t1, UH[v_, o_], R[v_, v_], tt),
            (2, 'ia,kc,ikac', dt1_R, UH[o_, v_], tt),
            (2, 'ijab,bj,ia', dHH, R[v_, o_], t1),
            (-2, 'ia,km,em,kiea', t1, UH[o_, o], R[v_, o], tt),
        ]
        norm = [
            (4, 'ia,ai', t1, UH[v_, o_]),
            (2, 'ia,ia', t1, dt1_H),
            (4, 'ijab,bj,ia', tt, UH[v_, o_], t1),
        ]
    val = N ** 2 * sum(c * contract(e, *ops) for c, e, *ops in rows)
    if normalization == 'full':
        val += N * N_R * sum(c * contract(e, *ops) for c, e, *ops in norm)
    return val


def ci_densities(t1, t2, O):
    """Normalised one- and two-particle densities over the active orbitals
    (CISD analytic_aats.py:634-666; CID :1461-1484 with t1 = None), Dirac
    order D_pqrs <-> <pq|rs>."""
    nv = t2.shape[2]
    n = O + nv
    o, v = slice(0, O), slice(O, n)
    c = np.conj
    s1 = 0 if t1 is None else 2 * contract('ia,ia->', c(t1), t1)
    N = 1 / np.sqrt(1 + s1 + contract('ijab,ijab->', c(t2), 2 * t2 - t2.swapaxes(2, 3)))
    t0n, t2n = N, t2 * N
    ttn = 2 * t2n - t2n.swapaxes(2, 3)
    dt = np.result_type(t2, t2 if t1 is None else t1, float)
    Dp = np.zeros((n, n), dtype=dt)
    Dr = np.zeros((n, n, n, n), dtype=dt)
    Dp[o, o] -= 2 * contract('jkab,ikab->ij', c(ttn), t2n)
    Dp[v, v] += 2 * contract('ijac,ijbc->ab', c(ttn), t2n)
    Dr[o, o, o, o] += contract('klab,ijab->ijkl', c(t2n), ttn)
    Dr[v, v, v, v] += contract('ijab,ijcd->abcd', c(t2n), ttn)
    Dr[v, o, o, v] += 2 * contract('jkac,ikbc->aijb', c(ttn), ttn)
    Dr[v, o, v, o] -= 4 * contract('jkac,ikbc->aibj', c(t2n), t2n)
    Dr[v, o, v, o] += 2 * contract('jkac,ikcb->aibj', c(t2n), t2n)
    Dr[v, o, v, o] += 2 * contract('jkca,ikbc->aibj', c(t2n), t2n)
    Dr[v, o, v, o] -= 4 * contract('jkca,ikcb->aibj', c(t2n), t2n)
    Dr[o, o, v, v] += np.conj(t0n) * ttn
    Dr[v, v, o, o] += c(ttn).transpose(2, 3, 0, 1) * t0n
    if t1 is not None:
        t1n = t1 * N
        Dp[o, o] -= 2 * contract('ja,ia->ij', c(t1n), t1n)
        Dp[v, v] += 2 * contract('ia,ib->ab', c(t1n), t1n)
        Dp[o, v] += 2 * np.conj(t0n) * t1n + 2 * contract('jb,ijab->ia', c(t1n), t2n - t2n.swapaxes(2, 3))
        Dp[v, o] += 2 * c(t1n.T) * t0n + 2 * contract('ijab,jb->ai', c(t2n - t2n.swapaxes(2, 3)), t1n)
        Dr[o, v, v, o] += 4 * contract('ja,ib->iabj', c(t1n), t1n)
        Dr[o, v, o, v] -= 2 * contract('ja,ib->iajb', c(t1n), t1n)
        Dr[v, o, v, v] += 2 * contract('ja,ijcb->aibc', c(t1n), ttn)
        Dr[o, v, o, o] -= 2 * contract('kjab,ib->iajk', c(ttn), t1n)
        Dr[v, v, v, o] += 2 * contract('jiab,jc->abci', c(ttn), t1n)
        Dr[o, o, o, v] -= 2 * contract('kb,ijba->ijka', c(t1n), ttn)
    return Dp, Dr


def ci_response(F, ERI, eps, O, E, t1, t2, dF, dERI, dE_guess, params):
    """First-order CISD (t1 given) or CID (t1 None) amplitudes for one
    perturbation (dF, dERI) of the active-space Fock matrix and <pq|rs>:
    the reference's Jacobi + DIIS iteration (analytic_aats.py:733-890,
    :980-1143; CID :1540-1660).  The residual is the perturbed Hamiltonian's
    CISD residual of (1, t) with E -> the projected derivative, plus the
    unperturbed one of (0, dt) with the converged correlation energy E
    (no driver); the projected derivative is the CISD energy expression of
    (dF, dERI) on t plus that of (F, ERI) on dt.  The guess is the
    perturbed residual of t without its driver, with the density gradient
    dE_guess in place of the projected derivative.
    Returns (dE_proj, dt1 or None, dt2, iterations)."""
    n = F.shape[0]
    o, v = slice(0, O), slice(O, n)
    D1, D2 = spatial._denoms(eps, O)
    singles = t1 is not None
    t1z = t1 if singles else np.zeros((O, n - O))
    drv1, drv2 = F.swapaxes(0, 1)[o, v], spatial._driver(ERI, O)
    p1, p2 = spatial.cisd_residual(dF, dERI, t1z, t2, dE_guess, O)
    dt2 = (p2 - spatial._driver(dERI, O)) / D2
    dt1 = (p1 - dF.swapaxes(0, 1)[o, v]) / D1 if singles else np.zeros_like(t1z)
    e0 = spatial.cisd_energy(dF, dERI, t1z, t2, O)

    def resid(a, b, s):
        q1, q2 = spatial.cisd_residual(dF, dERI, t1z, t2, s, O)
        r1, r2 = spatial.cisd_residual(F, ERI, a, b, E, O)
        return q1 + r1 - drv1, q2 + r2 - drv2

    if singles:
        dE, (dt1, dt2), it = spatial._iterate(resid, lambda a, b: e0 + spatial.cisd_energy(F, ERI, a, b, O),
                                              [dt1, dt2], [D1, D2], params)
        return dE, dt1, dt2, it
    z = np.zeros_like(t1z)
    dE, (dt2,), it = spatial._iterate(lambda b, s: (resid(z, b, s)[1],),
                                      lambda b: e0 + spatial.cisd_energy(F, ERI, z, b, O), [dt2], [D2], params)
    return dE, None, dt2, it


def ci_aat(parameters, method='CISD', normalization='full', orbitals='non-canonical', record=None):
    """[3N, 3] CISD or CID AAT (compute_CISD_AATs analytic_aats.py:553-1391,
    compute_CID_AATs :1392-1927).  The normalisation factor N is the full
    one whatever `normalization` says (the reference reassigns it before
    use, :628 / :1462); 'intermediate' drops the normalisation-derivative
    terms only.  `record`, when a dict, receives per-perturbation data
    (guess and converged energy derivatives, amplitudes) for tests."""
    ctx = Context(parameters)
    o = ctx.o
    f_, o_, v_, t_ = ctx.C_list
    eps = ctx.eps
    ERI = ctx.ERI
    O = ctx.no - ctx.nfzc
    Fa, Ea, ea = ctx.F[t_, t_], ERI[t_, t_, t_, t_], eps[t_]
    singles = method == 'CISD'
    if singles:
        E, t1, t2, _ = spatial.solve_cisd(Fa, Ea, ea, O, parameters)
        t1 = np.real_if_close(t1)
    else:
        E, t2, _ = spatial.solve_cid(Fa, Ea, ea, O, parameters)
        t1 = None
    t2 = np.real_if_close(t2)
    E = np.real_if_close(E)
    Dp, Dr = ci_densities(t1, t2, O)
    tt = 2 * t2 - t2.swapaxes(2, 3)
    norm2 = 1 + contract('ijab,ijab', np.conj(t2), tt) + (2 * contract('ia,ia', np.conj(t1), t1) if singles else 0)
    N = 1 / np.sqrt(norm2)

    def solve(df, dE):
        dF, dA = df[t_, t_], dE[t_, t_, t_, t_]
        guess = contract('pq,pq->', dF, Dp) + contract('pqrs,pqrs->', dA, Dr)
        out = ci_response(Fa, Ea, ea, O, E, t1, t2, dF, dA, guess, parameters)
        if record is not None:
            record.setdefault('guess', []).append(guess)
            record.setdefault('dE', []).append(out[0])
            record.setdefault('dt', []).append((out[1], out[2]))
        return out[1], np.real_if_close(out[2])

    U_H, dT1_dH, dT2_dH = [], [], []
    for b in range(3):
        U, h_core = ctx.magnetic(b, orbitals)
        dt1, dt2 = solve(_df_magnetic(ctx, U, h_core), _dERI(U, ERI, (-1, -1, 1, 1)))
        U_H.append(U)
        dT1_dH.append(dt1)
        dT2_dH.append(dt2)
    AAT = np.zeros((3 * ctx.natom, 3))
    for N1 in range(ctx.natom):
        for a in range(3):
            U_R, F1, S1, E1, half = ctx.nuclear(N1, a, orbitals)
            dt1, dt2 = solve(_df_nuclear(ctx, U_R, F1, S1), E1 + _dERI(U_R, ERI, (1, 1, 1, 1)))
            s = (contract('ijab,ijab', np.conj(dt2), tt) + contract('ijab,ijab', dt2, np.conj(tt)))
            if singles:
                s += 2 * contract('ia,ia', np.conj(dt1), t1) + 2 * contract('ia,ia', dt1, np.conj(t1))
            N_R = -(1 / np.sqrt(norm2 ** 3)) * 0.5 * s
            R = U_R + half.T
            for b in range(3):
                val = N ** 2 * 2 * contract('em,em', U_H[b][v_, o], R[v_, o])
                val += _doubles_part(ctx, orbitals, normalization, N, N_R, t2, dt2, dT2_dH[b], U_H[b], R)
                if singles:
                    val += _singles_part(ctx, orbitals, normalization, N, N_R, t1, t2, dt1, dt2, dT1_dH[b],
                                         dT2_dH[b], U_H[b], R)
                AAT[3 * N1 + a, b] += np.real(val)
    return AAT
