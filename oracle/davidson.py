"""Lowest CISD root by Davidson iteration and its energy gradients, NumPy
(oracle only; test infrastructure, never the product path).

The operator is the Hermitian spin-orbital CISD matrix (the reference's
ci_wfn.py:309-334 terms plus the P(ij)P(ab) f_bj t_ia coupling that the
reference's doubles residual leaves out; with canonical Hartree-Fock
orbitals, F_ov = 0, the two coincide).  Vectors are (c0, c1, c2) with c2
stored dense and antisymmetric; the inner product counts each determinant
once: conj(c0) d0 + sum conj(c1) d1 + 1/4 sum conj(c2) d2.

The gradient of the (variational) correlation energy with respect to F is
the normal-ordered one-particle density of the normalised vector, in the
form of the reference's adjoint sketch (ci_wfn.py:380-394):
    D_oo[i,j] = -sum_a conj(c1[j,a]) c1[i,a] - 1/2 sum conj(c2[j,k,a,b]) c2[i,k,a,b]
    D_vv[a,b] =  sum_i conj(c1[i,a]) c1[i,b] + 1/2 sum conj(c2[i,j,a,c]) c2[i,j,b,c]
    D_ov[i,a] =  conj(c0) c1[i,a] + sum conj(c1[j,b]) c2[i,j,a,b]
    D_vo[a,i] =  conj(c1[i,a]) c0 + sum conj(c2[i,j,a,b]) c1[j,b]
so that dE = sum_pq D[p,q] dF[p,q] for any one-electron perturbation.

The gradient with respect to the antisymmetrised integrals A = <pq||rs>
is the two-particle density G (two_particle_density; block layouts of the
device integrals in tpd_blocks).

Pinned by: the lowest eigenpair of the explicit CISD matrix (oracle.brute)
and, for D and G, Hellmann-Feynman derivatives of that matrix's lowest
eigenvalue with respect to each F[p,q] and along each antisymmetrised unit
direction of A (tests/test_oracle_davidson.py).
"""
import numpy as np

from .so import cisd_so_sigma_blocks, contract


def sigma(F, W, c0, t1, t2, O):
    """H_N (c0, t1, t2) -> (s0, s1, s2) of the Hermitian CISD matrix."""
    o, v = slice(0, O), slice(O, F.shape[0])
    s1, s2 = cisd_so_sigma_blocks(F, W, t1, t2, O)
    # cisd_so_sigma_blocks carries the reference driver with weight 1
    s1 = s1 + (c0 - 1.0) * F.swapaxes(0, 1)[o, v]
    s2 = s2 + (c0 - 1.0) * W['vvoo'].transpose(2, 3, 0, 1)
    Fvo = F[v, o]                                                     # F[O+b, j]
    X = contract('bj,ia->ijab', Fvo, t1)
    s2 = s2 + X - X.swapaxes(0, 1) - X.swapaxes(2, 3) + X.swapaxes(0, 1).swapaxes(2, 3)
    s0 = contract('ia,ia->', F[o, v], t1) + 0.25 * contract('ijab,ijab->', W['oovv'], t2)
    return s0, s1, s2


def dot(x, y):
    return np.conj(x[0]) * y[0] + np.vdot(x[1], y[1]) + 0.25 * np.vdot(x[2], y[2])


def davidson(F, W, O, tol=1e-12, rtol=1e-9, max_space=20, maxit=200):
    """Lowest root.  Returns (E, (c0, c1, c2) normalised, iterations)."""
    n = F.shape[0]
    V = n - O
    eps = np.real(np.diag(F))
    D1 = eps[O:][None, :] - eps[:O][:, None]
    D2 = (eps[O:][None, None, :, None] + eps[O:][None, None, None, :]
          - eps[:O][:, None, None, None] - eps[:O][None, :, None, None])
    B = [(1.0 + 0j, np.zeros((O, V), complex), np.zeros((O, O, V, V), complex))]
    S = [sigma(F, W, *B[0], O)]
    theta_old = None
    it = 0
    for it in range(1, maxit + 1):
        k = len(B)
        G = np.array([[dot(B[i], S[j]) for j in range(k)] for i in range(k)])
        w, U = np.linalg.eigh(0.5 * (G + G.conj().T))
        theta, y = w[0], U[:, 0]
        x = tuple(sum(y[j] * B[j][q] for j in range(k)) for q in range(3))
        r = tuple(sum(y[j] * (S[j][q] - theta * B[j][q]) for j in range(k)) for q in range(3))
        rn = np.sqrt(abs(dot(r, r)))
        if theta_old is not None and abs(theta - theta_old) < tol and rn < rtol:
            break
        theta_old = theta
        d2 = r[2] / (theta - D2)
        # keep the correction in the antisymmetric doubles space: rounding
        # outside it is amplified by small denominators and never removed
        d2 = 0.25 * (d2 - d2.swapaxes(0, 1) - d2.swapaxes(2, 3) + d2.swapaxes(0, 1).swapaxes(2, 3))
        d = (r[0] / theta if theta != 0 else r[0], r[1] / (theta - D1), d2)
        if k >= max_space:
            B, S = [x], [tuple(sum(y[j] * S[j][q] for j in range(k)) for q in range(3))]
        for _ in range(2):
            for b in B:
                c = dot(b, d)
                d = tuple(d[q] - c * b[q] for q in range(3))
        nd = np.sqrt(abs(dot(d, d)))
        d = tuple(q / nd for q in d)
        B.append(d)
        S.append(sigma(F, W, *d, O))
    nx = np.sqrt(abs(dot(x, x)))
    x = tuple(q / nx for q in x)
    return theta, x, it


def one_particle_density(c, O):
    """D[p,q] = dE/dF[p,q] of the normalised CISD vector c = (c0, c1, c2)."""
    c0, c1, c2 = c
    V = c1.shape[1]
    n = O + V
    D = np.zeros((n, n), complex)
    o, v = slice(0, O), slice(O, n)
    D[o, o] = -contract('ja,ia->ij', np.conj(c1), c1) - 0.5 * contract('jkab,ikab->ij', np.conj(c2), c2)
    D[v, v] = contract('ia,ib->ab', np.conj(c1), c1) + 0.5 * contract('ijac,ijbc->ab', np.conj(c2), c2)
    D[o, v] = np.conj(c0) * c1 + contract('jb,ijab->ia', np.conj(c1), c2)
    D[v, o] = np.conj(c1.T) * c0 + contract('ijab,jb->ai', np.conj(c2), c1)
    return D


def _antisym(G):
    return 0.25 * (G - G.swapaxes(0, 1) - G.swapaxes(2, 3) + G.swapaxes(0, 1).swapaxes(2, 3))


def two_particle_density(c, O):
    """G[p,q,r,s] = dE/dA[p,q,r,s] of the normalised CISD vector c = (c0, c1,
    c2), A = <pq||rs>, antisymmetric in (p,q) and in (r,s), so that
    dE = sum G dA for every antisymmetric dA and E = sum F D + sum A G.

    The blocks follow the reference's TPD sketch (ci_wfn.py:396-409, written
    there for the plain Dirac integrals) restated for the antisymmetrised
    blocks as the sigma equations read them (ci_wfn.py:309-334 plus the
    reference row); each block is the derivative of conj(c) . H c with
    respect to the block the sigma reads, then the whole tensor is
    antisymmetrised."""
    c0, c1, c2 = c
    V = c1.shape[1]
    n = O + V
    o, v = slice(0, O), slice(O, n)
    cc1, cc2 = np.conj(c1), np.conj(c2)
    G = np.zeros((n, n, n, n), complex)
    G[v, v, v, v] = 0.125 * contract('ijab,ijcd->abcd', cc2, c2)
    G[o, o, o, o] = 0.125 * contract('ijab,klab->klij', cc2, c2)
    G[o, v, v, o] = contract('ia,jb->jabi', cc1, c1) + contract('ijab,ikac->kbcj', cc2, c2)
    G[v, o, v, v] = 0.5 * contract('ia,ijcb->ajcb', cc1, c2)
    G[o, o, o, v] = -0.5 * contract('ia,kjab->kjib', cc1, c2)
    G[v, v, o, o] = 0.25 * c0 * cc2.transpose(2, 3, 0, 1)
    G[o, v, o, o] = -0.25 * contract('ijab,ka->kbij', cc2, c1)
    G[v, o, o, o] = -0.25 * contract('ijab,kb->akij', cc2, c1)
    G[v, v, v, o] = 0.25 * contract('ijab,ic->abcj', cc2, c1)
    G[v, v, o, v] = 0.25 * contract('ijab,jc->abic', cc2, c1)
    G[o, o, v, v] = 0.25 * np.conj(c0) * c2
    return _antisym(G)


def _pairs(N):
    return [(p, q) for p in range(N) for q in range(p + 1, N)]


def tpd_blocks(G, O):
    """The density in the packed block layouts of the device integrals
    (vvvv_p, ring, oooo_p, ovoo_p, vvvo_p, vovv_p, ooov, vvoo_t, oovv): each
    stored entry carries the sum over the full-tensor positions it stands
    for (4 for the packed blocks and ring, 2 for ooov, 1 for
    vvoo_t and oovv), so that dE = sum_blocks sum Gblk * dblk."""
    n = G.shape[0]
    V = n - O
    Po, Pv = _pairs(O), _pairs(V)
    oi = np.array([p for p, _ in Po], int), np.array([q for _, q in Po], int)
    vi = np.array([p for p, _ in Pv], int) + O, np.array([q for _, q in Pv], int) + O
    o, v = slice(0, O), slice(O, n)
    out = {}
    out['vvvv_p'] = 4 * G[vi[0][:, None], vi[1][:, None], vi[0][None, :], vi[1][None, :]]
    # ring[(j,b),(k,c)] = <kb||cj>
    out['ring'] = 4 * G[o, v, v, o].transpose(3, 1, 0, 2)
    # oooo_p[(i<j)][(k<l)] = <kl||ij>
    out['oooo_p'] = 4 * G[oi[0][None, :], oi[1][None, :], oi[0][:, None], oi[1][:, None]]
    # ovoo_p[k][b][(i<j)] = <kb||ij>
    out['ovoo_p'] = 4 * G[o, v][:, :, oi[0], oi[1]]
    # vvvo_p[(a<b)][j][c] = <ab||cj>
    out['vvvo_p'] = 4 * G[vi[0], vi[1]][:, v, o].transpose(0, 2, 1)
    # vovv_p[a][j][(c<b)] = <aj||cb>
    out['vovv_p'] = 4 * G[v, o][:, :, vi[0], vi[1]]
    out['ooov'] = 2 * G[o, o, o, v]
    out['vvoo_t'] = G[v, v, o, o].transpose(2, 3, 0, 1)
    out['oovv'] = G[o, o, v, v].copy()
    return out


def integral_blocks(A, O):
    """The same block layouts of the integrals themselves (A = <pq||rs>)."""
    n = A.shape[0]
    V = n - O
    Po, Pv = _pairs(O), _pairs(V)
    oi = np.array([p for p, _ in Po], int), np.array([q for _, q in Po], int)
    vi = np.array([p for p, _ in Pv], int) + O, np.array([q for _, q in Pv], int) + O
    o, v = slice(0, O), slice(O, n)
    return {
        'vvvv_p': A[vi[0][:, None], vi[1][:, None], vi[0][None, :], vi[1][None, :]],
        'ring': A[o, v, v, o].transpose(3, 1, 0, 2),
        'oooo_p': A[oi[0][None, :], oi[1][None, :], oi[0][:, None], oi[1][:, None]],
        'ovoo_p': A[o, v][:, :, oi[0], oi[1]],
        'vvvo_p': A[vi[0], vi[1]][:, v, o].transpose(0, 2, 1),
        'vovv_p': A[v, o][:, :, vi[0], vi[1]],
        'ooov': A[o, o, o, v].copy(),
        'vvoo_t': A[v, v, o, o].transpose(2, 3, 0, 1),
        'oovv': A[o, o, v, v].copy(),
    }
