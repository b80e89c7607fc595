"""Determinant-basis configuration interaction by second quantization
(oracle only; independent of every einsum in the restatement).

For n spin orbitals with the first O occupied in the reference determinant,
a Fock matrix F and antisymmetrized integrals A[p,q,r,s] = <pq||rs>, the
Hamiltonian is H = sum h_pq a+_p a_q + 1/4 sum <pq||rs> a+_p a+_q a_s a_r
with h_pq = F_pq - sum_m <pm||qm> (m occupied), so that F is the Fock
operator of the reference determinant.

Basis states are defined by operators acting on the reference:
    Phi0,  Phi_i^a = a+_a a_i Phi0,  Phi_ij^ab = a+_a a+_b a_j a_i Phi0 (i<j, a<b)
which is the convention of the amplitude equations in ci_wfn.py
(|Psi> = Phi0 + sum t_ia Phi_i^a + 1/4 sum t_ijab Phi_ij^ab).

Nothing here is a restatement of reference code: it is the textbook
definition of the CI matrix, used to pin oracle.so.
"""
import itertools

import numpy as np


def _apply_annihilate(bits, q):
    if not (bits >> q) & 1:
        return None, 0
    sign = -1 if bin(bits & ((1 << q) - 1)).count('1') % 2 else 1
    return bits & ~(1 << q), sign


def _apply_create(bits, p):
    if (bits >> p) & 1:
        return None, 0
    sign = -1 if bin(bits & ((1 << p) - 1)).count('1') % 2 else 1
    return bits | (1 << p), sign


def _apply_string(bits, ops):
    """Apply operators right-to-left; ops = [(p, is_create), ...] in written
    order (leftmost first)."""
    sign = 1
    for p, cre in reversed(ops):
        bits, s = (_apply_create if cre else _apply_annihilate)(bits, p)
        if bits is None:
            return None, 0
        sign *= s
    return bits, sign


class CISpace:
    def __init__(self, n, O, singles=True, doubles=True):
        self.n, self.O = n, O
        ref = (1 << O) - 1
        self.ref = ref
        states = [(ref, 1)]
        labels = [('0',)]
        occ = range(O)
        vir = range(O, n)
        if singles:
            for i in occ:
                for a in vir:
                    b, s = _apply_string(ref, [(a, True), (i, False)])
                    states.append((b, s))
                    labels.append(('S', i, a - O))
        if doubles:
            for i, j in itertools.combinations(occ, 2):
                for a, b_ in itertools.combinations(vir, 2):
                    b, s = _apply_string(ref, [(a, True), (b_, True), (j, False), (i, False)])
                    states.append((b, s))
                    labels.append(('D', i, j, a - O, b_ - O))
        self.states = states
        self.labels = labels
        self.index = {b: k for k, (b, _) in enumerate(states)}

    def __len__(self):
        return len(self.states)


def one_body(F, A, O):
    n = F.shape[0]
    h = F.copy()
    for m in range(O):
        h = h - A[:, m, :, m]
    return h


def hamiltonian(F, A, space):
    """Matrix of H in the operator-defined basis of `space`, and E_ref."""
    n, O = space.n, space.O
    h = one_body(F, A, O)
    N = len(space)
    H = np.zeros((N, N), dtype=np.complex128)
    for L, (bl, sl) in enumerate(space.states):
        occ = [p for p in range(n) if (bl >> p) & 1]
        out = {}
        # one-body
        for q in occ:
            b1, s1 = _apply_annihilate(bl, q)
            for p in range(n):
                b2, s2 = _apply_create(b1, p)
                if b2 is None:
                    continue
                out[b2] = out.get(b2, 0) + h[p, q] * s1 * s2
        # two-body: 1/4 sum <pq||rs> a+p a+q a_s a_r
        for r in occ:
            b1, s1 = _apply_annihilate(bl, r)
            for s in occ:
                if s == r:
                    continue
                b2, s2 = _apply_annihilate(b1, s)
                free = [p for p in range(n) if not (b2 >> p) & 1]
                for q in free:
                    b3, s3 = _apply_create(b2, q)
                    for p in free:
                        if p == q:
                            continue
                        b4, s4 = _apply_create(b3, p)
                        out[b4] = out.get(b4, 0) + 0.25 * A[p, q, r, s] * s1 * s2 * s3 * s4
        for b, val in out.items():
            K = space.index.get(b)
            if K is not None:
                H[K, L] += space.states[K][1] * sl * val
    E_ref = H[0, 0].real if abs(H[0, 0].imag) < 1e-12 else H[0, 0]
    return H, E_ref


def vector_from_amplitudes(space, t1, t2, c0=1.0):
    c = np.zeros(len(space), dtype=np.complex128)
    for k, lab in enumerate(space.labels):
        if lab[0] == '0':
            c[k] = c0
        elif lab[0] == 'S':
            c[k] = t1[lab[1], lab[2]]
        else:
            c[k] = t2[lab[1], lab[2], lab[3], lab[4]]
    return c


def amplitudes_from_vector(space, c, V):
    O = space.O
    t1 = np.zeros((O, V), dtype=np.complex128)
    t2 = np.zeros((O, O, V, V), dtype=np.complex128)
    c0 = None
    for k, lab in enumerate(space.labels):
        if lab[0] == '0':
            c0 = c[k]
        elif lab[0] == 'S':
            t1[lab[1], lab[2]] = c[k]
        else:
            _, i, j, a, b = lab
            t2[i, j, a, b] = c[k]
            t2[j, i, a, b] = -c[k]
            t2[i, j, b, a] = -c[k]
            t2[j, i, b, a] = c[k]
    return c0, t1, t2


def lowest_root(F, A, O, singles=True):
    """Lowest eigenpair of the CISD (or CID) matrix, as (E_corr, t1, t2)
    in intermediate normalization."""
    n = F.shape[0]
    V = n - O
    sp = CISpace(n, O, singles=singles, doubles=True)
    H, E_ref = hamiltonian(F, A, sp)
    w, U = np.linalg.eigh(H)
    c = U[:, 0]
    c0, t1, t2 = amplitudes_from_vector(sp, c / c[0], V)
    return w[0] - E_ref, t1, t2


def mp2_rspt2(F, A, O):
    """Second-order Rayleigh-Schroedinger energy with H0 = diag(F):
    sum over double excitations |<D|H|0>|^2 / (E0 - E_D)."""
    n = F.shape[0]
    sp = CISpace(n, O, singles=False, doubles=True)
    H, _ = hamiltonian(F, A, sp)
    eps = np.real(np.diag(F))
    e2 = 0.0
    for k, lab in enumerate(sp.labels):
        if lab[0] != 'D':
            continue
        _, i, j, a, b = lab
        d = eps[i] + eps[j] - eps[a + O] - eps[b + O]
        e2 += H[k, 0] * H[0, k] / d
    return e2
