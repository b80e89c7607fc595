"""Seeded synthetic spin-orbital Hamiltonians (host side, NumPy).

The synthetic workloads are defined by a counter-based generator so that any
block of the n^4 integral tensor can be produced independently, on the host
(here, for small sizes and for checking) or on the GPU (apy_synth_* in the C
ABI, same definition) without materialising the whole tensor:

  G(p,q,r,s)  complex normal, E|G|^2 = std^2, from splitmix64 of
              (seed, tag=1, counter=((p*n+q)*n+r)*n+s) via Box-Muller
  X(pqrs)    = (G(pqrs) + G(qpsr)) - (G(qprs) + G(pqsr))
  A[p,q,r,s] = 1/8 [ X(pqrs) + conj(X(rspq)) ]
              = <pq||rs>: antisymmetric in (pq) and in (rs), and Hermitian,
              A[p,q,r,s] = conj(A[r,s,p,q])
  eps_p       occupied U[-1.5,-0.3], virtual U[0.2,3.0] Eh   (tag 2)
  F           diag(eps) + (X + X^H)/2, X complex normal with std `noise`
              off the diagonal, zero on it                     (tag 3)

Perturbed instances (finite-difference displacements) add `scale` times a
second draw (tag 4 for integrals, 5 for the Fock matrix) given the symmetry
of the perturbation: "nuclear" = real and symmetric (real-symmetric Fock
part, real part of the integral draw), "field" = i times a real
antisymmetric one-electron operator (Fock only).
"""
import numpy as np

GOLD = np.uint64(0x9E3779B97F4A7C15)
C2 = np.uint64(0xD1B54A32D192ED03)
M1 = np.uint64(0xBF58476D1CE4E5B9)
M2 = np.uint64(0x94D049BB133111EB)
TWO_M53 = 1.0 / 9007199254740992.0

TAG_ERI, TAG_EPS, TAG_FOCK, TAG_DERI, TAG_DFOCK = 1, 2, 3, 4, 5


def _mix(z):
    z = (z ^ (z >> np.uint64(30))) * M1
    z = (z ^ (z >> np.uint64(27))) * M2
    return z ^ (z >> np.uint64(31))


def _hash2(seed, tag, counter):
    with np.errstate(over='ignore'):
        base = np.uint64(seed) * GOLD + np.uint64(tag) * C2
        h1 = _mix(base + np.asarray(counter, dtype=np.uint64) * np.uint64(2) + np.uint64(1))
        h2 = _mix(h1 + GOLD)
    return h1, h2


def uniform(seed, tag, counter):
    h1, _ = _hash2(seed, tag, counter)
    return (h1 >> np.uint64(11)).astype(np.float64) * TWO_M53


def cnormal(seed, tag, counter, std):
    """Complex normal with E|z|^2 = std^2 (Box-Muller on two 53-bit uniforms)."""
    h1, h2 = _hash2(seed, tag, counter)
    u1 = ((h1 >> np.uint64(11)).astype(np.float64) + 0.5) * TWO_M53
    u2 = (h2 >> np.uint64(11)).astype(np.float64) * TWO_M53
    r = np.sqrt(-2.0 * np.log(u1)) * (std / np.sqrt(2.0))
    return r * np.cos(2.0 * np.pi * u2) + 1j * (r * np.sin(2.0 * np.pi * u2))


def _g(seed, tag, n, p, q, r, s, std):
    cnt = ((p.astype(np.uint64) * np.uint64(n) + q) * np.uint64(n) + r) * np.uint64(n) + s
    return cnormal(seed, tag, cnt, std)


def eri_block(n, seed, ps, qs, rs, ss, std=0.05, tag=TAG_ERI):
    """A[p,q,r,s] = <pq||rs> for index ranges ps, qs, rs, ss (arrays or ranges)."""
    p, q, r, s = np.meshgrid(np.asarray(ps, dtype=np.uint64), np.asarray(qs, dtype=np.uint64),
                             np.asarray(rs, dtype=np.uint64), np.asarray(ss, dtype=np.uint64), indexing='ij')
    X1 = _xsum(seed, tag, n, p, q, r, s, std)
    X2 = _xsum(seed, tag, n, r, s, p, q, std)
    return 0.125 * (X1 + np.conj(X2))


def eri_elements(n, seed, p, q, r, s, std=0.05, tag=TAG_ERI):
    """A[p_k,q_k,r_k,s_k] = <pq||rs> for matching index arrays (scattered
    elements, no grid)."""
    p, q, r, s = (np.asarray(x, dtype=np.uint64) for x in (p, q, r, s))
    return 0.125 * (_xsum(seed, tag, n, p, q, r, s, std) + np.conj(_xsum(seed, tag, n, r, s, p, q, std)))


def _xsum(seed, tag, n, p, q, r, s, std):
    """(G(pqrs) + G(qpsr)) - (G(qprs) + G(pqsr)): grouped so that pair
    antisymmetry, zero diagonals and Hermiticity hold exactly in floating
    point (each swap only reorders commutative additions or negates)."""
    g = _g
    return ((g(seed, tag, n, p, q, r, s, std) + g(seed, tag, n, q, p, s, r, std))
            - (g(seed, tag, n, q, p, r, s, std) + g(seed, tag, n, p, q, s, r, std)))


def eri_full(n, seed, std=0.05):
    r = np.arange(n)
    return eri_block(n, seed, r, r, r, r, std)


def orbital_energies(O, V, seed):
    n = O + V
    u = uniform(seed, TAG_EPS, np.arange(n, dtype=np.uint64))
    eps = np.empty(n)
    eps[:O] = -1.5 + 1.2 * u[:O]
    eps[O:] = 0.2 + 2.8 * u[O:]
    return eps


def fock(O, V, seed, noise=1e-3):
    n = O + V
    eps = orbital_energies(O, V, seed)
    p, q = np.meshgrid(np.arange(n, dtype=np.uint64), np.arange(n, dtype=np.uint64), indexing='ij')
    X = cnormal(seed, TAG_FOCK, p * np.uint64(n) + q, noise)
    np.fill_diagonal(X, 0.0)
    F = np.diag(eps).astype(np.complex128) + 0.5 * (X + X.conj().T)
    return F, eps


def fock_perturbation(n, seed, kind):
    """One-electron perturbation direction: 'nuclear' real symmetric,
    'field' i * real antisymmetric (a Hermitian, purely imaginary operator)."""
    p, q = np.meshgrid(np.arange(n, dtype=np.uint64), np.arange(n, dtype=np.uint64), indexing='ij')
    X = cnormal(seed, TAG_DFOCK, p * np.uint64(n) + q, 1.0).real
    if kind == 'nuclear':
        return (0.5 * (X + X.T)).astype(np.complex128)
    if kind == 'field':
        return 0.5j * (X - X.T)
    raise ValueError(kind)


def eri_perturbation_full(n, seed):
    """Real part of a second antisymmetrized Hermitian draw: keeps the pair
    antisymmetry and, being real, the Hermiticity of a real-orbital
    displacement."""
    r = np.arange(n)
    return eri_block(n, seed, r, r, r, r, 1.0, tag=TAG_DERI).real.astype(np.complex128)


def system(O, V, seed, std=0.05, noise=1e-3):
    """Small full synthetic system: (F, A, eps) with A the full <pq||rs>."""
    F, eps = fock(O, V, seed, noise)
    A = eri_full(O + V, seed, std)
    return F, A, eps
