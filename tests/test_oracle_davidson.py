"""The Davidson oracle against the explicit CISD matrix, and its
one-particle density against Hellmann-Feynman derivatives of that matrix's
lowest eigenvalue.  CPU only."""
import numpy as np
import pytest

from apyib_amd import synthetic
from oracle import brute, davidson, so


def _system(O, V, seed, noise=0.05):
    """A synthetic Hermitian system with a sizeable F_ov block (non-Brillouin)."""
    F, A, eps = synthetic.system(O, V, seed, noise=noise)
    return F, A, so.blocks_from_eri(A / 2, O)        # blocks_from_eri antisymmetrises a Dirac tensor


@pytest.mark.parametrize("O,V,seed", [(2, 3, 1), (3, 3, 2), (4, 4, 3), (2, 6, 4)])
def test_davidson_lowest_root_matches_explicit_matrix(O, V, seed):
    F, A, W = _system(O, V, seed)
    E_ref, t1_ref, t2_ref = brute.lowest_root(F, A, O)
    E, (c0, c1, c2), _ = davidson.davidson(F, W, O, tol=1e-13, rtol=1e-10)
    assert abs(E - E_ref) < 1e-11
    assert np.abs(c1 / c0 - t1_ref).max() < 1e-8
    assert np.abs(c2 / c0 - t2_ref).max() < 1e-8


def test_sigma_is_hermitian_in_ci_metric():
    O, V = 3, 4
    F, A, W = _system(O, V, 7)
    rng = np.random.default_rng(0)

    def vec():
        c1 = rng.standard_normal((O, V)) + 1j * rng.standard_normal((O, V))
        c2 = rng.standard_normal((O, O, V, V)) + 1j * rng.standard_normal((O, O, V, V))
        c2 = c2 - c2.swapaxes(0, 1)
        c2 = c2 - c2.swapaxes(2, 3)
        return (complex(rng.standard_normal(), rng.standard_normal()), c1, c2)
    x, y = vec(), vec()
    hx, hy = davidson.sigma(F, W, *x, O), davidson.sigma(F, W, *y, O)
    assert abs(davidson.dot(x, hy) - np.conj(davidson.dot(y, hx))) < 1e-12


@pytest.mark.parametrize("O,V,seed", [(2, 3, 5), (3, 3, 6)])
def test_one_particle_density_is_energy_gradient(O, V, seed):
    """D[p,q] = d E_corr / d F[p,q]: c^H (dH/dF_pq) c - dE_ref/dF_pq with
    the explicit matrix and its normalised lowest eigenvector."""
    F, A, W = _system(O, V, seed)
    n = O + V
    sp = brute.CISpace(n, O, singles=True, doubles=True)
    H, _ = brute.hamiltonian(F, A, sp)
    w, U = np.linalg.eigh(H)
    u = U[:, 0]
    _, (c0, c1, c2), _ = davidson.davidson(F, W, O, tol=1e-13, rtol=1e-11)
    D = davidson.one_particle_density((c0, c1, c2), O)
    Z = np.zeros_like(A)
    for p in range(n):
        for q in range(n):
            E1 = np.zeros((n, n), complex)
            E1[p, q] = 1.0
            Hpq, Eref_pq = brute.hamiltonian(E1, Z, sp)
            g = np.vdot(u, Hpq @ u) - Eref_pq
            assert abs(D[p, q] - g) < 1e-9, (p, q)


def _unit_asym(n, p, q, r, s):
    X = np.zeros((n, n, n, n), complex)
    X[p, q, r, s] += 1.0
    X[q, p, r, s] -= 1.0
    X[p, q, s, r] -= 1.0
    X[q, p, s, r] += 1.0
    return X


@pytest.mark.parametrize("O,V,seed", [(2, 3, 5), (3, 3, 6)])
def test_two_particle_density_is_energy_gradient(O, V, seed):
    """G[p,q,r,s] = dE_corr/dA along the antisymmetrised unit direction at
    (p<q, r<s): c^H (dH) c - dE_ref with the explicit matrix (h depends on A
    through -sum_m <pm||qm>), every position of every o/v block, including
    directions that break Hermiticity (first-order perturbation theory holds
    for any direction at a Hermitian point)."""
    F, A, W = _system(O, V, seed)
    n = O + V
    sp = brute.CISpace(n, O, singles=True, doubles=True)
    H, _ = brute.hamiltonian(F, A, sp)
    w, U = np.linalg.eigh(H)
    u = U[:, 0]
    _, c, _ = davidson.davidson(F, W, O, tol=1e-13, rtol=1e-11)
    G = davidson.two_particle_density(c, O)
    Z = np.zeros((n, n), complex)
    for p in range(n):
        for q in range(p + 1, n):
            for r in range(n):
                for s in range(r + 1, n):
                    X = _unit_asym(n, p, q, r, s)
                    Hx, Ex = brute.hamiltonian(Z, X, sp)
                    g = np.vdot(u, Hx @ u) - Ex
                    assert abs(np.sum(G * X) - g) < 1e-9, (p, q, r, s)


def test_density_energy_identity_and_blocks():
    """E_corr = sum F D + sum A G (the reference sketch's check,
    ci_wfn.py:411-414), and the block layouts carry the same sum."""
    O, V = 3, 4
    F, A, W = _system(O, V, 8)
    E, c, _ = davidson.davidson(F, W, O, tol=1e-13, rtol=1e-11)
    D = davidson.one_particle_density(c, O)
    G = davidson.two_particle_density(c, O)
    assert abs(np.sum(F * D) + np.sum(A * G) - E) < 1e-12
    Gb, Ab = davidson.tpd_blocks(G, O), davidson.integral_blocks(A, O)
    assert abs(sum(np.sum(Gb[k] * Ab[k]) for k in Gb) - np.sum(A * G)) < 1e-13
    # G is antisymmetric in each index pair
    assert np.abs(G + G.swapaxes(0, 1)).max() < 1e-15 and np.abs(G + G.swapaxes(2, 3)).max() < 1e-15
