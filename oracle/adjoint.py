"""MP2 energy and its derivatives by the adjoint (lambda) method — oracle.

Orbital-invariant spin-orbital MP2: the amplitudes solve the Fock part of
the CID residual (ci_wfn.py:210-212, without the integral terms) driven by
<ab||ij>,
    R(t) = V + P(ab) t F_vv^T - P(ij) F_oo^T t = 0,     V[i,j,a,b] = <ab||ij>
and E = 1/4 sum W t with W = <ij||ab> (mp2_wfn.py:85).  At diagonal F this
is solve_MP2_SO (mp2_wfn.py:64-87) exactly.  Writing R = V + A t, the
adjoint lambda solves A^T lambda = -W/4 (the same operator with F -> F^T),
and
    dE/dW = t/4,   dE/dV = lambda,
    dE/dF_bc = 2 sum_ija lambda_ijab t_ijac,   dE/dF_kj = -2 sum_iab lambda_ijab t_ikab
(W, V treated as independent variables, no conjugation, as the reference
writes the energy).  This is the derivative contracted in analytic.py:320-324
(E_grad = ERI_grad term + f_grad terms) for the spin-orbital case.

Pinned by: central finite differences of `energy` along random directions
(tests/test_oracle_adjoint.py), and at diagonal F by the closed form
dE/dV = W/(4D), dE/dF_aa = -sum W V/(4D^2) ... (canonical MP2).
"""
import numpy as np

from .so import denominators


def fock_apply(F, t, O):
    """A t = P(ab) t F_vv^T - P(ij) F_oo^T t  (ci_wfn.py:211-212)."""
    o, v = slice(0, O), slice(O, F.shape[0])
    Y = np.einsum('ijae,be->ijab', t, F[v, v], optimize=True)
    Z = np.einsum('imab,mj->ijab', t, F[o, o], optimize=True)
    return Y - Y.swapaxes(2, 3) - Z + Z.swapaxes(0, 1)


def solve(F, drv, eps, O, tol=1e-13, maxit=200):
    """x with drv + A x = 0 by Jacobi iteration preconditioned with D."""
    _, D = denominators(eps[:O], eps[O:])
    x = drv / D
    for _ in range(maxit):
        r = drv + fock_apply(F, x, O)
        x = x + r / D
        if np.abs(r).max() < tol:
            break
    return x


def energy(F, W, V, eps, O):
    t = solve(F, V, eps, O)
    return 0.25 * np.sum(W * t), t


def gradients(F, W, V, eps, O):
    """(E, t, lam, dE/dF (n x n, only oo and vv blocks non-zero), dE/dW, dE/dV)."""
    E, t = energy(F, W, V, eps, O)
    lam = solve(F.T, 0.25 * W, eps, O)
    n = F.shape[0]
    G = np.zeros((n, n), dtype=np.complex128)
    G[O:, O:] = 2.0 * np.einsum('ijab,ijac->bc', lam, t, optimize=True)
    G[:O, :O] = -2.0 * np.einsum('ijab,ikab->kj', lam, t, optimize=True)
    return E, t, lam, G, 0.25 * t, lam
