"""Spin-adapted (closed-shell) MP2 energy derivative by the adjoint form — oracle.

Test infrastructure only: the checker of apy_mp2_sp_* on the device.

Closed-shell MP2 (mp2_wfn.py:42-59): amplitudes t[i,j,a,b] = V[i,j,a,b] / D
with the driver V[i,j,a,b] = <ab|ij> and D = e_i + e_j - e_a - e_b, energy
E = sum (2 W - W^T) t with W[i,j,a,b] = <ij|ab> and ^T the swap of a and b
(plain products, no conjugation, as the reference writes it).  With a
general (non-diagonal) Fock matrix the amplitudes solve
    R(t) = V + A t = 0,
    A t = sum_c F_ac t_ijcb + sum_c F_bc t_ijac - sum_k F_ki t_kjab - sum_k F_kj t_ikab
(the Fock terms of the reference's magnetic-field t2 response,
analytic_aats.py:346-350, with df -> F; analytic.py:326-331 writes the
transposed index order, the same operator for a symmetric F), which is
D-diagonal (A t = -D t) at diagonal F.

Adjoint: with lambda solving A^T lambda = -(2 W - W^T),
    dE = sum (2 dW - dW^T) t + sum lambda (dV + dA t).
At diagonal F, lambda = (2 W - W^T) / D; when W = V (the real-orbital case
the reference's analytic Hessian runs in), lambda = 2 t - t^T = u and, for
dW = dV = ERI_grad, dE is the reference's E_grad (analytic.py:320-324):
    2 u . ERI_grad - t_ijab u_kjab f_ik - t_ijab u_ikab f_jk
                  + t_ijab u_ijcb f_ac + t_ijab u_ijac f_bc.
`densities` gives its coefficients (dE/dERI_grad, dE/df_oo, dE/df_vv);
with the pair symmetry t_ijab = t_jiba the two f_oo (and the two f_vv)
terms are equal, so dE/df_oo = -2 T U^T over (j,a,b) and dE/df_vv =
2 sum_ijb t_ijab u_ijcb.  `t2_response` is analytic.py:326-331 and
`grad_via_response` the equivalent E_grad1 form (analytic.py:333-334).

Pinned by: central finite differences of `energy` (a non-canonical solve)
along random directions at non-diagonal F; the reference's E_grad and
E_grad1 forms agreeing with each other and with the general adjoint at
diagonal F with W = V; the densities against their defining sums
(tests/test_oracle_mp2_sp.py).
"""
import numpy as np


def denominators(eps, o):
    eo, ev = eps[:o], eps[o:]
    return eo.reshape(-1, 1, 1, 1) + eo.reshape(-1, 1, 1) - ev.reshape(-1, 1) - ev


def fock_apply(F, t, o):
    """A t of the closed-shell MP2 residual (the four Fock terms of
    analytic_aats.py:346-350 with df -> F)."""
    oo, vv = F[:o, :o], F[o:, o:]
    out = np.einsum('ac,ijcb->ijab', vv, t) + np.einsum('bc,ijac->ijab', vv, t)
    out -= np.einsum('ki,kjab->ijab', oo, t) + np.einsum('kj,ikab->ijab', oo, t)
    return out


def fock_apply_T(F, x, o):
    """A^T x under the plain bilinear form sum_ijab y A x (F -> F^T)."""
    return fock_apply(F.T, x, o)


def solve(F, V, eps, o, tol=1e-14, maxit=500):
    """t with V + A t = 0 (Jacobi, preconditioned with D)."""
    D = denominators(eps, o)
    t = V / D
    for _ in range(maxit):
        r = V + fock_apply(F, t, o)
        t = t + r / D
        if np.abs(r).max() < tol:
            break
    return t


def energy(F, W, V, eps, o, tol=1e-14):
    """E = sum (2 W - W^T) t (mp2_wfn.py:57) with t from the non-canonical solve."""
    t = solve(F, V, eps, o, tol)
    return np.sum((2 * W - W.swapaxes(2, 3)) * t), t


def adjoint(F, W, V, eps, o, tol=1e-14, maxit=500):
    """(t, lambda): lambda solves A^T lambda = -(2W - W^T)."""
    t = solve(F, V, eps, o, tol)
    D = denominators(eps, o)
    rhs = 2 * W - W.swapaxes(2, 3)
    lam = rhs / D
    for _ in range(maxit):
        r = rhs + fock_apply_T(F, lam, o)
        lam = lam + r / D
        if np.abs(r).max() < tol:
            break
    return t, lam


def directional(F, W, V, eps, o, dF, dW, dV):
    """dE along (dF, dW, dV) by the general adjoint."""
    t, lam = adjoint(F, W, V, eps, o)
    return (np.sum((2 * dW - dW.swapaxes(2, 3)) * t)
            + np.sum(lam * (dV + fock_apply(dF, t, o))))


def reference_grad(t, f_grad, eri_grad, o):
    """The reference's E_grad (analytic.py:320-324), term by term."""
    u = 2 * t - t.swapaxes(2, 3)
    fo, fv = f_grad[:o, :o], f_grad[o:, o:]
    g = np.sum(2 * u * eri_grad)
    g -= np.einsum('ijab,kjab,ik->', t, u, fo)
    g -= np.einsum('ijab,ikab,jk->', t, u, fo)
    g += np.einsum('ijab,ijcb,ac->', t, u, fv)
    g += np.einsum('ijab,ijac,bc->', t, u, fv)
    return g


def t2_response(t, f_grad, eri_grad, eps, o, form=0):
    """t2_grad of analytic.py:326-331 (form 0, the index order also of the
    nuclear first-order amplitudes, analytic_aats.py:446-451), or the
    magnetic-field order of analytic_aats.py:346-350 (form 1, = fock_apply);
    the two differ for a non-symmetric f_grad."""
    fo, fv = f_grad[:o, :o], f_grad[o:, o:]
    if form == 0:
        x = eri_grad - np.einsum('kjab,ik->ijab', t, fo) - np.einsum('ikab,kj->ijab', t, fo)
        x = x + np.einsum('ijcb,ac->ijab', t, fv) + np.einsum('ijac,cb->ijab', t, fv)
    else:
        x = eri_grad + fock_apply(f_grad, t, o)
    return x / denominators(eps, o)


def grad_via_response(t, f_grad, eri_grad, eri_oovv, eps, o):
    """E_grad1 of analytic.py:333-334."""
    tg = t2_response(t, f_grad, eri_grad, eps, o)
    return (np.sum(tg * (2 * eri_oovv - eri_oovv.swapaxes(2, 3)))
            + np.sum(t * (2 * eri_grad - eri_grad.swapaxes(2, 3))))


def densities(t, o):
    """(dE/dERI_grad, dE/df_oo, dE/df_vv) of the reference's E_grad for
    pair-symmetric t: 2u, -2 sum_jab t_ijab u_kjab, 2 sum_ijb t_ijab u_ijcb."""
    u = 2 * t - t.swapaxes(2, 3)
    poo = -2 * np.einsum('ijab,kjab->ik', t, u)
    pvv = 2 * np.einsum('ijab,ijcb->ac', t, u)
    return 2 * u, poo, pvv
