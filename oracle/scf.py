"""Complex RHF of apyib's hf_wfn.py restated in NumPy (oracle only).

Pinned by: the reference's golden RHF energy (test_002_RHF.py:21,
H2O/STO-3G at the geometry of data/molecules.py:2-10) through the integral
engine apyib_amd.gaussian; and by Roothaan's conditions (F C = S C eps, C^H S C
= 1) at convergence.
"""
import numpy as np
import scipy.linalg as la

from .so import solve_general_diis


def contract(expr, *ops):
    return np.einsum(expr, *ops, optimize=True)


def solve_scf(T, V, ERI, S, E_nuc, ndocc, params):
    """hf_wfn.py:36-170.  Returns (E_SCF electronic, C, eps, iterations)."""
    nbf = S.shape[0]
    H_core = T + V
    X = np.linalg.inv(la.sqrtm(S))                                            # :46
    F_p = X @ H_core @ X
    e, C_p = np.linalg.eigh(F_p)
    C = X @ C_p
    D = 2 * contract('mp,np->mn', C[:, :ndocc], np.conjugate(C[:, :ndocc]))   # :58
    E_SCF = contract('mn,mn->', D, H_core)                                    # :62-64
    eps = e
    F_iter = e_iter = None
    it = 0
    for it in range(1, params['max_iterations'] + 1):
        E_old, D_old = E_SCF, D
        F = H_core + 0.5 * contract('ls,mnls->mn', D, 2 * ERI - ERI.swapaxes(1, 2))   # :81
        if params['DIIS']:
            res = (X @ (S @ D @ F - np.conjugate((S @ D @ F).T)) @ X).ravel()   # :86
            Fv = F.ravel()
            if it == 1:
                F_iter = Fv.reshape(-1, 1)
                e_iter = res.reshape(-1, 1)
            Fv, e_iter, F_iter = solve_general_diis(res, Fv, e_iter, F_iter, it)
            F = Fv.reshape(nbf, nbf)
        eps, C_p = np.linalg.eigh(X @ F @ X)                                   # :96-98
        C = X @ C_p
        D = 2 * contract('mp,np->mn', C[:, :ndocc], np.conjugate(C[:, :ndocc]))
        E_SCF = 0.5 * contract('nm,mn->', D, H_core + F)                         # :104-107
        dE = E_SCF - E_old
        rms = np.sqrt(np.sum((D_old - D) ** 2))                                  # :117-121
        if it > 1 and abs(dE) < params['e_convergence'] and rms < params['d_convergence']:
            break
    return E_SCF, C, eps, it
