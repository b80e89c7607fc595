"""AO->MO->spin-orbital transforms, overlaps and phase of apyib's utils.py
restated in NumPy (oracle only).

Pinned by (tests/test_oracle_golden.py unless noted): the loop forms of
compute_F_SO / compute_ERI_SO (statements of utils.py:283-365 with explicit
spin integration) against the vectorized forms used elsewhere;
compute_ERI_MO against a direct four-index sum on small inputs;
compute_phase against its defining property (diagonal of the overlap made
real positive); compute_so_overlap through the spin-orbital determinants,
which must factor into products of the spatial alpha and beta ones
(tests/test_oracle_aat.py); all of them through the reference's golden
energies and AAT arrays they feed.
"""
import numpy as np


def contract(expr, *ops):
    return np.einsum(expr, *ops, optimize=True)


def get_slices(nbf, ndocc, nfzc, method):
    """Slices of utils.py:184-213."""
    f, o, v, t = slice(0, nfzc), slice(nfzc, ndocc), slice(ndocc, nbf), slice(nfzc, nbf)
    if method in ('RHF', 'MP2', 'CID', 'CISD'):
        I = [slice(0, nfzc), slice(0, ndocc - nfzc), slice(ndocc - nfzc, nbf - nfzc), slice(0, nbf - nfzc)]
    elif method in ('MP2_SO', 'CID_SO', 'CISD_SO'):
        I = [slice(0, 2 * nfzc), slice(0, 2 * ndocc - 2 * nfzc), slice(2 * ndocc - 2 * nfzc, 2 * nbf - 2 * nfzc),
             slice(0, 2 * nbf - 2 * nfzc)]
    else:
        raise ValueError(method)
    return [f, o, v, t], I


def compute_F_MO(h_AO, ERI_AO, C, C_list, freeze_core):
    """utils.py:217-254: frozen-core operator/energy, AO Fock, C^H F C."""
    f, o, v, t = C_list
    E_fc = 0
    if freeze_core:
        C_fc = C[:, f]
        D_fc = contract('mp,np->mn', C_fc, np.conjugate(C_fc))
        h_fc = h_AO + contract('ls,mnls->mn', D_fc, 2 * ERI_AO - ERI_AO.swapaxes(1, 2))
        E_fc = contract('nm,mn->', D_fc, h_AO + h_fc)
        h_AO = h_fc
    D = contract('mp,np->mn', C[:, o], np.conjugate(C[:, o]))
    F_AO = h_AO + contract('ls,mnls->mn', D, 2 * ERI_AO - ERI_AO.swapaxes(1, 2))
    F_MO = contract('ip,ij,jq->pq', np.conjugate(C[:, t]), F_AO, C[:, t])
    return F_MO, E_fc


def compute_ERI_MO(ERI_AO, C, C_list):
    """utils.py:258-279: four quarter transforms, bra-conjugated on p and r."""
    Ct = C[:, C_list[3]]
    X = contract('mnlg,gs->mnls', ERI_AO, Ct)
    X = contract('mnls,lr->mnrs', X, np.conjugate(Ct))
    X = contract('nq,mnrs->mqrs', Ct, X)
    return contract('mp,mqrs->pqrs', np.conjugate(Ct), X)


def _spin(p):
    return 1 if p % 2 == 0 else -1


def compute_F_SO_loop(F_MO):
    """utils.py:283-313 as written (explicit spin integration)."""
    n = 2 * F_MO.shape[0]
    F = np.zeros((n, n), dtype=F_MO.dtype)
    for p in range(n):
        for q in range(n):
            F[p, q] = F_MO[p // 2, q // 2] * max(_spin(p) * _spin(q), 0)
    return F


def compute_ERI_SO_loop(ERI_MO, p_range=None):
    """utils.py:317-365 as written (pure-Python four-index loop).  p_range
    (optional) restricts the outer index to a subsample; the other rows of
    the result stay zero."""
    n = [2 * s for s in ERI_MO.shape]
    out = np.zeros(n, dtype=ERI_MO.dtype)
    for p in (range(n[0]) if p_range is None else p_range):
        for q in range(n[1]):
            spq = max(_spin(p) * _spin(q), 0)
            for r in range(n[2]):
                for s in range(n[3]):
                    out[p, q, r, s] = ERI_MO[p // 2, q // 2, r // 2, s // 2] * spq * max(_spin(r) * _spin(s), 0)
    return out


def spin_delta(n):
    p = np.arange(2 * n)
    return (p[:, None] % 2 == p[None, :] % 2).astype(float)


def compute_F_SO(F_MO):
    """Vectorized compute_F_SO: F_SO = kron(F_MO, 1_2) masked by spin."""
    m = F_MO.shape[0]
    return np.repeat(np.repeat(F_MO, 2, axis=0), 2, axis=1) * spin_delta(m)


def compute_ERI_SO(ERI_MO):
    """Vectorized compute_ERI_SO (chemists' (pq|rs) in spin orbitals)."""
    m = ERI_MO.shape[0]
    X = ERI_MO
    for ax in range(4):
        X = np.repeat(X, 2, axis=ax)
    d = spin_delta(m)
    return X * d[:, :, None, None] * d[None, None, :, :]


def compute_mo_overlap(S_AO, C_bra, C_ket):
    """utils.py:370-388: C_bra^H S_AO C_ket (S_AO supplied by the caller)."""
    return contract('mp,mn,nq->pq', np.conjugate(C_bra), S_AO, C_ket)


def compute_so_overlap(S_MO):
    """utils.py:393-422 (vectorized spin-block expansion)."""
    return compute_F_SO(S_MO)


def compute_so_overlap_loop(S_MO):
    return compute_F_SO_loop(S_MO)


def compute_phase(S_AO, C0, C):
    """utils.py:427-446: divide column m of C by S_mm / sqrt(S_mm conj(S_mm)),
    S = C0^H S_AO C."""
    S1 = compute_mo_overlap(S_AO, C0, C)
    S2 = np.conjugate(S1.T)
    d = np.diagonal(S1)
    N = np.sqrt(d * np.diagonal(S2))
    phase = d / N
    return C * (phase ** -1)[None, :]
