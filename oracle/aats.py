"""Finite-difference atomic axial tensors of apyib's aats.py restated in
NumPy (oracle only): determinants of displaced-orbital overlaps contracted
with CI/MP2 amplitudes.

Spatial form (compute_spatial_aats, aats.py:646-1055):
  * compute_all_dets (aats.py:558-642): determinants of the occupied block of
    a spatial MO overlap after row replacements i -> a (bra) and column
    replacements k -> c (ket), up to two of each, for active i, k.
  * The nine overlap classes uu, up, un, pu, nu, pp, pn, np, nn contribute the
    same pattern of terms (0S, S0, 0D, D0, SS, SD, DS, DD) with class-specific
    bra/ket amplitudes, signs and normalisation factors; `spatial_class_terms`
    writes that pattern once.  As in the reference:
      - 0S and S0 carry the normalisation factor of the other side,
      - the reference-weight terms 0S, S0, 0D, D0 and the singles terms are
        evaluated for CISD only (for MP2 and CID the reference keeps only the
        HF term and DD; aats.py:713-1047 nests 0D/D0 under the CISD branch).
SO form (compute_SO_aats, aats.py:120-554): explicit spin-orbital
determinants (compute_SO_det, row/column swaps on the full overlap), summed
over i<j, a<b (equal to the reference's 1/4 sums over all index orders for
antisymmetric amplitudes) and vectorised over determinant batches.

Pinned by: the reference's golden FD AATs of (H2)2 (test_011_AAT.py:56-240,
test_013_AAT_parallel.py:37-189, 322-384, 393-772; MP2/CID/CISD, spatial and
SO, intermediate and full normalisation) through oracle/fd.py in
tests/test_oracle_aat.py, and compute_all_dets against its definition.
"""
import itertools

import numpy as np


def contract(expr, *ops):
    return np.einsum(expr, *ops, optimize=True)


def det_replaced(S, no, rows, cols):
    """det of S[occ, occ] after replacing occupied row i by row no + a and
    occupied column k by column no + c (the matrix the reference builds by
    swapping rows/columns i <-> no + a)."""
    r = list(range(no))
    c = list(range(no))
    for i, a in rows:
        r[i] = no + a
    for k, cc in cols:
        c[k] = no + cc
    return np.linalg.det(S[np.ix_(r, c)])


def compute_all_dets(S, no, nv, nf=0):
    """aats.py:558-642.  Returns, in the reference's order, det_S, ia_S, S_kc,
    iajb_S, S_kcld, ia_S_kc, iajb_S_kc, ia_S_kcld, iajb_S_kcld, with active
    occupied indices counted from nf (ia_S_kcld[i,a,k,c,l,d]: row i -> a,
    columns k -> c and l -> d)."""
    o = range(nf, no)
    na = no - nf
    sh1 = (na, nv)
    ia = np.zeros(sh1, complex)
    kc = np.zeros(sh1, complex)
    iajb = np.zeros(sh1 * 2, complex)
    kcld = np.zeros(sh1 * 2, complex)
    iakc = np.zeros(sh1 * 2, complex)
    iajbkc = np.zeros(sh1 * 3, complex)
    iakcld = np.zeros(sh1 * 3, complex)
    iajbkcld = np.zeros(sh1 * 4, complex)
    d0 = np.linalg.det(S[:no, :no])
    singles = list(itertools.product(o, range(nv)))
    pairs = [(i, a, j, b) for i, a in singles for j, b in singles if i != j and a != b]
    for i, a in singles:
        ia[i - nf, a] = det_replaced(S, no, [(i, a)], [])
        kc[i - nf, a] = det_replaced(S, no, [], [(i, a)])
        for k, c in singles:
            iakc[i - nf, a, k - nf, c] = det_replaced(S, no, [(i, a)], [(k, c)])
    for i, a, j, b in pairs:
        I, J = i - nf, j - nf
        iajb[I, a, J, b] = det_replaced(S, no, [(i, a), (j, b)], [])
        kcld[I, a, J, b] = det_replaced(S, no, [], [(i, a), (j, b)])
        for k, c in singles:
            iajbkc[I, a, J, b, k - nf, c] = det_replaced(S, no, [(i, a), (j, b)], [(k, c)])
            iakcld[k - nf, c, I, a, J, b] = det_replaced(S, no, [(k, c)], [(i, a), (j, b)])
        for k, c, l, d in pairs:
            iajbkcld[I, a, J, b, k - nf, c, l - nf, d] = det_replaced(S, no, [(i, a), (j, b)], [(k, c), (l, d)])
    return d0, ia, kc, iajb, kcld, iakc, iajbkc, iakcld, iajbkcld


def spatial_class_terms(dets, bra, ket, n_bra, n_ket):
    """The terms one overlap class adds (aats.py:713-1047).  bra / ket =
    (t1, t2) or None where that side is the reference determinant; n_bra /
    n_ket multiply the 0S / S0 terms."""
    S, ia, kc, iajb, kcld, iakc, iajbkc, iakcld, iajbkcld = dets
    out = {k: 0 for k in ('0S', '0D', 'S0', 'D0', 'SS', 'SD', 'DS', 'DD')}
    asym = lambda t: t - np.swapaxes(t, 2, 3)
    if bra is None:
        K1, K2 = ket
        out['0S'] = 2 * contract('kc,kc->', K1, kc) * S * n_bra
        out['0D'] = (0.5 * contract('klcd,kcld->', asym(K2), kcld) * S
                     + contract('ld,ld->', contract('klcd,kc->ld', K2, kc), kc))
        return out
    B1, B2 = bra
    if ket is None:
        out['S0'] = 2 * contract('ia,ia->', B1, ia) * S * n_ket
        out['D0'] = (0.5 * contract('ijab,iajb->', asym(B2), iajb) * S
                     + contract('jb,jb->', contract('ijab,ia->jb', B2, ia), ia))
        return out
    K1, K2 = ket
    out['SS'] = 2 * contract('ia,kc,iakc->', B1, K1, iakc) * S + 2 * contract('ia,ia->', B1, ia) * contract(
        'kc,kc->', K1, kc)
    out['DS'] = (0.5 * contract('ijab,kc,iajbkc->', asym(B2), K1, iajbkc) * S
                 + 0.5 * contract('ijab,iajb->', asym(B2), iajb) * contract('kc,kc->', K1, kc)
                 + 2 * contract('ijab,kc,iakc,jb->', B2, K1, iakc, ia))
    out['SD'] = (0.5 * contract('ia,klcd,iakcld->', B1, asym(K2), iakcld) * S
                 + 0.5 * contract('ia,ia->', B1, ia) * contract('klcd,kcld->', asym(K2), kcld)
                 + 2 * contract('ia,klcd,iakc,ld->', B1, K2, iakc, kc))
    out['DD'] = (0.125 * contract('ijab,klcd,iajbkcld->', asym(B2), asym(K2), iajbkcld) * S
                 + 0.125 * contract('ijab,iajb->', asym(B2), iajb) * contract('klcd,kcld->', asym(K2), kcld)
                 + 0.5 * contract('ijab,klcd,iajbkc,ld->', asym(B2), K2, iajbkc, kc)
                 + 0.25 * contract('ijab,klcd,iakcld,jb->', B2, asym(K2), iakcld, ia)
                 + 0.25 * contract('ijab,klcd,ia,jbkcld->', B2, asym(K2), ia, iakcld)
                 + 1.0 * contract('ijab,klcd,iakc,jbld->', B2, K2, iakc, iakc))
    return out


def _spatial_norm(T, singles):
    """aats.py:652-669: 1/sqrt(t0 + 2|t1|^2 + 2 t2*.t2 - t2*.t2^T) (t0 enters
    unsquared)."""
    t0, t1, t2 = T
    x = t0 + (2 * contract('ijab,ijab->', np.conjugate(t2), t2) - contract('ijab,ijba->', np.conjugate(t2), t2))
    if singles:
        x = x + 2 * contract('ia,ia->', np.conjugate(t1), t1)
    return 1 / np.sqrt(x)


def _so_norm(T, singles):
    """aats.py:134-157: 1/sqrt(t0^2 + |t1|^2 + 1/4 |t2|^2)."""
    t0, t1, t2 = T
    x = t0 ** 2 + 0.25 * contract('ijab,ijab->', np.conjugate(t2), t2)
    if singles:
        x = x + contract('ia,ia->', np.conjugate(t1), t1)
    return 1 / np.sqrt(x)


def normalizations(method, T, alpha, beta, normalization):
    """(N, N_np, N_nn, N_mp, N_mn) of one AAT element."""
    sets = (T['u'], T['np'][alpha], T['nn'][alpha], T['mp'][beta], T['mn'][beta])
    if method == 'RHF' or normalization == 'intermediate':
        return (1.0,) * 5
    if method.endswith('_SO'):
        return tuple(_so_norm(t, method == 'CISD_SO') for t in sets)
    return tuple(_spatial_norm(t, method == 'CISD') for t in sets)


def spatial_aat(method, no, nv, nf, overlaps, T, dR, dH, alpha, beta, normalization='full', dets_cache=None,
                split=False):
    """One AAT element of compute_spatial_aats (aats.py:646-1055).

    overlaps: {'uu': S, 'up': [S]*3, 'un', 'pu': [S]*3N, 'nu',
    'pp': [[S]*3]*3N, 'pn', 'np', 'nn'} spatial MO overlaps; T: {'u': [t0,t1,t2],
    'np': [...]*3N, 'nn', 'mp': [...]*3, 'mn'}.  dets_cache (dict) reuses
    compute_all_dets of shared overlaps across elements.  split=True returns
    {'00', '0S', 'S0', '0D', 'D0', 'SS', 'SD', 'DS', 'DD'} -> the reference's
    I_00, I_0S, ... (aats.py:677-1047) instead of their sum."""
    N, N_np, N_nn, N_mp, N_mn = normalizations(method, T, alpha, beta, normalization)
    d2 = lambda S: np.linalg.det(S[:no, :no]) ** 2
    parts = {k: 0.0 for k in SPLIT_TERMS}
    parts['00'] = (d2(overlaps['pp'][alpha][beta]) * N_np * N_mp - d2(overlaps['pn'][alpha][beta]) * N_np * N_mn
                   - d2(overlaps['np'][alpha][beta]) * N_nn * N_mp + d2(overlaps['nn'][alpha][beta]) * N_nn * N_mn)
    if method != 'RHF':
        cisd = method == 'CISD'
        na = no - nf
        z1 = np.zeros((na, nv), complex)
        u, tnp, tnn, tmp, tmn = T['u'], T['np'][alpha], T['nn'][alpha], T['mp'][beta], T['mn'][beta]
        s1 = lambda t: t if cisd else z1
        t1, t2 = N * s1(u[1]), N * u[2]
        dHk = (N_mp * s1(tmp[1]) - N_mn * s1(tmn[1]), N_mp * tmp[2] - N_mn * tmn[2])
        dRc = (np.conjugate(N_np * s1(tnp[1]) - N_nn * s1(tnn[1])), np.conjugate(N_np * tnp[2] - N_nn * tnn[2]))
        tk = (t1, t2)
        tb = (np.conjugate(t1), np.conjugate(t2))
        classes = [
            (('uu',), dRc, dHk, 1, 1, +1, None),
            (('up', beta), dRc, tk, 1, N_mp, +1, 'S0D0'),
            (('un', beta), dRc, tk, 1, N_mn, -1, 'S0D0'),
            (('pu', alpha), tb, dHk, N_np, 1, +1, '0S0D'),
            (('nu', alpha), tb, dHk, N_nn, 1, -1, '0S0D'),
            (('pp', alpha, beta), tb, tk, N_np, N_mp, +1, 'all'),
            (('pn', alpha, beta), tb, tk, N_np, N_mn, -1, 'all'),
            (('np', alpha, beta), tb, tk, N_nn, N_mp, -1, 'all'),
            (('nn', alpha, beta), tb, tk, N_nn, N_mn, +1, 'all'),
        ]
        for key, bra, ket, nb, nk, sign, ref in classes:
            S = overlaps[key[0]]
            for idx in key[1:]:
                S = S[idx]
            if dets_cache is not None and key in dets_cache:
                dets = dets_cache[key]
            else:
                dets = compute_all_dets(S, no, nv, nf)
                if dets_cache is not None:
                    dets_cache[key] = dets
            both = spatial_class_terms(dets, bra, ket, nb, nk)
            terms = ['DD']
            if cisd:
                terms += ['SS', 'SD', 'DS']
                if ref in ('S0D0', 'all'):
                    terms += ['S0', 'D0']
                    both.update({k: v for k, v in spatial_class_terms(dets, bra, None, nb, nk).items()
                                 if k in ('S0', 'D0')})
                if ref in ('0S0D', 'all'):
                    terms += ['0S', '0D']
                    both.update({k: v for k, v in spatial_class_terms(dets, None, ket, nb, nk).items()
                                 if k in ('0S', '0D')})
            for k in terms:
                parts[k] = parts[k] + sign * both[k]
    return _finish(parts, dR, dH, split)


SPLIT_TERMS = ('00', '0S', 'S0', '0D', 'D0', 'SS', 'SD', 'DS', 'DD')


def _finish(parts, dR, dH, split):
    f = 1 / (4 * dR * dH)
    if split:
        return {k: f * np.imag(v) for k, v in parts.items()}
    return f * np.imag(sum(parts[k] for k in SPLIT_TERMS))


# ----------------------------------------------------------------------------------------------------------
# Spin-orbital form

def _so_states(nocc, nvir, singles):
    """Bra/ket configurations: reference, singles (i, a) and doubles
    (i<j, a<b); a, b are virtual indices counted from nocc."""
    st = [()]
    if singles:
        st += [(i, a) for i in range(nocc) for a in range(nvir)]
    st += [(i, a, j, b) for i in range(nocc) for j in range(i + 1, nocc)
           for a in range(nvir) for b in range(a + 1, nvir)]
    return st


def _so_index_lists(states, nocc):
    """compute_SO_det's swap sequence (aats.py:120-130) as index lists: the
    occupied positions i, j take rows/columns nocc + a, nocc + b (i != j and
    a != b, so the swaps are replacements)."""
    out = np.tile(np.arange(nocc), (len(states), 1))
    for s, st in enumerate(states):
        for x in range(0, len(st), 2):
            out[s, st[x]] = nocc + st[x + 1]
    return out


def so_det_matrix(S, nocc, states):
    """D[B, K] = compute_SO_det(S, B, K) for every bra/ket configuration."""
    R = _so_index_lists(states, nocc)
    M = S[R[:, None, :, None], R[None, :, None, :]]
    return np.linalg.det(M)


def _so_coeffs(states, c0, t1, t2):
    out = np.zeros(len(states), complex)
    for s, st in enumerate(states):
        if len(st) == 0:
            out[s] = c0
        elif len(st) == 2:
            out[s] = t1[st[0], st[1]]
        else:
            out[s] = t2[st[0], st[2], st[1], st[3]]
    return out


def so_aat(method, nocc, nvir, overlaps, T, dR, dH, alpha, beta, normalization='full', det_cache=None,
           split=False):
    """One AAT element of compute_SO_aats (aats.py:520-554) from the SO
    overlaps (the reference's compute_so_overlap of each MO overlap).
    split=True returns the bra-segment x ket-segment parts {'00', '0S', ...}
    -> compute_SO_I_00, compute_SO_I_0S, ... (aats.py:161-516)."""
    N, N_np, N_nn, N_mp, N_mn = normalizations(method, T, alpha, beta, normalization)
    singles = method == 'CISD_SO'
    states = _so_states(nocc, nvir, singles) if method != 'RHF' else [()]
    z1 = np.zeros((nocc, nvir), complex)
    z2 = np.zeros((nocc, nocc, nvir, nvir), complex)
    amp = lambda t, k: (t[k] if (k == 2 or singles) and not np.isscalar(t[k]) else (z1 if k == 1 else z2))
    u, tnp, tnn, tmp, tmn = T['u'], T['np'][alpha], T['nn'][alpha], T['mp'][beta], T['mn'][beta]
    t = _so_coeffs(states, 1.0, amp(u, 1), amp(u, 2))
    dH_ = _so_coeffs(states, 0.0, amp(tmp, 1) - amp(tmn, 1), amp(tmp, 2) - amp(tmn, 2))
    dR_ = np.conjugate(_so_coeffs(states, 0.0, amp(tnp, 1) - amp(tnn, 1), amp(tnp, 2) - amp(tnn, 2)))
    tc = np.conjugate(t)

    def D(key):
        S = overlaps[key[0]]
        for idx in key[1:]:
            S = S[idx]
        if det_cache is not None and key in det_cache:
            return det_cache[key]
        val = so_det_matrix(S, nocc, states)
        if det_cache is not None:
            det_cache[key] = val
        return val

    seg = {'0': slice(0, 1), 'S': slice(1, 1 + (nocc * nvir if singles else 0)),
           'D': slice(1 + (nocc * nvir if singles else 0), len(states))}
    parts = {k: 0.0 for k in SPLIT_TERMS}

    def add(bra, M, ket):
        for k in SPLIT_TERMS:
            b, c = seg[k[0]], seg[k[1]]
            parts[k] = parts[k] + bra[b] @ M[b, c] @ ket[c]

    add(tc, D(('pp', alpha, beta)) * N_np * N_mp - D(('pn', alpha, beta)) * N_np * N_mn
        - D(('np', alpha, beta)) * N_nn * N_mp + D(('nn', alpha, beta)) * N_nn * N_mn, t)
    if method != 'RHF':
        add(dR_, D(('uu',)) * N * N, dH_)
        add(dR_, D(('up', beta)) * N * N_mp - D(('un', beta)) * N * N_mn, t)
        add(tc, D(('pu', alpha)) * N_np * N - D(('nu', alpha)) * N_nn * N, dH_)
    return _finish(parts, dR, dH, split)
