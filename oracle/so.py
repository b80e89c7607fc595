"""Spin-orbital CISD / CID / MP2 of apyib restated in NumPy (oracle only).

Input convention is the reference's after ``ERI_SO.swapaxes(1,2)``
(ci_wfn.py:284): ``eri`` is the Dirac tensor <pq|rs> over all n spin
orbitals, occupied first, NOT antisymmetrized.  Antisymmetrized blocks are
formed per term as ``eri[blk] - eri.swapaxes(2,3)[blk]`` exactly as the
reference writes them.  A caller holding an antisymmetrized tensor
A = <pq||rs> passes ``eri = A/2``, for which that difference is A exactly.

``opt_einsum.contract`` (the reference's contraction engine) is replaced by
``numpy.einsum(optimize=True)``; both evaluate the same sums.

Parity pins (tests/test_oracle_so.py):
  cisd_so_residual / cid_so_residual: per-sigma against the determinant-
      basis Hamiltonian of oracle.brute on random amplitudes (every term);
  solve_cisd_so / solve_cid_so: converged energy and amplitudes against the
      lowest eigenpair of the brute-force CISD / CID matrix;
  solve_mp2_so: against second-order Rayleigh-Schroedinger energy built by
      brute force from determinants;
  solve_general_diis: against a direct least-squares statement of DIIS
      (minimise |sum c_i e_i| subject to sum c_i = 1);
  cisd_so_sigma_blocks: against cisd_so_residual;
  all of them, end to end, against the reference's golden energies
      (H2O/STO-3G MP2 and CID, (H2)2/STO-3G CISD) through oracle.integrals.
"""
import numpy as np


def contract(expr, *ops):
    return np.einsum(expr, *ops, optimize=True)


def asym(eri, s0, s1, s2, s3):
    """<pq||rs> block as the reference forms it: eri[..] - eri.swapaxes(2,3)[..]."""
    return eri[s0, s1, s2, s3] - eri.swapaxes(2, 3)[s0, s1, s2, s3]


def _slices(O, n):
    return slice(0, O), slice(O, n)


def driver_t2(eri, O):
    """<ab||ij> laid out [i,j,a,b] (ci_wfn.py:319, mp2_wfn.py:82)."""
    n = eri.shape[0]
    o, v = _slices(O, n)
    X = eri.swapaxes(0, 2).swapaxes(1, 3)
    return X[o, o, v, v] - X.swapaxes(2, 3)[o, o, v, v]


def denominators(eps_o, eps_v):
    """D_ia and D_ijab (ci_wfn.py:270-271) from spin-orbital energies."""
    D_ia = eps_o.reshape(-1, 1) - eps_v
    D_ijab = eps_o.reshape(-1, 1, 1, 1) + eps_o.reshape(-1, 1, 1) - eps_v.reshape(-1, 1) - eps_v
    return D_ia, D_ijab


def cisd_so_energy(F, eri, t1, t2, O):
    """E = sum t1*F_ov + 1/4 sum t2*<ij||ab>, no conjugation (ci_wfn.py:355)."""
    o, v = _slices(O, F.shape[0])
    return contract('ia,ia->', t1, F[o, v]) + 0.25 * contract('ijab,ijab->', t2, asym(eri, o, o, v, v))


def cisd_so_residual(F, eri, t1, t2, E, O):
    """Residuals r_T1, r_T2 of ci_wfn.py:309-334 (sigma minus E*t)."""
    o, v = _slices(O, F.shape[0])
    dt = np.result_type(F, eri, t1, t2)
    r1 = F.swapaxes(0, 1)[o, v].astype(dt)                                      # :309
    r1 -= contract('ji,ja->ia', F[o, o], t1)                                    # :310
    r1 += contract('ab,ib->ia', F[v, v], t1)                                    # :311
    r1 += contract('jabi,jb->ia', asym(eri, o, v, v, o), t1)                    # :312
    r1 += contract('jb,ijab->ia', F[o, v], t2)                                  # :313
    r1 += 0.5 * contract('ajcb,ijcb->ia', asym(eri, v, o, v, v), t2)            # :314
    r1 -= 0.5 * contract('kjib,kjab->ia', asym(eri, o, o, o, v), t2)            # :315
    r1 -= E * t1                                                                # :316

    r2 = driver_t2(eri, O).astype(dt)                                           # :319
    r2 -= contract('kbij,ka->ijab', asym(eri, o, v, o, o), t1)                  # :320
    r2 -= contract('akij,kb->ijab', asym(eri, v, o, o, o), t1)                  # :321
    r2 += contract('abcj,ic->ijab', asym(eri, v, v, v, o), t1)                  # :322
    r2 += contract('abic,jc->ijab', asym(eri, v, v, o, v), t1)                  # :323
    r2 += contract('bc,ijac->ijab', F[v, v], t2)                                # :324
    r2 += contract('ac,ijcb->ijab', F[v, v], t2)                                # :325
    r2 -= contract('kj,ikab->ijab', F[o, o], t2)                                # :326
    r2 -= contract('ki,kjab->ijab', F[o, o], t2)                                # :327
    r2 += 0.5 * contract('klij,klab->ijab', asym(eri, o, o, o, o), t2)          # :328
    r2 += 0.5 * contract('abcd,ijcd->ijab', asym(eri, v, v, v, v), t2)          # :329
    W = asym(eri, o, v, v, o)
    r2 += contract('kbcj,ikac->ijab', W, t2)                                    # :330
    r2 += contract('kbci,kjac->ijab', W, t2)                                    # :331
    r2 += contract('kacj,ikcb->ijab', W, t2)                                    # :332
    r2 += contract('kaci,kjcb->ijab', W, t2)                                    # :333
    r2 -= E * t2                                                                # :334
    return r1, r2


def cid_so_energy(eri, t2, O):
    o, v = _slices(O, eri.shape[0])
    return 0.25 * contract('ijab,ijab->', asym(eri, o, o, v, v), t2)            # ci_wfn.py:237


def cid_so_residual(F, eri, t2, E, O):
    """r_T2 of ci_wfn.py:210-220."""
    o, v = _slices(O, F.shape[0])
    r2 = driver_t2(eri, O).astype(np.result_type(F, eri, t2))                   # :210
    r2 += contract('ijae,be->ijab', t2, F[v, v]) + contract('ijeb,ae->ijab', t2, F[v, v])   # :211
    r2 -= contract('imab,mj->ijab', t2, F[o, o]) + contract('mjab,mi->ijab', t2, F[o, o])   # :212
    r2 += 0.5 * contract('mnab,mnij->ijab', t2, asym(eri, o, o, o, o))          # :213
    r2 += 0.5 * contract('ijef,abef->ijab', t2, asym(eri, v, v, v, v))          # :214
    W = asym(eri, o, v, v, o)
    r2 += contract('imae,mbej->ijab', t2, W)                                    # :215
    r2 += contract('mjae,mbei->ijab', t2, W)                                    # :216
    r2 += contract('imeb,maej->ijab', t2, W)                                    # :217
    r2 += contract('mjeb,maei->ijab', t2, W)                                    # :218
    r2 -= E * t2                                                                # :220
    return r2


def solve_general_diis(res_vec, t_vec, e_iter, t_iter, iteration, max_diis=7):
    """DIIS extrapolation of utils.py:104-140.

    e_iter / t_iter are (length, nvec) arrays.  The oldest columns are
    dropped while more than max_diis are stored (:109-112), the new pair is
    appended from the second iteration on (:117-119), B = E^H E is bordered
    with -1 (:121-128) and solved against (0,..,0,-1) (:131-135).
    """
    while e_iter.shape[1] > max_diis:
        e_iter = e_iter[:, 1:]
        t_iter = t_iter[:, 1:]
    if iteration != 1:
        e_iter = np.hstack((e_iter, res_vec.reshape(-1, 1)))
        t_iter = np.hstack((t_iter, t_vec.reshape(-1, 1)))
    nv = e_iter.shape[1]
    B = np.zeros((nv + 1, nv + 1), dtype=np.result_type(e_iter.dtype, np.float64))
    B[:nv, :nv] = contract('om,on->mn', np.conjugate(e_iter), e_iter)
    B[-1, :] = -1
    B[:, -1] = -1
    B[-1, -1] = 0
    A = np.zeros(nv + 1)
    A[-1] = -1
    c = np.linalg.solve(B, A)
    return t_iter @ c[:-1], e_iter, t_iter


def _converged(iteration, dE, rms_list, params):
    """Convergence test of ci_wfn.py:371-374: |dE| < e_conv and every rms
    (sqrt of an un-conjugated sum of squares, a complex number) compared to
    d_convergence in NumPy's complex ordering (real part first)."""
    if iteration <= 1:
        return False
    if not abs(dE) < params['e_convergence']:
        return False
    return all(np.complex128(r) < params['d_convergence'] for r in rms_list)


def solve_cisd_so(F, eri, eps_o, eps_v, O, params, t1=None, t2=None):
    """solve_CISD_SO of ci_wfn.py:263-416 (Jacobi + DIIS).

    Returns (E_CISD, t1, t2, iterations performed)."""
    o, v = _slices(O, F.shape[0])
    D_ia, D_ijab = denominators(eps_o, eps_v)
    if t1 is None:
        t1 = F.swapaxes(0, 1)[o, v] / D_ia                                      # :287
    if t2 is None:
        t2 = driver_t2(eri, O) / D_ijab                                         # :290
    t1 = np.array(t1, dtype=np.complex128)
    t2 = np.array(t2, dtype=np.complex128)
    E = cisd_so_energy(F, eri, t1, t2, O)                                       # :293
    e_iter = t_iter = None
    it = 0
    for it in range(1, params['max_iterations'] + 1):
        E_old, t1_old, t2_old = E, t1.copy(), t2.copy()
        r1, r2 = cisd_so_residual(F, eri, t1, t2, E, O)
        t1 = t1 + r1 / D_ia                                                     # :336
        t2 = t2 + r2 / D_ijab                                                   # :337
        if params['DIIS']:
            res_vec = np.concatenate((r1.ravel(), r2.ravel()))
            t_vec = np.concatenate((t1.ravel(), t2.ravel()))
            if it == 1:
                t_iter = t_vec.reshape(-1, 1)
                e_iter = res_vec.reshape(-1, 1)
            t_vec, e_iter, t_iter = solve_general_diis(res_vec, t_vec, e_iter, t_iter, it)
            t1 = t_vec[:t1.size].reshape(t1.shape)
            t2 = t_vec[t1.size:].reshape(t2.shape)
        E = cisd_so_energy(F, eri, t1, t2, O)                                   # :355
        d1 = t1_old - t1
        d2 = t2_old - t2
        rms1 = np.sqrt(contract('ia,ia->', d1, d1))                             # :361-362
        rms2 = np.sqrt(contract('ijab,ijab->', d2, d2))                         # :364-365
        if _converged(it, E_old - E, [rms1, rms2], params):
            break
    return E, t1, t2, it


def solve_cid_so(F, eri, eps_o, eps_v, O, params):
    """solve_CID_SO of ci_wfn.py:171-259.  Returns (E_CID, t2, iterations)."""
    _, D_ijab = denominators(eps_o, eps_v)
    t2 = driver_t2(eri, O) / D_ijab                                             # :194
    E = cid_so_energy(eri, t2, O)                                               # :197
    e_iter = t_iter = None
    it = 0
    for it in range(1, params['max_iterations'] + 1):
        E_old, t2_old = E, t2.copy()
        r2 = cid_so_residual(F, eri, t2, E, O)
        t2 = t2 + r2 / D_ijab                                                   # :221
        if params['DIIS']:
            res_vec = r2.ravel()
            t_vec = t2.ravel()
            if it == 1:
                t_iter = t_vec.reshape(-1, 1)
                e_iter = res_vec.reshape(-1, 1)
            t_vec, e_iter, t_iter = solve_general_diis(res_vec, t_vec, e_iter, t_iter, it)
            t2 = t_vec.reshape(t2.shape)
        E = cid_so_energy(eri, t2, O)                                           # :237
        d2 = t2_old - t2
        rms2 = np.sqrt(contract('ijab,ijab->', d2, d2))                         # :243-244
        if _converged(it, E_old - E, [rms2], params):
            break
    return E, t2, it


def solve_mp2_so(eri, eps_o, eps_v, O):
    """solve_MP2_SO of mp2_wfn.py:64-87: t2 = <ab||ij>/D, E = 1/4 sum <ij||ab> t2."""
    _, D_ijab = denominators(eps_o, eps_v)
    o, v = _slices(O, eri.shape[0])
    t2 = driver_t2(eri, O) / D_ijab                                             # :82
    E = 0.25 * contract('ijab,ijab->', asym(eri, o, o, v, v), t2)               # :85
    return E, t2


# ---------------------------------------------------------------------------
# Block form: the same sigma from precomputed antisymmetrized blocks, for
# sizes where the full n^4 tensor does not fit (CPU baseline at O=20, V=160).
# ---------------------------------------------------------------------------

BLOCK_NAMES = ('oooo', 'ooov', 'oovv', 'ovoo', 'ovvo', 'vooo', 'vovv', 'vvoo', 'vvov', 'vvvo', 'vvvv')


def blocks_from_eri(eri, O):
    """Antisymmetrized Dirac blocks <pq||rs> keyed by o/v pattern."""
    o, v = _slices(O, eri.shape[0])
    sl = {'o': o, 'v': v}
    return {k: asym(eri, sl[k[0]], sl[k[1]], sl[k[2]], sl[k[3]]) for k in BLOCK_NAMES}


def cisd_so_sigma_blocks(F, W, t1, t2, O):
    """r_T1 + E t1 and r_T2 + E t2 of cisd_so_residual from blocks W (the
    E-independent part, i.e. the CISD sigma with unit reference weight)."""
    o, v = _slices(O, F.shape[0])
    s1 = F.swapaxes(0, 1)[o, v].copy()
    s1 -= contract('ji,ja->ia', F[o, o], t1)
    s1 += contract('ab,ib->ia', F[v, v], t1)
    s1 += contract('jabi,jb->ia', W['ovvo'], t1)
    s1 += contract('jb,ijab->ia', F[o, v], t2)
    s1 += 0.5 * contract('ajcb,ijcb->ia', W['vovv'], t2)
    s1 -= 0.5 * contract('kjib,kjab->ia', W['ooov'], t2)
    s2 = W['vvoo'].transpose(2, 3, 0, 1).copy()
    s2 -= contract('kbij,ka->ijab', W['ovoo'], t1)
    s2 -= contract('akij,kb->ijab', W['vooo'], t1)
    s2 += contract('abcj,ic->ijab', W['vvvo'], t1)
    s2 += contract('abic,jc->ijab', W['vvov'], t1)
    s2 += contract('bc,ijac->ijab', F[v, v], t2)
    s2 += contract('ac,ijcb->ijab', F[v, v], t2)
    s2 -= contract('kj,ikab->ijab', F[o, o], t2)
    s2 -= contract('ki,kjab->ijab', F[o, o], t2)
    s2 += 0.5 * contract('klij,klab->ijab', W['oooo'], t2)
    s2 += 0.5 * contract('abcd,ijcd->ijab', W['vvvv'], t2)
    s2 += contract('kbcj,ikac->ijab', W['ovvo'], t2)
    s2 += contract('kbci,kjac->ijab', W['ovvo'], t2)
    s2 += contract('kacj,ikcb->ijab', W['ovvo'], t2)
    s2 += contract('kaci,kjcb->ijab', W['ovvo'], t2)
    return s1, s2


def blocks_from_packed(host, O, V):
    """Dense antisymmetrized blocks W (keys of BLOCK_NAMES) from the packed
    device layouts of apy_so_ints copied to the host (flat arrays keyed by
    vvvv_p, ring, oooo_p, ovoo_p, vvvo_p, vovv_p, ooov, vvoo_t, oovv).  Each
    packed pair (p<q) fills the (p,q) entry and its negated (q,p) image; the
    diagonals stay zero.  Used where the full n^4 tensor is too large to
    build (O=20, V=160: the dense vvvv block alone is 10.5 GB).

    Pinned by tests/test_oracle_so.py::test_blocks_from_packed_inverts_layouts
    (the host image of the device layouts of a small full tensor maps back
    to blocks_from_eri of that tensor, exactly)."""
    Po, Pv = O * (O - 1) // 2, V * (V - 1) // 2
    io, jo = np.triu_indices(O, 1)
    ia, ib = np.triu_indices(V, 1)
    W = {}
    W['oovv'] = host['oovv'].reshape(O, O, V, V)
    W['vvoo'] = host['vvoo_t'].reshape(O, O, V, V).transpose(2, 3, 0, 1).copy()
    W['ovvo'] = host['ring'].reshape(O, V, O, V).transpose(2, 1, 3, 0).copy()
    Pw = host['vovv_p'].reshape(V, O, Pv)                # [a][j][(c<b)] = <aj||cb>
    vovv = np.zeros((V, O, V, V), dtype=np.complex128)
    vovv[:, :, ia, ib] = Pw
    vovv[:, :, ib, ia] = -Pw
    W['vovv'] = vovv
    W['ooov'] = host['ooov'].reshape(O, O, O, V)
    P = host['oooo_p'].reshape(Po, Po)                  # [(i<j)][(k<l)] = <kl||ij>
    X = np.zeros((O, O, O, O), dtype=np.complex128)      # X[i,j,k,l] = <kl||ij>
    Xv = X.reshape(O * O, O * O)
    r1, r2 = io * O + jo, jo * O + io
    Xv[np.ix_(r1, r1)] = P
    Xv[np.ix_(r1, r2)] = -P
    Xv[np.ix_(r2, r1)] = -P
    Xv[np.ix_(r2, r2)] = P
    W['oooo'] = X.transpose(2, 3, 0, 1).copy()
    Pk = host['ovoo_p'].reshape(O, V, Po)                # [k][b][(i<j)] = <kb||ij>
    ovoo = np.zeros((O, V, O, O), dtype=np.complex128)
    ovoo[:, :, io, jo] = Pk
    ovoo[:, :, jo, io] = -Pk
    W['ovoo'] = ovoo
    W['vooo'] = -ovoo.transpose(1, 0, 2, 3).copy()
    Pq = host['vvvo_p'].reshape(Pv, O, V)                # [(a<b)][j][c] = <ab||cj>
    vvvo = np.zeros((V, V, V, O), dtype=np.complex128)
    vvvo[ia, ib] = Pq.transpose(0, 2, 1)
    vvvo[ib, ia] = -Pq.transpose(0, 2, 1)
    W['vvvo'] = vvvo
    W['vvov'] = -vvvo.transpose(0, 1, 3, 2).copy()
    Pvv = host['vvvv_p'].reshape(Pv, Pv)                 # [(a<b)][(c<d)] = <ab||cd>
    vvvv = np.zeros((V * V, V * V), dtype=np.complex128)
    ra, rb = ia * V + ib, ib * V + ia
    vvvv[np.ix_(ra, ra)] = Pvv
    vvvv[np.ix_(ra, rb)] = -Pvv
    vvvv[np.ix_(rb, ra)] = -Pvv
    vvvv[np.ix_(rb, rb)] = Pvv
    W['vvvv'] = vvvv.reshape(V, V, V, V)
    return W
