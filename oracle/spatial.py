"""Spin-adapted closed-shell CID / CISD / MP2 of apyib restated in NumPy
(oracle only).

`eri` is the Dirac tensor <pq|rs> over spatial active orbitals (the
reference's ERI_MO.swapaxes(1,2), ci_wfn.py:64, 433), F the MO Fock matrix,
O the number of occupied spatial orbitals.

Pinned by: the reference's golden energies (H2O/STO-3G MP2 test_003:21 and
CID test_004:21) in tests/test_oracle_golden.py, and against the spin-orbital
oracle (spatial amplitudes = the alpha-beta spin-orbital amplitudes, the
energies equal) in tests/test_oracle_spatial.py.
"""
import numpy as np

from .so import _converged, solve_general_diis


def contract(expr, *ops):
    return np.einsum(expr, *ops, optimize=True)


def _sl(O, n):
    return slice(0, O), slice(O, n)


def _driver(eri, O):
    o, v = _sl(O, eri.shape[0])
    return eri.swapaxes(0, 2).swapaxes(1, 3)[o, o, v, v]                  # <ab|ij> at [i,j,a,b]


def _denoms(eps, O):
    eo, ev = eps[:O], eps[O:]
    return eo.reshape(-1, 1) - ev, eo.reshape(-1, 1, 1, 1) + eo.reshape(-1, 1, 1) - ev.reshape(-1, 1) - ev


def solve_mp2(eri, eps, O):
    """mp2_wfn.py:42-59."""
    o, v = _sl(O, eri.shape[0])
    _, D = _denoms(eps, O)
    t2 = _driver(eri, O) / D
    E = contract('ijab,ijab->', 2 * eri[o, o, v, v] - eri.swapaxes(2, 3)[o, o, v, v], t2)
    return E, t2


def cid_residual(F, eri, t2, E, O):
    """ci_wfn.py:83-93."""
    o, v = _sl(O, F.shape[0])
    r = 0.5 * _driver(eri, O).astype(np.result_type(F, eri, t2))
    r += contract('ijae,be->ijab', t2, F[v, v])
    r -= contract('imab,mj->ijab', t2, F[o, o])
    r += 0.5 * contract('mnab,mnij->ijab', t2, eri[o, o, o, o])
    r += 0.5 * contract('ijef,abef->ijab', t2, eri[v, v, v, v])
    r += contract('imae,mbej->ijab', t2 - t2.swapaxes(2, 3), eri[o, v, v, o])
    r += contract('imae,mbej->ijab', t2, eri[o, v, v, o] - eri.swapaxes(2, 3)[o, v, v, o])
    r -= contract('mjae,mbie->ijab', t2, eri[o, v, o, v])
    r = r + r.swapaxes(0, 1).swapaxes(2, 3)
    r -= E * t2
    return r


def cid_energy(eri, t2, O):
    o, v = _sl(O, eri.shape[0])
    return contract('ijab,ijab->', 2 * eri[o, o, v, v] - eri.swapaxes(2, 3)[o, o, v, v], t2)


def cisd_energy(F, eri, t1, t2, O):
    o, v = _sl(O, F.shape[0])
    return (2.0 * contract('ia,ia->', t1, F[o, v])
            + contract('ijab,ijab->', t2, 2.0 * eri[o, o, v, v] - eri.swapaxes(2, 3)[o, o, v, v]))


def cisd_residual(F, eri, t1, t2, E, O):
    """ci_wfn.py:457-483."""
    o, v = _sl(O, F.shape[0])
    dt = np.result_type(F, eri, t1, t2)
    X = eri.swapaxes(2, 3)
    r1 = F.swapaxes(0, 1)[o, v].astype(dt)
    r1 -= contract('ji,ja->ia', F[o, o], t1)
    r1 += contract('ab,ib->ia', F[v, v], t1)
    r1 += contract('jabi,jb->ia', 2.0 * eri[o, v, v, o] - X[o, v, v, o], t1)
    r1 += contract('jb,ijab->ia', F[o, v], 2.0 * t2 - t2.swapaxes(2, 3))
    r1 += contract('ajbc,ijbc->ia', 2.0 * eri[v, o, v, v] - X[v, o, v, v], t2)
    r1 -= contract('kjib,kjab->ia', 2.0 * eri[o, o, o, v] - X[o, o, o, v], t2)
    r1 -= E * t1
    r2 = _driver(eri, O).astype(dt)
    r2 += contract('abcj,ic->ijab', eri[v, v, v, o], t1)
    r2 += contract('abic,jc->ijab', eri[v, v, o, v], t1)
    r2 -= contract('kbij,ka->ijab', eri[o, v, o, o], t1)
    r2 -= contract('akij,kb->ijab', eri[v, o, o, o], t1)
    r2 += contract('ac,ijcb->ijab', F[v, v], t2)
    r2 += contract('bc,ijac->ijab', F[v, v], t2)
    r2 -= contract('ki,kjab->ijab', F[o, o], t2)
    r2 -= contract('kj,ikab->ijab', F[o, o], t2)
    r2 += contract('klij,klab->ijab', eri[o, o, o, o], t2)
    r2 += contract('abcd,ijcd->ijab', eri[v, v, v, v], t2)
    r2 -= contract('kbcj,ikca->ijab', eri[o, v, v, o], t2)
    r2 += contract('kaci,kjcb->ijab', 2.0 * eri[o, v, v, o] - X[o, v, v, o], t2)
    r2 -= contract('kbic,kjac->ijab', eri[o, v, o, v], t2)
    r2 -= contract('kaci,kjbc->ijab', eri[o, v, v, o], t2)
    r2 += contract('kbcj,ikac->ijab', 2.0 * eri[o, v, v, o] - X[o, v, v, o], t2)
    r2 -= contract('kajc,ikcb->ijab', eri[o, v, o, v], t2)
    r2 -= E * t2
    return r1, r2


def _iterate(resid, energy, t_list, D_list, params):
    E = energy(*t_list)
    e_iter = t_iter = None
    it = 0
    for it in range(1, params['max_iterations'] + 1):
        E_old = E
        old = [t.copy() for t in t_list]
        rs = resid(*t_list, E)
        t_list = [t + r / D for t, r, D in zip(t_list, rs, D_list)]
        if params['DIIS']:
            res_vec = np.concatenate([r.ravel() for r in rs])
            t_vec = np.concatenate([t.ravel() for t in t_list])
            if it == 1:
                t_iter = t_vec.reshape(-1, 1)
                e_iter = res_vec.reshape(-1, 1)
            t_vec, e_iter, t_iter = solve_general_diis(res_vec, t_vec, e_iter, t_iter, it)
            out, off = [], 0
            for t in t_list:
                out.append(t_vec[off:off + t.size].reshape(t.shape))
                off += t.size
            t_list = out
        E = energy(*t_list)
        rms = [np.sqrt(contract(''.join('ijab'[:t.ndim]) + ',' + ''.join('ijab'[:t.ndim]) + '->', o_ - t, o_ - t))
               for o_, t in zip(old, t_list)]
        if _converged(it, E_old - E, rms, params):
            break
    return E, t_list, it


def solve_cid(F, eri, eps, O, params):
    """ci_wfn.py:51-167.  Returns (E, t2, iterations)."""
    _, D = _denoms(eps, O)
    t2 = (_driver(eri, O) / D).astype(np.complex128)
    E, (t2,), it = _iterate(lambda t, E: (cid_residual(F, eri, t, E, O),), lambda t: cid_energy(eri, t, O),
                            [t2], [D], params)
    return E, t2, it


def solve_cisd(F, eri, eps, O, params):
    """ci_wfn.py:420-574.  Returns (E, t1, t2, iterations)."""
    o, v = _sl(O, F.shape[0])
    D1, D2 = _denoms(eps, O)
    t1 = (F.swapaxes(0, 1)[o, v] / D1).astype(np.complex128)
    t2 = (_driver(eri, O) / D2).astype(np.complex128)
    E, (t1, t2), it = _iterate(lambda a, b, E: cisd_residual(F, eri, a, b, E, O),
                               lambda a, b: cisd_energy(F, eri, a, b, O), [t1, t2], [D1, D2], params)
    return E, t1, t2, it
