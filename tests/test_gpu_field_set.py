"""The one-electron sigma L(dF) (apy_cisd_so_fock_sigma) and the field-set
Davidson that shares one subspace among the +/- one-electron-displaced
instances of a Hermitian CISD problem (FieldSetDavidson).

L(dF) is checked three ways: against the oracle's Hermitian CISD sigma with
every two-electron block zero (the F_SO terms of ci_wfn.py:309-334 plus the
f_ov couplings), as the difference of two full GPU sigmas whose Fock
matrices differ by dF (the sigma is affine in F), and through <c|L(dF)|c> =
sum D_pq dF_pq with D the one-particle density of the Davidson root (pinned
by Hellmann-Feynman in the oracle tests).  The field set is checked against
separate cold-start solves of each instance, with and without subspace
collapses, and one instance's vector as an eigenpair of the oracle matrix."""
import ctypes

import numpy as np
import pytest
import torch

from apyib_amd import _lib, synthetic
from apyib_amd.davidson import CISDDavidson, FieldSetDavidson
from apyib_amd.kernels import stream_of, workspace
from apyib_amd.so_ci import SOCI, SOIntegrals
from oracle import davidson as odav
from oracle import so

pytestmark = pytest.mark.gpu

CDT = torch.complex128


def _vp(t):
    return ctypes.c_void_p(t.data_ptr())


def _random_vector(O, V, seed):
    rng = np.random.default_rng(seed)
    c0 = complex(rng.standard_normal() + 1j * rng.standard_normal())
    c1 = rng.standard_normal((O, V)) + 1j * rng.standard_normal((O, V))
    c2 = rng.standard_normal((O, O, V, V)) + 1j * rng.standard_normal((O, O, V, V))
    c2 = c2 - c2.transpose(1, 0, 2, 3)
    c2 = c2 - c2.transpose(0, 1, 3, 2)
    return c0, c1, c2


def _pack(lib, O, V, c, dev):
    c0, c1, c2 = c
    dense = torch.from_numpy(np.concatenate([c1.ravel(), c2.ravel(), [c0]])).to(dev)
    packed = torch.zeros(int(lib.apy_ci_packed_length(O, V)), dtype=CDT, device=dev)
    _lib.check(lib.apy_ci_pack(O, V, _vp(dense), _vp(packed), stream_of(dev)), "apy_ci_pack")
    return packed


def _unpack(lib, O, V, packed, dev):
    L = O * V + O * O * V * V
    dense = torch.empty(L + 1, dtype=CDT, device=dev)
    _lib.check(lib.apy_ci_unpack(O, V, _vp(packed), _vp(dense), stream_of(dev)), "apy_ci_unpack")
    d = dense.cpu().numpy()
    return d[L], d[:O * V].reshape(O, V), d[O * V:L].reshape(O, O, V, V)


def _fock_sigma(lib, O, V, dFs, packed, dev):
    Po = O * (O - 1) // 2
    hoo = []
    for d in dFs:
        h = torch.zeros(Po * Po, dtype=CDT, device=dev)
        _lib.check(lib.apy_cisd_so_pair_operator(O, V, _vp(d), None, _vp(h), stream_of(dev)), "pair operator")
        hoo.append(h)
    outs = [torch.full_like(packed, float('nan')) for _ in dFs]
    nf = len(dFs)
    args = _lib.CisdSoFockSigmaArgs(O, V, nf, (ctypes.c_void_p * nf)(*[d.data_ptr() for d in dFs]),
                                    (ctypes.c_void_p * nf)(*[h.data_ptr() for h in hoo]), packed.data_ptr(),
                                    (ctypes.c_void_p * nf)(*[o.data_ptr() for o in outs]))
    nb = int(lib.apy_cisd_so_fock_sigma_workspace(O, V))
    ws = workspace(nb, dev)
    _lib.check(lib.apy_cisd_so_fock_sigma(ctypes.byref(args), _vp(ws), nb, stream_of(dev)), "apy_cisd_so_fock_sigma")
    return outs


def _hermitian(n, seed):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    return 0.5 * (X + X.conj().T)


@pytest.mark.parametrize("O,V,seed", [(2, 3, 0), (4, 6, 1), (5, 9, 2), (6, 33, 3)])
def test_fock_sigma_matches_oracle_with_zero_integrals(dev, O, V, seed):
    lib = _lib.load()
    n = O + V
    c = _random_vector(O, V, seed)
    dFs = [_hermitian(n, 10 * seed + q) for q in range(2)] + [np.zeros((n, n), complex)]
    dFs[1] = dFs[1] + 0.3 * np.random.default_rng(seed).standard_normal((n, n))   # non-Hermitian too: L is linear
    outs = _fock_sigma(lib, O, V, [torch.from_numpy(d).to(dev) for d in dFs], _pack(lib, O, V, c, dev), dev)
    Wz = so.blocks_from_eri(np.zeros((n,) * 4, complex), O)
    for dF, out in zip(dFs, outs):
        ref = odav.sigma(dF, Wz, *c, O)
        got = _unpack(lib, O, V, out, dev)
        scale = max(np.abs(ref[1]).max(), np.abs(ref[2]).max(), abs(ref[0]), 1e-300)
        err = max(abs(got[0] - ref[0]), np.abs(got[1] - ref[1]).max(), np.abs(got[2] - ref[2]).max())
        assert err <= 1e-13 * max(scale, 1.0), (err, scale)


def test_fock_sigma_is_the_difference_of_full_sigmas(dev):
    """sigma(F0 + dF) - sigma(F0) = L(dF) through the full packed GPU sigma."""
    O, V, seed = 6, 14, 4
    lib = _lib.load()
    n = O + V
    F0, eps = synthetic.fock(O, V, seed)
    ints = SOIntegrals.synthetic(O, V, seed, device=dev)
    dF = synthetic.fock_perturbation(n, 77, 'field') + synthetic.fock_perturbation(n, 78, 'nuclear')
    c = _random_vector(O, V, seed)
    p = _pack(lib, O, V, c, dev)
    N = p.numel() - 1
    sig = []
    for F in (F0, F0 + dF):
        s = SOCI(ints, torch.from_numpy(F).to(dev), eps, singles=True, fov_coupling=True)
        out = torch.zeros_like(p)
        s.sigma(p[:N], out[:N], c0_dev=p[N:].data_ptr(), packed=True)
        s.energy(p[:N], out=out[N:], packed=True)
        sig.append(out)
    diff = (sig[1] - sig[0]).cpu().numpy()
    L = _fock_sigma(lib, O, V, [torch.from_numpy(dF).to(dev)], p, dev)[0].cpu().numpy()
    scale = np.abs(sig[0].cpu().numpy()).max()
    assert np.abs(diff - L).max() <= 1e-13 * scale


def test_fock_sigma_expectation_is_the_density_contraction(dev):
    """<c|L(dF)|c> = sum_pq D_pq dF_pq for the normalised Davidson root c."""
    O, V, seed = 5, 12, 6
    lib = _lib.load()
    n = O + V
    F0, eps = synthetic.fock(O, V, seed)
    ints = SOIntegrals.synthetic(O, V, seed, device=dev)
    dav = CISDDavidson(SOCI(ints, torch.from_numpy(F0).to(dev), eps, singles=True, fov_coupling=True))
    dav.solve(e_conv=1e-13, r_conv=1e-9)
    D = dav.gradient_fock()
    x = dav.x                                             # the normalised packed root
    for q, kind in enumerate(('field', 'nuclear')):
        dF = torch.from_numpy(synthetic.fock_perturbation(n, 90 + q, kind)).to(dev)
        L = _fock_sigma(lib, O, V, [dF], x, dev)[0]
        lhs = complex(torch.vdot(x, L).cpu())
        rhs = complex(torch.sum(D * dF).cpu())
        assert abs(lhs - rhs) <= 1e-13 * max(1.0, float(torch.sum((D * dF).abs()).cpu())), (lhs, rhs)


def _field_system(O, V, seed, dev, h, ndir=3, kind='field'):
    F0, eps = synthetic.fock(O, V, seed)
    ints = SOIntegrals.synthetic(O, V, seed, device=dev)
    n = O + V
    dirs = [synthetic.fock_perturbation(n, seed * 1000 + 500 + k, kind) for k in range(ndir)]
    inst = [(d, s * h) for d in range(ndir) for s in (+1, -1)]
    return F0, eps, ints, dirs, inst


@pytest.mark.parametrize("O,V,seed,h,max_space,seeded", [(6, 14, 1, 1e-4, 24, False), (6, 14, 1, 1e-4, 24, True),
                                                          (6, 14, 2, 1e-2, 8, False), (8, 20, 3, 1e-4, 8, False),
                                                          (8, 20, 3, 1e-4, 20, True)])
def test_field_set_matches_separate_solves(dev, O, V, seed, h, max_space, seeded):
    """h = 1e-2 with max_space 8: distinct instances and a collapse every few
    iterations; h = 1e-4: +/- pairs share their first-order directions, so the
    set needs fewer full sigmas than the separate solves."""
    F0, eps, ints, dirs, inst = _field_system(O, V, seed, dev, h)
    conv = dict(e_conv=1e-13, r_conv=1e-9)
    fs = FieldSetDavidson(ints, torch.from_numpy(F0).to(dev), eps, [torch.from_numpy(d).to(dev) for d in dirs],
                          inst, max_space=max_space)
    E = fs.solve(**conv, seed_subspace=seeded)
    assert fs.converged
    print("field set: base iterations", fs.base_iterations, "shared", fs.iterations, "sigmas", fs.nsigma)
    nsig_sep = 0
    for k, (d, s) in enumerate(inst):
        dav = CISDDavidson(SOCI(ints, torch.from_numpy(F0 + s * dirs[d]).to(dev), eps, singles=True,
                                fov_coupling=True))
        Ek = dav.solve(**conv)
        nsig_sep += dav.nsigma
        assert abs(E[k] - Ek) < 1e-10, (k, E[k], Ek)
    print("separate solves: sigmas", nsig_sep)
    if h <= 1e-4:
        assert fs.nsigma < nsig_sep


def test_field_set_vector_is_an_eigenpair_of_the_oracle(dev):
    O, V, seed, h = 5, 11, 7, 1e-3
    F0, eps, ints, dirs, inst = _field_system(O, V, seed, dev, h)
    fs = FieldSetDavidson(ints, torch.from_numpy(F0).to(dev), eps, [torch.from_numpy(d).to(dev) for d in dirs],
                          inst, max_space=10)
    E = fs.solve(e_conv=1e-13, r_conv=1e-9)
    lib = _lib.load()
    W = so.blocks_from_eri(synthetic.eri_full(O + V, seed) / 2, O)
    for k in (1, 4):
        d, s = inst[k]
        c = _unpack(lib, O, V, fs.X[k], dev)
        sv = odav.sigma(F0 + s * dirs[d], W, *c, O)
        assert abs(odav.dot(c, c).real - 1.0) < 1e-12
        assert abs(odav.dot(c, sv).real - E[k]) < 1e-11
        res = tuple(a - E[k] * b for a, b in zip(sv, c))
        assert np.sqrt(abs(odav.dot(res, res))) < 1e-8
