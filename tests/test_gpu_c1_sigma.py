"""Parity at the O=20, V=160 spin-orbital CISD workload itself (seeded
complex128 <pq||rs>, std 0.05 Eh; seeded Fock matrix): the device-generated
integral blocks against the host generator, one CISD sigma/residual against
the oracle's block form of ci_wfn.py:309-334 (the pair-packed ladder, the
wide split-K plans and the pack/unpack index paths at full size), and the
Davidson lowest root checked by the oracle sigma of the returned vector.

Host memory: about 17 GB for the dense oracle blocks (vvvv alone 10.5 GB)."""
import json
import os

import numpy as np
import pytest
import torch

from apyib_amd import _lib, synthetic
from apyib_amd.davidson import CISDDavidson
from apyib_amd.kernels import plans_load, stream_of
from apyib_amd.so_ci import SOCI, SOIntegrals
from oracle import davidson as odav
from oracle import so
from tests.gpu_layouts import block_element_indices
from tests.helpers import rand_amplitudes

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(900)]

O, V, SEED = 20, 160, 1
PLANS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "apyib_amd", "plans_gfx950.json")


@pytest.fixture(scope="module")
def c1(dev):
    ints = SOIntegrals.synthetic(O, V, SEED, device=dev)
    F, eps = synthetic.fock(O, V, SEED)
    host = {k: ints[k].cpu().numpy() for k in ints.blocks}
    W = so.blocks_from_packed(host, O, V)
    del host
    yield ints, F, eps, W
    del W, ints
    torch.cuda.empty_cache()


def _clear_plans():
    _lib.check(_lib.load().apy_zgemm_tune_clear(), "apy_zgemm_tune_clear")


def test_c1_device_blocks_match_generator(c1):
    """4096 seeded positions of every device block against the host
    generator evaluated at the same full-tensor indices (n = 180)."""
    ints = c1[0]
    rng = np.random.default_rng(5)
    for name, blk in ints.blocks.items():
        flat = rng.integers(0, blk.numel(), 4096)
        flat[:2] = (0, blk.numel() - 1)                 # both ends of the layout
        got = blk[torch.from_numpy(flat).to(blk.device)].cpu().numpy()
        p, q, r, s = block_element_indices(name, flat, O, V)
        want = synthetic.eri_elements(O + V, SEED, p, q, r, s)
        assert np.abs(got - want).max() <= 1e-16, name


@pytest.mark.parametrize("plans", ["model", "saved"])
def test_c1_residual_step_matches_oracle(c1, dev, plans):
    """The reference's residual r = sigma - E t and Jacobi update t + r/D
    (apy_cisd_so_step, residual mode, no f_vo coupling) at O=20, V=160,
    with the cost-model GEMM plans and with the saved tuned plans."""
    ints, F, eps, W = c1
    _clear_plans()
    if plans == "saved":
        plans_load(json.load(open(PLANS)))
    try:
        t1, t2 = rand_amplitudes(O, V, 11, scale=0.01)
        E = -0.37 + 0.011j
        solver = SOCI(ints, torch.from_numpy(F).to(dev), eps, singles=True)
        t = torch.from_numpy(np.concatenate([t1.ravel(), t2.ravel()])).to(dev)
        tn, r = solver.new_vector(), solver.new_vector()
        solver.step(t, tn, r, energy=torch.tensor([E], dtype=torch.complex128, device=dev))
        g1, g2 = (x.cpu().numpy() for x in solver.split(r))
        n1, n2 = (x.cpu().numpy() for x in solver.split(tn))
    finally:
        _clear_plans()
    s1, s2 = so.cisd_so_sigma_blocks(F, W, t1, t2, O)
    r1, r2 = s1 - E * t1, s2 - E * t2
    scale = max(np.abs(r1).max(), np.abs(r2).max())
    assert np.abs(g1 - r1).max() <= 1e-12 * scale
    assert np.abs(g2 - r2).max() <= 1e-12 * scale
    D1, D2 = so.denominators(eps[:O], eps[O:])
    assert np.abs(n1 - (t1 + r1 / D1)).max() <= 1e-12 * np.abs(t1 + r1 / D1).max()
    want = t2 + r2 / D2
    assert np.abs(n2 - want).max() <= 1e-12 * np.abs(want).max()


def test_c1_hermitian_sigma_and_packed_layout(c1, dev):
    """sigma mode with the f_vo coupling and reference weight c0 != 1 (the
    Hermitian CISD matrix Davidson applies) against the oracle, and the
    packed subspace layout at full size: pack -> unpack restores a dense
    antisymmetric vector exactly."""
    ints, F, eps, W = c1
    lib = _lib.load()
    t1, t2 = rand_amplitudes(O, V, 12, scale=0.01)
    c0 = 0.83 - 0.12j
    herm = SOCI(ints, torch.from_numpy(F).to(dev), eps, singles=True, fov_coupling=True)
    x = torch.from_numpy(np.concatenate([t1.ravel(), t2.ravel(), [c0]])).to(dev)
    out = torch.zeros_like(x)
    herm.sigma(x[:herm.L], out[:herm.L], c0=c0)
    herm.energy(x[:herm.L], out=out[herm.L:])
    s0, s1, s2 = odav.sigma(F, W, c0, t1, t2, O)
    g = out.cpu().numpy()
    scale = max(np.abs(s1).max(), np.abs(s2).max())
    assert np.abs(g[:O * V] - s1.ravel()).max() <= 1e-12 * scale
    assert np.abs(g[O * V:herm.L] - s2.ravel()).max() <= 1e-12 * scale
    assert abs(g[herm.L] - s0) <= 1e-12 * max(1.0, abs(s0))

    Lp = int(lib.apy_ci_packed_length(O, V))
    packed = torch.zeros(Lp, dtype=torch.complex128, device=dev)
    back = torch.zeros_like(x)
    st = stream_of(dev)
    _lib.check(lib.apy_ci_pack(O, V, _lib.c_vp(x.data_ptr()), _lib.c_vp(packed.data_ptr()), st), "apy_ci_pack")
    _lib.check(lib.apy_ci_unpack(O, V, _lib.c_vp(packed.data_ptr()), _lib.c_vp(back.data_ptr()), st),
               "apy_ci_unpack")
    assert torch.equal(back, x)
    io, jo = np.triu_indices(O, 1)
    ia, ib = np.triu_indices(V, 1)
    p = packed.cpu().numpy()
    want = np.concatenate([t1.ravel(), t2[io, jo][:, ia, ib].ravel(), [c0]])
    assert np.array_equal(p, want)


def test_c1_davidson_root_is_an_eigenpair_of_the_oracle(c1, dev):
    """Davidson lowest root at O=20, V=160 (saved plans, as timed): the
    returned normalised vector is an eigenvector of the oracle's Hermitian
    CISD matrix with the returned energy as its Rayleigh quotient."""
    ints, F, eps, W = c1
    _clear_plans()
    plans_load(json.load(open(PLANS)))
    try:
        herm = SOCI(ints, torch.from_numpy(F).to(dev), eps, singles=True, fov_coupling=True)
        dav = CISDDavidson(herm, max_space=12)
        E = dav.solve(e_conv=1e-11, r_conv=1e-8, maxit=60)
        c = dav.c.cpu().numpy()
    finally:
        _clear_plans()
    L = O * V + O * O * V * V
    c1v, c2v, c0 = c[:O * V].reshape(O, V), c[O * V:L].reshape(O, O, V, V), c[L]
    s0, s1, s2 = odav.sigma(F, W, c0, c1v, c2v, O)
    cv, sv = (c0, c1v, c2v), (s0, s1, s2)
    norm = odav.dot(cv, cv).real
    assert abs(norm - 1.0) < 1e-12
    rq = odav.dot(cv, sv).real / norm
    assert abs(rq - E) <= 1e-10 * max(1.0, abs(E))
    res = tuple(s - E * x for s, x in zip(sv, cv))
    assert np.sqrt(abs(odav.dot(res, res))) < 1e-7


def test_c1_gauss3m_sigma_and_davidson(c1, dev):
    """The three-product form of the PPL and ring products (SOCI gauss3m) at
    O=20, V=160: the Hermitian sigma against the oracle at the same 1e-12
    relative bound as the four-product form, and the Davidson lowest root
    within 1e-10 Eh of the four-product solve."""
    ints, F, eps, W = c1
    t1, t2 = rand_amplitudes(O, V, 13, scale=0.01)
    c0 = 0.91 + 0.07j
    g3 = SOCI(ints, torch.from_numpy(F).to(dev), eps, singles=True, fov_coupling=True, gauss3m=True)
    x = torch.from_numpy(np.concatenate([t1.ravel(), t2.ravel()])).to(dev)
    out = torch.zeros_like(x)
    g3.sigma(x, out, c0=c0)
    _, s1, s2 = odav.sigma(F, W, c0, t1, t2, O)
    g = out.cpu().numpy()
    scale = max(np.abs(s1).max(), np.abs(s2).max())
    assert np.abs(g[:O * V] - s1.ravel()).max() <= 1e-12 * scale
    assert np.abs(g[O * V:] - s2.ravel()).max() <= 1e-12 * scale
    _clear_plans()
    plans_load(json.load(open(PLANS)))
    try:
        E3 = CISDDavidson(g3, max_space=12).solve_device(e_conv=1e-12, r_conv=1e-7, maxit=60, graph=True)
        herm = SOCI(ints, torch.from_numpy(F).to(dev), eps, singles=True, fov_coupling=True)
        E4 = CISDDavidson(herm, max_space=12).solve_device(e_conv=1e-12, r_conv=1e-7, maxit=60, graph=True)
    finally:
        _clear_plans()
    assert abs(E3 - E4) <= 1e-10
