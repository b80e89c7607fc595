"""Spin-adapted closed-shell CISD / CID on the GPU (apy_cisd_sp_step,
apy_cisd_sp_energy) against the spatial oracle (oracle/spatial.py, the
reference's ci_wfn.py:420-574 and :51-167 restated): one residual step on
random pair-symmetric amplitudes, the energy, and whole Jacobi + DIIS solves
iterate for iterate."""
import numpy as np
import pytest
import torch

from apyib_amd.spatial_ci import SPIntegrals, SpatialCI
from oracle import spatial
from tests.helpers import PARAMS

pytestmark = pytest.mark.gpu


def _d(x, dev):
    return torch.from_numpy(np.ascontiguousarray(x)).to(dev)


def _system(o, v, seed, std=0.05, noise=1e-3):
    """Chemists' tensor with the particle-exchange symmetry (pq|rs) = (rs|pq)
    of any MO integral set, a Fock matrix with diagonal eps and small
    off-diagonal couplings, and the orbital energies."""
    rng = np.random.default_rng(seed)
    n = o + v
    E = std * (rng.standard_normal((n,) * 4) + 1j * rng.standard_normal((n,) * 4))
    E = 0.5 * (E + E.transpose(2, 3, 0, 1))
    eps = np.concatenate([np.sort(rng.uniform(-2.0, -0.5, o)), np.sort(rng.uniform(0.3, 2.0, v))])
    F = np.diag(eps).astype(complex) + noise * (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    return E, F, eps


def _amps(o, v, seed):
    rng = np.random.default_rng(seed)
    t1 = 0.05 * (rng.standard_normal((o, v)) + 1j * rng.standard_normal((o, v)))
    t2 = 0.05 * (rng.standard_normal((o, o, v, v)) + 1j * rng.standard_normal((o, o, v, v)))
    t2 = 0.5 * (t2 + t2.transpose(1, 0, 3, 2))
    return t1, t2


@pytest.mark.parametrize("o,v,seed", [(1, 4, 8), (3, 1, 9), (2, 3, 1), (4, 7, 2), (6, 17, 3)])
@pytest.mark.parametrize("singles", [True, False])
def test_residual_step_matches_oracle(dev, o, v, seed, singles):
    E4, F, eps = _system(o, v, seed)
    eri = E4.swapaxes(1, 2)                                   # <pq|rs> = (pr|qs)
    t1, t2 = _amps(o, v, seed + 10)
    if not singles:
        t1 = np.zeros_like(t1)
    Ec = 0.02 - 0.003j
    ints = SPIntegrals.from_mo(_d(E4, dev), o, singles=singles)
    ci = SpatialCI(ints, _d(F, dev), eps, singles=singles)
    t = _d(np.concatenate([t1.ravel(), t2.ravel()]), dev)
    tn, r = ci.new_vector(), ci.new_vector()
    ci.step(t, tn, r, energy=torch.tensor([Ec], dtype=torch.complex128, device=dev))
    g1, g2 = (x.cpu().numpy() for x in ci.split(r))
    if singles:
        r1, r2 = spatial.cisd_residual(F, eri, t1, t2, Ec, o)
        assert np.abs(g1 - r1).max() < 1e-13 * max(1.0, np.abs(r1).max())
    else:
        r2 = spatial.cid_residual(F, eri, t2, Ec, o)
        assert not np.any(g1)
    assert np.abs(g2 - r2).max() < 1e-13 * max(1.0, np.abs(r2).max())
    D1, D2 = spatial._denoms(eps, o)
    n1, n2 = (x.cpu().numpy() for x in ci.split(tn))
    assert np.abs(n2 - (t2 + r2 / D2)).max() < 1e-12
    if singles:
        assert np.abs(n1 - (t1 + r1 / D1)).max() < 1e-12
        Eref = spatial.cisd_energy(F, eri, t1, t2, o)
    else:
        Eref = spatial.cid_energy(eri, t2, o)
    Eg = complex(ci.energy(t).cpu()[0])
    assert abs(Eg - Eref) < 1e-14 * max(1.0, abs(Eref)) + 1e-15


@pytest.mark.parametrize("o,v,seed", [(3, 5, 4), (5, 12, 5)])
def test_cisd_solve_matches_oracle(dev, o, v, seed):
    E4, F, eps = _system(o, v, seed)
    eri = E4.swapaxes(1, 2)
    Er, r1, r2, it_ref = spatial.solve_cisd(F, eri, eps, o, PARAMS)
    ci = SpatialCI(SPIntegrals.from_mo(_d(E4, dev), o), _d(F, dev), eps)
    E, t1, t2, it = ci.solve(PARAMS)
    assert it == it_ref
    assert abs(E - Er) < 1e-10
    assert np.abs(t1.cpu().numpy() - r1).max() < 1e-9 and np.abs(t2.cpu().numpy() - r2).max() < 1e-9


@pytest.mark.parametrize("o,v,seed", [(3, 5, 6), (5, 12, 7)])
def test_cid_solve_matches_oracle(dev, o, v, seed):
    E4, F, eps = _system(o, v, seed)
    eri = E4.swapaxes(1, 2)
    Er, r2, it_ref = spatial.solve_cid(F, eri, eps, o, PARAMS)
    ci = SpatialCI(SPIntegrals.from_mo(_d(E4, dev), o, singles=False), _d(F, dev), eps, singles=False)
    E, _, t2, it = ci.solve(PARAMS)
    assert it == it_ref
    assert abs(E - Er) < 1e-10
    assert np.abs(t2.cpu().numpy() - r2).max() < 1e-9


@pytest.mark.parametrize("singles", [True, False])
def test_folded_ring_operator_matches_separate_fvv_product(dev, singles):
    """apy_cisd_sp_step with w21_f (F_vv folded into the ring operator) and
    with the separate F_vv product (w21_f null) agree, with a non-diagonal F."""
    import ctypes
    from apyib_amd import _lib
    from apyib_amd.kernels import stream_of, workspace
    o, v = 5, 11
    E4, F, eps = _system(o, v, 21, noise=2e-2)
    ints = SPIntegrals.from_mo(_d(E4, dev), o, singles=singles)
    ci = SpatialCI(ints, _d(F, dev), eps, singles=singles)
    t1, t2 = _amps(o, v, 22)
    if not singles:
        t1 = np.zeros_like(t1)
    t = _d(np.concatenate([t1.ravel(), t2.ravel()]), dev)
    E = torch.tensor([0.01 + 0.002j], dtype=torch.complex128, device=dev)
    tn_f, r_f = ci.new_vector(), ci.new_vector()
    ci.step(t, tn_f, r_f, energy=E)
    tn, r = ci.new_vector(), ci.new_vector()
    a = _lib.CisdSpArgs(o, v, ci.fock.data_ptr(), ci.eps.data_ptr(), ctypes.pointer(ints.struct), ci.hoo.data_ptr(),
                        ci.g.data_ptr(), int(singles), E.data_ptr(), t.data_ptr(), tn.data_ptr(), r.data_ptr())
    ws = workspace(ci.ws_bytes, dev)
    assert ci.lib.apy_cisd_sp_step(ctypes.byref(a), ctypes.c_void_p(ws.data_ptr()), ci.ws_bytes, stream_of(dev)) == 0
    scale = float(torch.abs(r).max().cpu())
    assert float(torch.abs(r_f - r).max().cpu()) < 1e-13 * max(1.0, scale)
    assert float(torch.abs(tn_f - tn).max().cpu()) < 1e-12


@pytest.mark.parametrize("o,v,seed", [(1, 4, 8), (4, 7, 2), (6, 17, 3), (10, 40, 4)])
@pytest.mark.parametrize("singles", [True, False])
def test_gauss3m_residual_step_matches_oracle(dev, o, v, seed, singles):
    """SpatialCI(gauss3m=True), the ladder and ring products in the
    three-product complex form: one residual step against the oracle at the
    bound the four-product form meets, and the Jacobi + DIIS solve to the
    same energy."""
    E4, F, eps = _system(o, v, seed)
    eri = E4.swapaxes(1, 2)
    t1, t2 = _amps(o, v, seed + 10)
    if not singles:
        t1 = np.zeros_like(t1)
    Ec = 0.02 - 0.003j
    ints = SPIntegrals.from_mo(_d(E4, dev), o, singles=singles)
    ci = SpatialCI(ints, _d(F, dev), eps, singles=singles, gauss3m=True)
    t = _d(np.concatenate([t1.ravel(), t2.ravel()]), dev)
    tn, r = ci.new_vector(), ci.new_vector()
    ci.step(t, tn, r, energy=torch.tensor([Ec], dtype=torch.complex128, device=dev))
    g1, g2 = (x.cpu().numpy() for x in ci.split(r))
    if singles:
        r1, r2 = spatial.cisd_residual(F, eri, t1, t2, Ec, o)
        assert np.abs(g1 - r1).max() < 1e-13 * max(1.0, np.abs(r1).max())
    else:
        r2 = spatial.cid_residual(F, eri, t2, Ec, o)
    assert np.abs(g2 - r2).max() < 1e-13 * max(1.0, np.abs(r2).max())
    if o * v <= 60:
        E3 = ci.solve(dict(PARAMS))[0]
        E4s = SpatialCI(ints, _d(F, dev), eps, singles=singles).solve(dict(PARAMS))[0]
        assert abs(E3 - E4s) < 1e-11
