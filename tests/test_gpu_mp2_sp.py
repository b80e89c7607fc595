"""Spin-adapted MP2 energy derivatives on the GPU (apy_mp2_sp_*, the forms of
analytic.py:320-335) against the oracle, whose general adjoint is pinned by
central finite differences of non-canonical closed-shell MP2 solves
(tests/test_oracle_mp2_sp.py)."""
import numpy as np
import pytest
import torch

from apyib_amd import mp2_grad
from oracle import mp2_sp

pytestmark = pytest.mark.gpu


def _system(o, v, seed):
    """Pair-symmetric t = W / D from complex integrals with (pq|rs) = (rs|pq)."""
    rng = np.random.default_rng(seed)
    n = o + v
    E4 = 0.1 * (rng.standard_normal((n,) * 4) + 1j * rng.standard_normal((n,) * 4))
    E4 = 0.5 * (E4 + E4.transpose(2, 3, 0, 1))
    W = E4[:o, o:, :o, o:].transpose(0, 2, 1, 3).copy()
    eps = np.concatenate([np.sort(rng.uniform(-1.5, -0.4, o)), np.sort(rng.uniform(0.3, 2.5, v))])
    return W, eps, W / mp2_sp.denominators(eps, o)


def _dirs(o, v, nd, seed, pair_sym=True):
    rng = np.random.default_rng(seed)
    n = o + v
    f = rng.standard_normal((nd, n, n)) + 1j * rng.standard_normal((nd, n, n))
    e = rng.standard_normal((nd, o, o, v, v)) + 1j * rng.standard_normal((nd, o, o, v, v))
    if pair_sym:
        e = 0.5 * (e + e.transpose(0, 2, 1, 4, 3))
    return f, e


def _d(x, dev):
    return torch.from_numpy(np.ascontiguousarray(x)).to(dev)


@pytest.mark.parametrize("o,v,nd,seed", [(4, 6, 3, 1), (5, 9, 2, 2), (10, 40, 2, 3)])
@pytest.mark.parametrize("form", [0, 1])
def test_t2_response_matches_oracle(dev, o, v, nd, seed, form):
    _, eps, t = _system(o, v, seed)
    f, e = _dirs(o, v, nd, seed + 50, pair_sym=False)
    got = mp2_grad.t2_response(_d(t, dev), _d(f, dev), _d(e, dev), eps, form=form).cpu().numpy()
    for q in range(nd):
        ref = mp2_sp.t2_response(t, f[q], e[q], eps, o, form=form)
        assert np.abs(got[q] - ref).max() <= 1e-13 * max(1.0, np.abs(ref).max()), (form, q)
    # no driver: the Fock terms alone; one unbatched direction
    one = mp2_grad.t2_response(_d(t, dev), _d(f[0], dev), None, eps, form=form).cpu().numpy()
    ref = mp2_sp.t2_response(t, f[0], np.zeros_like(t), eps, o, form=form)
    assert one.shape == t.shape
    assert np.abs(one - ref).max() <= 1e-13 * max(1.0, np.abs(ref).max())


@pytest.mark.parametrize("o,v,seed", [(3, 5, 4), (6, 11, 5), (12, 48, 6)])
def test_densities_match_oracle(dev, o, v, seed):
    _, _, t = _system(o, v, seed)
    d_eri, d_oo, d_vv = mp2_grad.gradient_densities(_d(t, dev))
    r_eri, r_oo, r_vv = mp2_sp.densities(t, o)
    assert np.abs(d_eri.cpu().numpy() - r_eri).max() <= 1e-14 * max(1.0, np.abs(r_eri).max())
    assert np.abs(d_oo.cpu().numpy() - r_oo).max() <= 1e-13 * max(1.0, np.abs(r_oo).max())
    assert np.abs(d_vv.cpu().numpy() - r_vv).max() <= 1e-13 * max(1.0, np.abs(r_vv).max())


@pytest.mark.parametrize("o,v,nd,seed", [(4, 7, 5, 7), (9, 30, 3, 8)])
def test_energy_gradient_matches_reference_forms(dev, o, v, nd, seed):
    """E_grad of every direction from the device densities against the
    reference's term-by-term E_grad and its E_grad1 form (through the device
    t2 response), and the f-only and ERI-only parts."""
    W, eps, t = _system(o, v, seed)
    f, e = _dirs(o, v, nd, seed + 9)
    dens = mp2_grad.gradient_densities(_d(t, dev))
    g = mp2_grad.energy_gradient(dens, _d(f, dev), _d(e, dev)).cpu().numpy()
    gf = mp2_grad.energy_gradient(dens, f_grad=_d(f, dev)).cpu().numpy()
    ge = mp2_grad.energy_gradient(dens, eri_grad=_d(e, dev)).cpu().numpy()
    tg = mp2_grad.t2_response(_d(t, dev), _d(f, dev), _d(e, dev), eps, form=0).cpu().numpy()
    for q in range(nd):
        ref = mp2_sp.reference_grad(t, f[q], e[q], o)
        scale = max(1.0, abs(ref))
        assert abs(g[q] - ref) <= 1e-12 * scale
        assert abs(gf[q] + ge[q] - ref) <= 1e-12 * scale
        assert abs(ge[q] - np.sum(2 * (2 * t - t.swapaxes(2, 3)) * e[q])) <= 1e-12 * scale
        g1 = np.sum(tg[q] * (2 * W - W.swapaxes(2, 3))) + np.sum(t * (2 * e[q] - e[q].swapaxes(2, 3)))
        assert abs(g1 - ref) <= 1e-12 * scale


def test_energy_gradient_matches_central_differences(dev):
    """At canonical orbitals with <ij|ab> = <ab|ij> (the reference's
    setting), the device E_grad along (f_grad, ERI_grad) against central
    differences of the oracle's non-canonical MP2 energy with F + h f_grad and
    both integral arrays + h ERI_grad."""
    o, v = 3, 5
    W, eps, t = _system(o, v, 11)
    f, e = _dirs(o, v, 2, 12)
    dens = mp2_grad.gradient_densities(_d(t, dev))
    g = mp2_grad.energy_gradient(dens, _d(f, dev), _d(e, dev)).cpu().numpy()
    F = np.diag(eps).astype(complex)
    h = 1e-5
    for q in range(2):
        Ep, _ = mp2_sp.energy(F + h * f[q], W + h * e[q], W + h * e[q], eps, o)
        Em, _ = mp2_sp.energy(F - h * f[q], W - h * e[q], W - h * e[q], eps, o)
        fd = (Ep - Em) / (2 * h)
        assert abs(g[q] - fd) <= 1e-8 * max(1.0, abs(fd)), (q, g[q], fd)


def test_errors(dev):
    _, eps, t = _system(2, 3, 1)
    with pytest.raises(ValueError):
        mp2_grad.energy_gradient(mp2_grad.gradient_densities(_d(t, dev)))
