"""MP2 energy adjoint on the GPU: parity with the oracle, central finite
differences along Hermitian directions, and the autograd wrapper."""
import numpy as np
import pytest
import torch

from apyib_amd import synthetic
from apyib_amd.adjoint import MP2Adjoint, mp2_energy
from apyib_amd.so_ci import MP2_BLOCKS, SOIntegrals
from oracle import adjoint

pytestmark = pytest.mark.gpu


def _host_system(O, V, seed):
    F, A, eps = synthetic.system(O, V, seed)
    W = A[:O, :O, O:, O:].copy()
    Vd = A[O:, O:, :O, :O].transpose(2, 3, 0, 1).copy()
    return F, W, Vd, eps


def _d(x, dev):
    return torch.from_numpy(np.ascontiguousarray(x)).to(dev)


@pytest.mark.parametrize("O,V,seed", [(4, 6, 1), (6, 9, 2), (3, 5, 7)])
def test_energy_and_gradients_match_oracle(dev, O, V, seed):
    F, W, Vd, eps = _host_system(O, V, seed)
    E_ref, t_ref, lam_ref, GF_ref, GW_ref, GV_ref = adjoint.gradients(F, W, Vd, eps, O)
    s = MP2Adjoint(_d(F, dev), _d(W, dev), _d(Vd, dev), eps, O)
    E = complex(s.energy().cpu())
    GF, GW, GV = s.gradients()
    assert abs(E - E_ref) < 1e-13
    assert np.abs(GF.cpu().numpy() - GF_ref).max() < 1e-12
    assert np.abs(GW.cpu().numpy() - GW_ref).max() < 1e-13
    assert np.abs(GV.cpu().numpy() - GV_ref).max() < 1e-12


def _herm_dirs(O, V, rng, dev):
    n = O + V
    X = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    dF = 0.5 * (X + X.conj().T)
    Y = rng.standard_normal((O, O, V, V)) + 1j * rng.standard_normal((O, O, V, V))
    Y = Y - Y.transpose(1, 0, 2, 3)
    Y = Y - Y.transpose(0, 1, 3, 2)
    # a Hermitian two-electron direction: d<ab||ij> = conj(d<ij||ab>)
    return _d(dF, dev), _d(Y, dev), _d(Y.conj(), dev)


def _dir_check(dev, F, W, Vd, eps, O, ndir, seed, h=1e-5):
    """Largest |FD - adjoint| over the directions, relative to the natural
    scale of the directional sum, sum |dE/dx| |dx| (a direction along which
    E barely changes would otherwise make the relative error meaningless)."""
    tol = 1e-14
    s = MP2Adjoint(F, W, Vd, eps, O, tol=tol)
    E0 = complex(s.energy().cpu())
    GF, GW, GV = s.gradients()
    rng = np.random.default_rng(seed)
    worst = 0.0
    for _ in range(ndir):
        dF, dW, dV = _herm_dirs(O, F.shape[0] - O, rng, dev)
        an = complex((torch.sum(GF * dF) + torch.sum(GW * dW) + torch.sum(GV * dV)).cpu())
        scale = float((torch.sum((GF * dF).abs()) + torch.sum((GW * dW).abs()) + torch.sum((GV * dV).abs())).cpu())
        Ep = complex(MP2Adjoint(F + h * dF, W + h * dW, Vd + h * dV, eps, O, tol=tol).energy().cpu())
        Em = complex(MP2Adjoint(F - h * dF, W - h * dW, Vd - h * dV, eps, O, tol=tol).energy().cpu())
        fd = (Ep - Em) / (2 * h)
        assert abs(fd.imag) < 1e-8 * scale and abs(an.imag) < 1e-10 * scale   # E stays real
        worst = max(worst, abs(fd - an) / scale)
    return E0, worst


def test_directional_derivatives_small(dev):
    O, V = 6, 10
    F, W, Vd, eps = _host_system(O, V, 11)
    _, worst = _dir_check(dev, _d(F, dev), _d(W, dev), _d(Vd, dev), eps, O, ndir=200, seed=3)
    assert worst < 1e-7


def test_directional_derivatives_o30_v270(dev):
    """MP2 adjoint at O=30, V=270 (seeded complex128 integrals generated on
    the device), 2 instances x 4 Hermitian directions, h = 1e-5."""
    O, V = 30, 270
    for seed in (1, 2):
        ints = SOIntegrals.synthetic(O, V, seed, device=dev, blocks=MP2_BLOCKS)
        F, eps = synthetic.fock(O, V, seed)
        W = ints['oovv'].view(O, O, V, V)
        Vd = ints['vvoo_t'].view(O, O, V, V)
        _, worst = _dir_check(dev, _d(F, dev), W, Vd, eps, O, ndir=4, seed=seed)
        assert worst < 1e-7


def test_autograd_function_gradients(dev):
    O, V = 4, 7
    F, W, Vd, eps = _host_system(O, V, 5)
    Ft = _d(F, dev).requires_grad_(True)
    Wt = _d(W, dev).requires_grad_(True)
    Vt = _d(Vd, dev).requires_grad_(True)
    E = mp2_energy(Ft, Wt, Vt, eps, O)
    gF, gW, gV = torch.autograd.grad(E.real, [Ft, Wt, Vt])
    _, _, _, GF, GW, GV = adjoint.gradients(F, W, Vd, eps, O)
    # holomorphic E: torch returns conj(dE/dx) for the real part
    assert np.abs(gF.cpu().numpy() - GF.conj()).max() < 1e-12
    assert np.abs(gW.cpu().numpy() - GW.conj()).max() < 1e-13
    assert np.abs(gV.cpu().numpy() - GV.conj()).max() < 1e-12


@pytest.mark.parametrize("O,V,seed", [(4, 6, 1), (6, 9, 2)])
def test_hermitian_lambda_matches_oracle(dev, O, V, seed):
    """MP2Adjoint(hermitian=True) (lambda = conj(t)/4, no lambda solve) on
    Hermitian systems: the gradients of the oracle's solved lambda."""
    F, W, Vd, eps = _host_system(O, V, seed)
    assert np.allclose(Vd, W.conj(), rtol=0, atol=1e-15)
    _, _, lam_ref, GF_ref, GW_ref, GV_ref = adjoint.gradients(F, W, Vd, eps, O)
    s = MP2Adjoint(_d(F, dev), _d(W, dev), _d(Vd, dev), eps, O, hermitian=True)
    s.energy()
    GF, GW, GV = s.gradients()
    assert np.abs(GF.cpu().numpy() - GF_ref).max() < 1e-12
    assert np.abs(GW.cpu().numpy() - GW_ref).max() < 1e-13
    assert np.abs(GV.cpu().numpy() - GV_ref).max() < 1e-12
