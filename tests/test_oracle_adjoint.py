"""The MP2 adjoint oracle against central finite differences and the
canonical closed form.  CPU only."""
import numpy as np
import pytest

from apyib_amd import synthetic
from oracle import adjoint, so


def _system(O, V, seed):
    F, A, eps = synthetic.system(O, V, seed)
    W = A[:O, :O, O:, O:].copy()                       # <ij||ab>
    Vd = A[O:, O:, :O, :O].transpose(2, 3, 0, 1).copy()  # <ab||ij> at [i,j,a,b]
    return F, W, Vd, eps


def _herm_dir(n, rng):
    X = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    return 0.5 * (X + X.conj().T)


def _asym_dir(O, V, rng):
    X = rng.standard_normal((O, O, V, V)) + 1j * rng.standard_normal((O, O, V, V))
    X = X - X.transpose(1, 0, 2, 3)
    return X - X.transpose(0, 1, 3, 2)


@pytest.mark.parametrize("O,V,seed", [(4, 6, 1), (6, 5, 2)])
def test_gradients_match_central_differences(O, V, seed):
    F, W, Vd, eps = _system(O, V, seed)
    E, t, lam, GF, GW, GV = adjoint.gradients(F, W, Vd, eps, O)
    rng = np.random.default_rng(seed)
    h = 1e-5
    for _ in range(4):
        dF = _herm_dir(O + V, rng)
        dW, dV = _asym_dir(O, V, rng), _asym_dir(O, V, rng)
        Ep, _ = adjoint.energy(F + h * dF, W + h * dW, Vd + h * dV, eps, O)
        Em, _ = adjoint.energy(F - h * dF, W - h * dW, Vd - h * dV, eps, O)
        fd = (Ep - Em) / (2 * h)
        an = np.sum(GF * dF) + np.sum(GW * dW) + np.sum(GV * dV)
        assert abs(fd - an) < 1e-8 * max(1.0, abs(an))


def test_diagonal_fock_reduces_to_canonical_mp2():
    O, V = 4, 6
    F, W, Vd, eps = _system(O, V, 3)
    Fd = np.diag(eps).astype(np.complex128)
    E, t, lam, GF, GW, GV = adjoint.gradients(Fd, W, Vd, eps, O)
    _, D = so.denominators(eps[:O], eps[O:])
    A = synthetic.eri_full(O + V, 3)
    E_can, t_can = so.solve_mp2_so(A / 2, eps[:O], eps[O:], O)
    assert abs(E - E_can) < 1e-14 and np.abs(t - t_can).max() < 1e-14
    assert np.abs(GV - W / (4 * D)).max() < 1e-14
    # dE/d eps_a through F_aa: sum over pairs containing a of -W V / (4 D^2) * (-1 per virtual)
    dEdeps = np.zeros(O + V, dtype=complex)
    x = W * Vd / (4 * D ** 2)
    for i in range(O):
        dEdeps[i] = -(x[i].sum() + x[:, i].sum())
    for a in range(V):
        dEdeps[O + a] = x[:, :, a].sum() + x[:, :, :, a].sum()
    assert np.abs(np.diag(GF) - dEdeps).max() < 1e-13


def test_lambda_of_hermitian_system_is_conjugated_amplitudes():
    """For F = F^H and <ab||ij> = conj(<ij||ab>) the lambda equations
    (F^T, driver <ij||ab>/4) are the complex conjugate of the amplitude
    equations scaled by 1/4, so lambda = conj(t)/4 (the identity
    apy_mp2_adjoint_sums uses when it is given no lambda); a non-Hermitian
    Fock matrix breaks it."""
    O, V = 4, 7
    F, A, eps = synthetic.system(O, V, 12)
    W = A[:O, :O, O:, O:]
    Vd = A[O:, O:, :O, :O].transpose(2, 3, 0, 1)
    assert np.allclose(Vd, W.conj(), rtol=0, atol=1e-15) and np.allclose(F, F.conj().T, rtol=0, atol=1e-15)
    _, t, lam, _, _, _ = adjoint.gradients(F, W, Vd, eps, O)
    assert np.abs(lam - 0.25 * t.conj()).max() < 1e-13 * np.abs(t).max()
    F2 = F.copy()
    F2[O, O + 1] += 0.05j                     # no longer Hermitian
    _, t2, lam2, _, _, _ = adjoint.gradients(F2, W, Vd, eps, O)
    assert np.abs(lam2 - 0.25 * t2.conj()).max() > 1e-6 * np.abs(t2).max()
