"""The spin-adapted MP2 adjoint oracle (analytic.py:320-335 forms) against
central finite differences of a non-canonical closed-shell MP2 solve, and
the reference's three forms (E_grad, E_grad1, the densities) against each
other.  CPU only."""
import numpy as np
import pytest

from oracle import mp2_sp


def _system(o, v, seed, offdiag=0.05):
    """Chemists' integrals with the particle-exchange symmetry only (complex
    orbitals), W = <ij|ab>, V = <ab|ij> at [i,j,a,b], a non-diagonal F."""
    rng = np.random.default_rng(seed)
    n = o + v
    E4 = 0.1 * (rng.standard_normal((n,) * 4) + 1j * rng.standard_normal((n,) * 4))
    E4 = 0.5 * (E4 + E4.transpose(2, 3, 0, 1))              # (pq|rs) = (rs|pq)
    W = E4[:o, o:, :o, o:].transpose(0, 2, 1, 3).copy()     # <ij|ab> = (ia|jb)
    V = E4[o:, :o, o:, :o].transpose(1, 3, 0, 2).copy()     # <ab|ij> = (ai|bj) at [i,j,a,b]
    eps = np.concatenate([np.sort(rng.uniform(-1.5, -0.4, o)), np.sort(rng.uniform(0.3, 2.5, v))])
    X = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    F = np.diag(eps).astype(complex) + offdiag * X
    return F, W, V, eps


def _pair_sym(x):
    return 0.5 * (x + x.transpose(1, 0, 3, 2))


def test_system_has_pair_symmetry():
    F, W, V, eps = _system(3, 5, 1)
    assert np.abs(V - V.transpose(1, 0, 3, 2)).max() < 1e-15
    assert np.abs(W - W.transpose(1, 0, 3, 2)).max() < 1e-15
    assert np.abs(W - V).max() > 1e-2                         # complex orbitals: W != V


@pytest.mark.parametrize("o,v,seed", [(3, 5, 1), (4, 4, 2)])
def test_adjoint_matches_central_differences(o, v, seed):
    F, W, V, eps = _system(o, v, seed)
    rng = np.random.default_rng(seed + 100)
    n = o + v
    dF = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    dW = _pair_sym(rng.standard_normal(W.shape) + 1j * rng.standard_normal(W.shape))
    dV = _pair_sym(rng.standard_normal(V.shape) + 1j * rng.standard_normal(V.shape))
    an = mp2_sp.directional(F, W, V, eps, o, dF, dW, dV)
    h = 1e-5
    Ep, _ = mp2_sp.energy(F + h * dF, W + h * dW, V + h * dV, eps, o)
    Em, _ = mp2_sp.energy(F - h * dF, W - h * dW, V - h * dV, eps, o)
    fd = (Ep - Em) / (2 * h)
    assert abs(an - fd) <= 1e-8 * max(1.0, abs(fd)), (an, fd)


@pytest.mark.parametrize("o,v,seed", [(3, 5, 3), (5, 3, 4)])
def test_reference_forms_agree_at_canonical_orbitals(o, v, seed):
    """E_grad (analytic.py:320-324), E_grad1 (:333-334, through the t2
    response :326-331), the densities, and the general adjoint with
    dW = dV = ERI_grad, at diagonal F with W = V (the reference's
    real-orbital setting)."""
    _, W, _, eps = _system(o, v, seed)
    F = np.diag(eps).astype(complex)
    V = W.copy()
    t = mp2_sp.solve(F, V, eps, o)
    assert np.abs(t - V / mp2_sp.denominators(eps, o)).max() < 1e-15     # mp2_wfn.py:54
    rng = np.random.default_rng(seed + 7)
    n = o + v
    fg = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    eg = _pair_sym(rng.standard_normal(W.shape) + 1j * rng.standard_normal(W.shape))
    g_ref = mp2_sp.reference_grad(t, fg, eg, o)
    g_resp = mp2_sp.grad_via_response(t, fg, eg, W, eps, o)
    g_adj = mp2_sp.directional(F, W, V, eps, o, fg, eg, eg)
    d_eri, poo, pvv = mp2_sp.densities(t, o)
    g_den = np.sum(d_eri * eg) + np.sum(poo * fg[:o, :o]) + np.sum(pvv * fg[o:, o:])
    scale = max(1.0, abs(g_ref))
    assert abs(g_resp - g_ref) <= 1e-13 * scale
    assert abs(g_adj - g_ref) <= 1e-12 * scale
    assert abs(g_den - g_ref) <= 1e-13 * scale


def test_densities_against_their_sums():
    """dE/df_oo and dE/df_vv from the two terms each of the reference form,
    term by term (no pair-symmetry shortcut) for pair-symmetric t."""
    _, W, _, eps = _system(3, 4, 5)
    t = W / mp2_sp.denominators(eps, 3)
    u = 2 * t - t.swapaxes(2, 3)
    poo = -np.einsum('ijab,kjab->ik', t, u) - np.einsum('jiab,jkab->ik', t, u)
    pvv = np.einsum('ijab,ijcb->ac', t, u) + np.einsum('ijba,ijbc->ac', t, u)
    d_eri, p_oo, p_vv = mp2_sp.densities(t, 3)
    assert np.abs(d_eri - 2 * u).max() == 0
    assert np.abs(p_oo - poo).max() < 1e-14
    assert np.abs(p_vv - pvv).max() < 1e-14


def test_response_forms():
    """Form 0 is the literal analytic.py:327-330 order, form 1 the literal
    analytic_aats.py:346-350 order; they differ for a non-symmetric f_grad
    and agree for a symmetric one."""
    o, v = 3, 4
    _, W, _, eps = _system(o, v, 6)
    t = W / mp2_sp.denominators(eps, o)
    rng = np.random.default_rng(8)
    n = o + v
    f = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    eg = rng.standard_normal(t.shape) + 1j * rng.standard_normal(t.shape)
    D = mp2_sp.denominators(eps, o)
    fo, fv = f[:o, :o], f[o:, o:]
    lit0 = eg.copy()
    lit0 -= np.einsum('kjab,ik->ijab', t, fo)
    lit0 -= np.einsum('ikab,kj->ijab', t, fo)
    lit0 += np.einsum('ijcb,ac->ijab', t, fv)
    lit0 += np.einsum('ijac,cb->ijab', t, fv)
    lit1 = eg + np.einsum('ac,ijcb->ijab', fv, t) + np.einsum('bc,ijac->ijab', fv, t)
    lit1 -= np.einsum('ki,kjab->ijab', fo, t) + np.einsum('kj,ikab->ijab', fo, t)
    r0 = mp2_sp.t2_response(t, f, eg, eps, o, form=0)
    r1 = mp2_sp.t2_response(t, f, eg, eps, o, form=1)
    assert np.abs(r0 - lit0 / D).max() < 1e-13
    assert np.abs(r1 - lit1 / D).max() < 1e-13
    assert np.abs(r0 - r1).max() > 1e-2
    fs = 0.5 * (f + f.T)
    assert np.abs(mp2_sp.t2_response(t, fs, eg, eps, o, 0) - mp2_sp.t2_response(t, fs, eg, eps, o, 1)).max() < 1e-13
