"""The two complex128 MP2 finite-difference instance sets at their own sizes:
A = 24 instances at O=10, V=38 (a 3-atom system: 9 nuclear + 3 field
directions, +/- 1e-4) and B = 36 instances at O 14-18, V 54-66 (18 directions,
each with its own seeded size).  The batched solve (all instances per
launch) against the plain per-instance loop on the same kernels at 1e-12,
instances against the oracle's orbital-invariant MP2 (fin_diff.py:267-372
runs one such solve per displacement; mp2_wfn.py:64-87 at diagonal F), and
the adjoint directional derivatives against the central differences of the
set and against the oracle's lambda-based gradient."""
import numpy as np
import pytest
import torch

from apyib_amd.batched import Instance, MP2NCBatch, fd_mp2_instances, mp2_adjoint_directional, plain_solve
from oracle import adjoint as oadj

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(600)]

SEED, H = 1, 1e-4


def set_sizes(name):
    """The (O, V) schedule of set B: 18 seeded draws, O in [14, 18], V in [54, 66]."""
    if name == "A":
        return dict(O=10, V=38, natom=3, sizes=None)
    rng = np.random.default_rng(SEED + 17)
    return dict(O=16, V=60, natom=5, sizes=[(int(rng.integers(14, 19)), int(rng.integers(54, 67))) for _ in range(18)])


@pytest.fixture(scope="module", params=["A", "B"])
def fdset(request, dev):
    cfg = set_sizes(request.param)
    insts, dirs = fd_mp2_instances(cfg["O"], cfg["V"], SEED, cfg["natom"], h=H, device=dev, sizes=cfg["sizes"])
    return request.param, insts, dirs


def _host(I):
    n4 = (I.O, I.O, I.V, I.V)
    return (I.F.cpu().numpy(), I.w.cpu().numpy().reshape(n4), I.drv.cpu().numpy().reshape(n4),
            I.eps.cpu().numpy())


def test_set_shape(fdset):
    name, insts, dirs = fdset
    assert len(insts) == (24 if name == "A" else 36) and len(dirs) == len(insts) // 2
    if name == "B":
        assert len({(I.O, I.V) for I in insts}) > 1
        assert all(14 <= I.O <= 18 and 54 <= I.V <= 66 for I in insts)


def test_batched_equals_plain_loop(fdset):
    _, insts, _ = fdset
    Ep, its = plain_solve(insts, 1e-12)
    b = MP2NCBatch(insts)
    b.solve(1e-12, maxit=400)
    Eb = b.energies()
    scale = float(torch.abs(Ep).max().cpu())
    assert float(torch.abs(Eb - Ep).max().cpu()) <= 1e-12 * max(1.0, scale)
    # amplitudes too
    ref = MP2NCBatch([insts[5]])
    ref.solve_unbatched_steps(1e-12)
    assert float(torch.abs(b.amplitudes(5) - ref.amplitudes(0)).max().cpu()) < 1e-11


@pytest.mark.parametrize("k", [0, 7, 23])
def test_instances_match_oracle(fdset, k):
    _, insts, _ = fdset
    I = insts[k]
    b = MP2NCBatch(insts)
    b.solve(1e-13, maxit=400)
    E = complex(b.energies()[k].cpu())
    F, W, Vd, eps = _host(I)
    Er, tr = oadj.energy(F, W, Vd, eps, I.O)
    assert abs(E - Er) <= 1e-10
    assert np.abs(b.amplitudes(k).cpu().numpy() - tr).max() <= 1e-9


def _base(insts, d):
    P, M = insts[2 * d], insts[2 * d + 1]
    return Instance(P.O, P.V, 0.5 * (P.F + M.F), P.eps, 0.5 * (P.drv + M.drv), 0.5 * (P.w + M.w))


def test_adjoint_matches_differences_of_the_set(fdset):
    """For every direction d, the adjoint derivative at the base point of
    that direction's system against (E(+h) - E(-h)) / 2h of the batched set,
    relative to sum |dE/dx| |dx| (the natural scale of the sum)."""
    name, insts, dirs = fdset
    b = MP2NCBatch(insts)
    b.solve(1e-13, maxit=400)
    E = b.energies().cpu().numpy()
    errs = []
    for d in range(len(dirs)):
        P, M = insts[2 * d], insts[2 * d + 1]
        base = _base(insts, d)
        dF = ((P.F - M.F) / (2 * H)).reshape(-1)
        dW = (P.w - M.w) / (2 * H)
        dV = (P.drv - M.drv) / (2 * H)
        an, (G, t, lam) = mp2_adjoint_directional(base, [dF], [dW], [dV], 1e-13)
        fd = (E[2 * d] - E[2 * d + 1]) / (2 * H)
        scale = float((torch.sum((G * dF).abs()) + 0.25 * torch.sum((t * dW).abs())
                       + torch.sum((lam * dV).abs())).cpu())
        errs.append(abs(an[0] - fd) / scale)
    assert max(errs) < 1e-7, (name, max(errs))


def test_adjoint_gradients_match_oracle(fdset):
    _, insts, _ = fdset
    base = _base(insts, 1)
    n = base.O + base.V
    rng = np.random.default_rng(3)
    dF = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    dFd = torch.from_numpy(dF.reshape(-1)).to(base.F.device)
    an, (G, t, lam) = mp2_adjoint_directional(base, [dFd], None, None, 1e-13)
    F, W, Vd, eps = _host(base)
    Er, tr, lr, Gr, _, _ = oadj.gradients(F, W, Vd, eps, base.O)
    assert np.abs(G.cpu().numpy().reshape(n, n) - Gr).max() <= 1e-10 * max(1.0, np.abs(Gr).max())
    assert np.abs(lam.cpu().numpy().reshape(lr.shape) - lr).max() <= 1e-10 * max(1.0, np.abs(lr).max())
    assert abs(an[0] - np.sum(Gr * dF)) <= 1e-10 * max(1.0, np.sum(np.abs(Gr * dF)))
