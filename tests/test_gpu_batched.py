"""Batched finite-difference MP2 instance sets on the GPU: the one-launch
batched iteration against the same kernels run per instance, against the
oracle, and the finite-difference derivatives it gives against the adjoint
gradient."""
import numpy as np
import pytest
import torch

from apyib_amd import _lib, synthetic
from apyib_amd.adjoint import MP2Adjoint
from apyib_amd.batched import Instance, MP2NCBatch, fd_mp2_instances, plain_solve
from oracle import adjoint

pytestmark = pytest.mark.gpu


def _host_instance(O, V, seed, dev):
    F, A, eps = synthetic.system(O, V, seed)
    W = A[:O, :O, O:, O:].copy()
    Vd = A[O:, O:, :O, :O].transpose(2, 3, 0, 1).copy()
    d = lambda x: torch.from_numpy(np.ascontiguousarray(x)).to(dev)
    inst = Instance(O, V, d(F), d(np.asarray(eps, dtype=np.float64)), d(Vd).reshape(-1), d(W).reshape(-1))
    return inst, (F, W, Vd, eps)


def test_batched_equals_per_instance_bitwise(dev):
    """Uneven sizes; the same number of iterations batched (graph) and one
    instance at a time gives identical amplitudes."""
    sizes = [(4, 6), (3, 9), (6, 5), (2, 11), (5, 7)]
    insts = [_host_instance(O, V, 10 + k, dev)[0] for k, (O, V) in enumerate(sizes)]
    b = MP2NCBatch(insts)
    b.solve(tol=0.0, maxit=12, chunk=6)          # exactly 12 iterations
    for k, I in enumerate(insts):
        s = MP2NCBatch([I])
        for q in range(12):
            s.step(q % 2)
        assert torch.equal(b.x[k], s.x[0])
    assert b.iterations == 12


def test_batched_converges_to_oracle(dev):
    sizes = [(4, 6), (3, 9), (6, 5)]
    data = [_host_instance(O, V, 20 + k, dev) for k, (O, V) in enumerate(sizes)]
    b = MP2NCBatch([d[0] for d in data])
    b.solve(tol=1e-13, maxit=400, chunk=8)
    E = b.energies().cpu().numpy()
    for k, (_, (F, W, Vd, eps)) in enumerate(data):
        E_ref, t_ref = adjoint.energy(F, W, Vd, eps, sizes[k][0])
        assert abs(E[k] - E_ref) < 1e-12
        assert np.abs(b.amplitudes(k).cpu().numpy() - t_ref).max() < 1e-11


def test_lambda_batch_matches_adjoint(dev):
    inst, (F, W, Vd, eps) = _host_instance(5, 8, 3, dev)
    lam_inst = Instance(5, 8, inst.F, inst.eps, inst.w, None)
    b = MP2NCBatch([lam_inst], transpose_f=True, drv_scale=0.25)
    b.solve(tol=1e-14, maxit=400)
    _, _, lam_ref, _, _, _ = adjoint.gradients(F, W, Vd, eps, 5)
    assert np.abs(b.amplitudes(0).cpu().numpy() - lam_ref).max() < 1e-12


def test_plain_loop_energies_match_batched(dev):
    insts, _ = fd_mp2_instances(6, 14, 5, natom=1, device=dev)
    b = MP2NCBatch(insts)
    b.solve(tol=1e-12, maxit=400)
    Eb = b.energies()
    Ep, its = plain_solve(insts, tol=1e-12)
    assert torch.abs(Eb - Ep).max().item() < 1e-12
    assert len(its) == len(insts)


def test_fd_derivatives_match_adjoint(dev):
    """Central differences from the batched +/- set against the adjoint
    directional derivatives at the base point (h = 1e-4; the central
    difference error is O(h^2))."""
    O, V, seed, h = 6, 14, 8, 1e-4
    insts, dirs = fd_mp2_instances(O, V, seed, natom=1, h=h, device=dev)
    b = MP2NCBatch(insts)
    b.solve(tol=1e-14, maxit=600)
    E = b.energies().cpu().numpy()
    fd = [(E[2 * d] - E[2 * d + 1]) / (2 * h) for d in range(len(dirs))]
    # base point = average of the +/- pair of direction 0 (F and integrals)
    Ip, Im = insts[0], insts[1]
    F0 = 0.5 * (Ip.F + Im.F)
    W0 = (0.5 * (Ip.w + Im.w)).view(O, O, V, V)
    V0 = (0.5 * (Ip.drv + Im.drv)).view(O, O, V, V)
    s = MP2Adjoint(F0, W0, V0, Ip.eps, O, tol=1e-14)
    s.energy()
    GF, GW, GV = s.gradients()
    for d in range(len(dirs)):
        P, M = insts[2 * d], insts[2 * d + 1]
        dF = (P.F - M.F) / (2 * h)
        dW = ((P.w - M.w) / (2 * h)).view(O, O, V, V)
        dV = ((P.drv - M.drv) / (2 * h)).view(O, O, V, V)
        an = complex((torch.sum(GF * dF) + torch.sum(GW * dW) + torch.sum(GV * dV)).cpu())
        scale = float((torch.sum((GF * dF).abs()) + torch.sum((GW * dW).abs()) + torch.sum((GV * dV).abs())).cpu())
        assert abs(fd[d] - an) < 1e-7 * scale, (d, fd[d], an)


def test_graph_replay_equals_direct_launches(dev):
    """The same chunk replayed from a captured HIP graph (several replays,
    with eager work in between) gives the directly launched iterates."""
    sizes = [(4, 6), (3, 9)]
    insts = [_host_instance(O, V, 40 + k, dev)[0] for k, (O, V) in enumerate(sizes)]
    a, b = MP2NCBatch(insts), MP2NCBatch(insts)
    a.solve(tol=0.0, maxit=12, chunk=4, graph=True)
    b.solve(tol=0.0, maxit=12, chunk=4, graph=False)
    for k in range(len(insts)):
        assert torch.equal(a.x[k], b.x[k])
    assert torch.equal(a.rnorm, b.rnorm)


def test_adjoint_directional_matches_fd(dev):
    from apyib_amd.batched import mp2_adjoint_directional
    O, V, seed, h = 5, 12, 9, 1e-4
    insts, dirs = fd_mp2_instances(O, V, seed, natom=1, h=h, device=dev)
    b = MP2NCBatch(insts)
    b.solve(tol=1e-14, maxit=600)
    E = b.energies().cpu().numpy()
    P, M = insts[0], insts[1]
    base = Instance(O, V, 0.5 * (P.F + M.F), P.eps, 0.5 * (P.drv + M.drv), 0.5 * (P.w + M.w))
    nd = len(dirs)
    dF = [((insts[2 * k].F - insts[2 * k + 1].F) / (2 * h)).reshape(-1) for k in range(nd)]
    dW = [(insts[2 * k].w - insts[2 * k + 1].w) / (2 * h) for k in range(nd)]
    dV = [(insts[2 * k].drv - insts[2 * k + 1].drv) / (2 * h) for k in range(nd)]
    an, (G, t, lam) = mp2_adjoint_directional(base, dF, dW, dV, tol=1e-14)
    for k in range(nd):
        fd = (E[2 * k] - E[2 * k + 1]) / (2 * h)
        scale = float((torch.sum((G * dF[k]).abs()) + 0.25 * torch.sum((t * dW[k]).abs())
                       + torch.sum((lam * dV[k]).abs())).cpu())
        assert abs(an[k] - fd) < 1e-7 * scale


@pytest.mark.parametrize("transpose", [False, True], ids=["amplitudes", "lambda"])
def test_batched_wide_instances(dev, transpose):
    """Instances spanning one to three 32-wide virtual tiles (ragged last
    tiles) and single-pair occupied spaces in one batch converge to the
    oracle's amplitudes (and lambda with F^T), and a graph-replayed chunk
    equals direct launches bit for bit."""
    sizes = [(3, 50), (2, 6), (2, 49), (4, 70), (2, 33)]
    data = [_host_instance(O, V, 40 + k, dev) for k, (O, V) in enumerate(sizes)]
    if transpose:
        insts = [Instance(I.O, I.V, I.F, I.eps, I.w, None) for I, _ in data]
        b = MP2NCBatch(insts, transpose_f=True, drv_scale=0.25)
    else:
        b = MP2NCBatch([d[0] for d in data])
    b.solve(tol=1e-13, maxit=400, chunk=8)
    for k, (_, (F, W, Vd, eps)) in enumerate(data):
        if transpose:
            _, _, want, _, _, _ = adjoint.gradients(F, W, Vd, eps, sizes[k][0])
        else:
            _, want = adjoint.energy(F, W, Vd, eps, sizes[k][0])
        got = b.amplitudes(k).cpu().numpy()
        assert np.abs(got - want).max() < 1e-11 * max(1.0, np.abs(want).max()), k
    g = MP2NCBatch([d[0] for d in data])
    h = MP2NCBatch([d[0] for d in data])
    g.solve(tol=0.0, maxit=8, chunk=4, graph=True)
    h.solve(tol=0.0, maxit=8, chunk=4, graph=False)
    for k in range(len(sizes)):
        assert torch.equal(g.x[k], h.x[k])


def test_packed_storage_round_trip_and_warm_start(dev):
    """The packed amplitudes expand to the antisymmetric dense tensor and
    pack back exactly; a batch started from converged amplitudes stays at
    them (residual at rounding level after one chunk)."""
    sizes = [(4, 9), (3, 40)]
    data = [_host_instance(O, V, 60 + k, dev) for k, (O, V) in enumerate(sizes)]
    b = MP2NCBatch([d[0] for d in data])
    b.solve(tol=1e-13, maxit=400, chunk=8)
    for k in range(len(sizes)):
        t = b.amplitudes(k)
        assert torch.equal(t, -t.transpose(0, 1)) and torch.equal(t, -t.transpose(2, 3))
        w = MP2NCBatch([data[k][0]])
        w.start_from(0, t)
        assert torch.equal(w.x[0], b.x[k])
        w.solve(tol=0.0, maxit=2, chunk=2)
        assert float(w.rnorm.max().cpu()) ** 0.5 < 1e-12


@pytest.mark.parametrize("tol,maxit", [(1e-12, 400), (0.0, 7), (0.0, 8), (1e-3, 400)])
def test_persistent_solve_equals_direct_launches(dev, tol, maxit):
    """One cooperative launch for the whole solve (apy_mp2_nc_batched_solve)
    against the step-by-step solve with a host test after each step: the
    same stopping step and bit-identical amplitudes and energies, for uneven
    sizes with amplitude and lambda instances in one batch, odd and even
    step counts (x_new copied back) and the maxit stop."""
    sizes = [(4, 6), (3, 9), (6, 5), (2, 11), (5, 7), (7, 40), (3, 66)]
    insts = []
    for k, (O, V) in enumerate(sizes):
        inst, _ = _host_instance(O, V, 40 + k, dev)
        if k % 3 == 2:                                   # a lambda instance: F^T and the driver <ij||ab>/4
            inst = Instance(O, V, inst.F, inst.eps, inst.w, inst.w, transpose_f=True, drv_scale=0.25)
        insts.append(inst)
    d = MP2NCBatch(insts)
    itd = d.solve(tol=tol, maxit=maxit)
    p = MP2NCBatch(insts)
    itp = p.solve(tol=tol, maxit=maxit, persistent=True)
    assert itd == itp
    for k in range(len(insts)):
        assert torch.equal(d.x[k], p.x[k])
    assert torch.equal(d.energies(), p.energies())
    assert torch.equal(d.rnorm, p.rnorm)


def test_persistent_solve_set_b_sizes(dev):
    """The 36-instance uneven set at its own sizes (O 14-18, V 54-66): the
    persistent solve against the direct launches, bit for bit."""
    rng = np.random.default_rng(18)
    sizes = [(int(rng.integers(14, 19)), int(rng.integers(54, 67))) for _ in range(18)]
    insts = fd_mp2_instances(16, 60, 1, 5, h=1e-4, device=dev, sizes=sizes)[0]
    d = MP2NCBatch(insts)
    itd = d.solve(tol=1e-12, maxit=400)
    p = MP2NCBatch(insts)
    itp = p.solve(tol=1e-12, maxit=400, persistent=True)
    assert itd == itp and len(insts) == 36
    assert torch.equal(d.energies(), p.energies())
    for k in range(len(insts)):
        assert torch.equal(d.x[k], p.x[k])


def test_adjoint_sums_match_separate_densities_and_dots(dev):
    """apy_mp2_adjoint_sums (every base's densities and directional sums in
    one call) against apy_mp2_nc_densities on the unpacked amplitudes and
    torch sums of the three terms: uneven bases, more than 16 directions
    on one base (chunked dots), and bases without the dW or dV terms."""
    from apyib_amd.batched import mp2_adjoint_set
    sizes = [(4, 9), (3, 11), (5, 7)]
    bases = [_host_instance(O, V, 70 + k, dev)[0] for k, (O, V) in enumerate(sizes)]
    g = torch.Generator(device="cpu").manual_seed(5)
    rnd = lambda n: torch.complex(torch.randn(n, generator=g, dtype=torch.float64),
                                  torch.randn(n, generator=g, dtype=torch.float64)).to(dev)
    ndirs = [19, 3, 5]
    directions = []
    for b, (B, K) in enumerate(zip(bases, ndirs)):
        n, d2 = B.O + B.V, B.O * B.O * B.V * B.V
        dF = [rnd(n * n) for _ in range(K)]
        dW = [rnd(d2) for _ in range(K)] if b != 1 else None
        dV = [rnd(d2) for _ in range(K)] if b != 2 else None
        directions.append((dF, dW, dV))
    out, dens = mp2_adjoint_set(bases, directions, tol=1e-13)
    lib = _lib.load()
    for b, (B, (dF, dW, dV)) in enumerate(zip(bases, directions)):
        G, t, lam = dens[b]
        n = B.O + B.V
        G2 = torch.empty(n * n, dtype=torch.complex128, device=dev)
        nb = int(lib.apy_mp2_nc_density_workspace(B.O, B.V))
        ws = torch.empty(max(nb, 1), dtype=torch.uint8, device=dev)
        assert lib.apy_mp2_nc_densities(B.O, B.V, _lib.c_vp(t.data_ptr()), _lib.c_vp(lam.data_ptr()),
                                        _lib.c_vp(G2.data_ptr()), _lib.c_vp(ws.data_ptr()), nb,
                                        _lib.c_vp(torch.cuda.current_stream(dev).cuda_stream)) == 0
        assert torch.equal(G, G2)
        for q in range(len(dF)):
            want = torch.sum(G * dF[q])
            if dW is not None:
                want = want + 0.25 * torch.sum(t * dW[q])
            if dV is not None:
                want = want + torch.sum(lam * dV[q])
            want = complex(want.cpu())
            assert abs(out[b][q] - want) <= 1e-12 * max(1.0, abs(want)), (b, q)


def test_adjoint_hermitian_lambda_matches_solved_lambda(dev):
    """mp2_adjoint_set(hermitian=True) (amplitudes only, lambda = conj(t)/4
    inside apy_mp2_adjoint_sums) against the solved lambda equations on
    Hermitian synthetic bases: same densities and directional sums to the
    convergence threshold."""
    from apyib_amd.batched import mp2_adjoint_set
    sizes = [(4, 9), (3, 11), (6, 13)]
    bases = [_host_instance(O, V, 80 + k, dev)[0] for k, (O, V) in enumerate(sizes)]
    for B in bases:
        assert torch.allclose(B.drv, B.w.conj(), rtol=0, atol=1e-15)
        assert torch.allclose(B.F, B.F.conj().T, rtol=0, atol=1e-15)
    g = torch.Generator(device="cpu").manual_seed(7)
    rnd = lambda n: torch.complex(torch.randn(n, generator=g, dtype=torch.float64),
                                  torch.randn(n, generator=g, dtype=torch.float64)).to(dev)
    directions = [([rnd((B.O + B.V) ** 2) for _ in range(4)], [rnd(B.O ** 2 * B.V ** 2) for _ in range(4)],
                   [rnd(B.O ** 2 * B.V ** 2) for _ in range(4)]) for B in bases]
    a, da = mp2_adjoint_set(bases, directions, tol=1e-14)
    h, dh = mp2_adjoint_set(bases, directions, tol=1e-14, hermitian=True)
    for b in range(len(bases)):
        scale = max(1.0, float(np.abs(a[b]).max()))
        assert np.abs(a[b] - h[b]).max() <= 1e-11 * scale
        assert float((da[b].G - dh[b].G).abs().max().cpu()) <= 1e-11 * max(1.0, float(da[b].G.abs().max().cpu()))
        assert float((da[b].lam - dh[b].lam).abs().max().cpu()) <= 1e-12
