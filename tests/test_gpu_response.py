"""First-order (linear-response) CISD / CID amplitudes on the GPU
(spatial_ci.ResponseCI: apy_cisd_sp_step with its response term) against
the oracle's restatement of analytic_aats.py:733-890 / :1540-1660
(oracle.analytic.ci_response): one step with a non-symmetric derivative
Fock matrix and non-Hermitian derivative integrals, and whole solves."""
import numpy as np
import pytest
import torch

from apyib_amd.spatial_ci import ResponseCI, SPIntegrals, SpatialCI
from oracle import analytic, spatial
from tests.helpers import PARAMS
from tests.test_gpu_spatial import _amps, _d, _system

pytestmark = pytest.mark.gpu


def _perturbation(o, v, seed, scale=0.05):
    """A pair-symmetric chemists' derivative tensor with no Hermiticity and
    a full non-symmetric derivative Fock matrix."""
    rng = np.random.default_rng(seed)
    n = o + v
    dE = scale * (rng.standard_normal((n,) * 4) + 1j * rng.standard_normal((n,) * 4))
    dE = 0.5 * (dE + dE.transpose(2, 3, 0, 1))
    dF = scale * (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    return dE, dF


def _vec(t1, t2):
    return np.concatenate([t1.ravel(), t2.ravel()])


@pytest.mark.parametrize("o,v,seed", [(2, 3, 1), (4, 7, 2), (5, 13, 3)])
@pytest.mark.parametrize("singles", [True, False])
def test_response_step_matches_oracle(dev, o, v, seed, singles):
    E4, F, eps = _system(o, v, seed)
    dE4, dF = _perturbation(o, v, seed + 20)
    A, dA = E4.swapaxes(1, 2), dE4.swapaxes(1, 2)          # <pq|rs> = (pr|qs)
    t1, t2 = _amps(o, v, seed + 10)
    x1, x2 = _amps(o, v, seed + 30)
    if not singles:
        t1 = np.zeros_like(t1)
        x1 = np.zeros_like(x1)
    Ec, s = -0.07 + 0.002j, 0.013 - 0.004j
    ints = SPIntegrals.from_mo(_d(E4, dev), o, singles=singles)
    dints = SPIntegrals.from_mo(_d(dE4, dev), o, singles=singles)
    r = ResponseCI(ints, _d(F, dev), eps, singles, _d(_vec(t1, t2), dev), Ec, dints, _d(dF, dev))
    e0 = spatial.cisd_energy(dF, dA, t1, t2, o)
    assert abs(r.shift - e0) < 1e-13 * max(1.0, abs(e0))
    x = _d(_vec(x1, x2), dev)
    xn, out = r.new_vector(), r.new_vector()
    r.step(x, xn, out, energy=torch.tensor([s], dtype=torch.complex128, device=dev))
    g1, g2 = (y.cpu().numpy() for y in r.split(out))
    q1, q2 = spatial.cisd_residual(dF, dA, t1, t2, s + e0, o)
    p1, p2 = spatial.cisd_residual(F, A, x1, x2, Ec, o)
    r1 = q1 + p1 - F.T[:o, o:]
    r2 = q2 + p2 - spatial._driver(A, o)
    tol = 1e-13 * max(1.0, np.abs(r2).max())
    assert np.abs(g2 - r2).max() < tol
    if singles:
        assert np.abs(g1 - r1).max() < tol
    else:
        assert np.abs(g1).max() == 0.0
    D1, D2 = spatial._denoms(eps, o)
    n1, n2 = (y.cpu().numpy() for y in r.split(xn))
    assert np.abs(n2 - (x2 + r2 / D2)).max() < 1e-12 * max(1.0, np.abs(n2).max())
    if singles:
        assert np.abs(n1 - (x1 + r1 / D1)).max() < 1e-12 * max(1.0, np.abs(n1).max())


@pytest.mark.parametrize("singles", [True, False])
def test_response_solve_matches_oracle(dev, singles):
    """Converged first-order amplitudes and projected energy derivative vs
    the oracle (different starting guesses; the solution is unique)."""
    o, v, seed = 4, 9, 5
    E4, F, eps = _system(o, v, seed)
    dE4, dF = _perturbation(o, v, seed + 20, scale=0.02)
    A, dA = E4.swapaxes(1, 2), dE4.swapaxes(1, 2)
    params = dict(PARAMS)
    ints = SPIntegrals.from_mo(_d(E4, dev), o, singles=singles)
    ci = SpatialCI(ints, _d(F, dev), eps, singles=singles)
    E, _, _, _ = ci.solve(params)
    t_ref = ci.t.clone()
    t1 = ci.split(t_ref)[0].cpu().numpy() if singles else None
    t2 = ci.split(t_ref)[1].cpu().numpy()
    dints = SPIntegrals.from_mo(_d(dE4, dev), o, singles=singles)
    r = ResponseCI(ints, _d(F, dev), eps, singles, t_ref, E, dints, _d(dF, dev))
    dE, g1, g2, it = r.solve(params)
    assert it < params['max_iterations']
    want, w1, w2, _ = analytic.ci_response(F, A, eps, o, E, t1, t2, dF, dA, 0.0, params)
    assert abs(dE - want) < 1e-10 * max(1.0, abs(want))
    assert np.abs(g2.cpu().numpy() - w2).max() < 1e-9 * max(1.0, np.abs(w2).max())
    if singles:
        assert np.abs(g1.cpu().numpy() - w1).max() < 1e-9 * max(1.0, np.abs(w1).max())


def _args(ci, x, xn, out, energy, rb, rt, re):
    from apyib_amd import _lib
    import ctypes
    return _lib.CisdSpArgs(ci.O, ci.V, ci.fock.data_ptr(), ci.eps.data_ptr(), ctypes.pointer(ci.ints.struct),
                           ci.hoo.data_ptr(), ci.g.data_ptr(), int(ci.singles), energy.data_ptr(), x.data_ptr(),
                           xn.data_ptr(), out.data_ptr(), rb.data_ptr(), rt.data_ptr(), re.data_ptr(), 0.0, 0.0)


@pytest.mark.parametrize("singles", [True, False])
def test_batched_step_and_energy_match_single_calls(dev, singles):
    """apy_cisd_sp_step_batch over 3 instances (own vectors, response
    vectors, energies, response scalars and shifts) against three
    apy_cisd_sp_step calls; apy_cisd_sp_energy_batch against
    apy_cisd_sp_energy."""
    import ctypes
    from apyib_amd import _lib
    from apyib_amd.kernels import stream_of, workspace
    lib = _lib.load()
    o, v, nb = 4, 11, 3
    E4, F, eps = _system(o, v, 8)
    ci = SpatialCI(SPIntegrals.from_mo(_d(E4, dev), o, singles=singles), _d(F, dev), eps, singles=singles)
    L = ci.L

    def vec(seed):
        a1, a2 = _amps(o, v, seed)
        return _d(_vec(a1 if singles else 0 * a1, a2), dev)

    X = torch.stack([vec(40 + q) for q in range(nb)])
    B = torch.stack([vec(50 + q) for q in range(nb)])
    rt = vec(60)
    en = torch.tensor([0.01 - 0.02j, -0.03 + 0.001j, 0.2], dtype=torch.complex128, device=dev)
    re = torch.tensor([0.004j, -0.01, 0.02 + 0.02j], dtype=torch.complex128, device=dev)
    sh = torch.tensor([0.5, -0.25j, 0.1 + 0.1j], dtype=torch.complex128, device=dev)
    Xn, R = torch.zeros_like(X), torch.zeros_like(X)
    a = _args(ci, X, Xn, R, en, B, rt, re)
    nbytes = int(lib.apy_cisd_sp_batch_workspace(o, v, int(singles), nb))
    ws = workspace(nbytes, dev)
    assert lib.apy_cisd_sp_step_batch(ctypes.byref(a), nb, ctypes.c_void_p(sh.data_ptr()),
                                      ctypes.c_void_p(ws.data_ptr()), nbytes, stream_of(dev)) == 0
    for q in range(nb):
        xn, r = torch.zeros(L, dtype=torch.complex128, device=dev), torch.zeros(L, dtype=torch.complex128, device=dev)
        req = re[q:q + 1] + sh[q:q + 1]
        a1 = _args(ci, X[q], xn, r, en[q:q + 1].clone(), B[q].clone(), rt, req)
        n1 = int(lib.apy_cisd_sp_workspace(o, v, int(singles)))
        w1 = workspace(n1, dev)
        assert lib.apy_cisd_sp_step(ctypes.byref(a1), ctypes.c_void_p(w1.data_ptr()), n1, stream_of(dev)) == 0
        scale = max(1.0, r.abs().max().item())
        assert (R[q] - r).abs().max().item() < 1e-13 * scale
        assert (Xn[q] - xn).abs().max().item() < 1e-12 * max(1.0, xn.abs().max().item())
    energies = torch.zeros(nb, dtype=torch.complex128, device=dev)
    nb_bytes = nb * int(lib.apy_cisd_sp_energy_workspace())
    ws = workspace(nb_bytes, dev)
    assert lib.apy_cisd_sp_energy_batch(o, v, int(singles), nb, ctypes.c_void_p(ci.g.data_ptr()),
                                        ctypes.c_void_p(ci.ints["l"].data_ptr()), ctypes.c_void_p(X.data_ptr()),
                                        ctypes.c_void_p(energies.data_ptr()), ctypes.c_void_p(ws.data_ptr()),
                                        nb_bytes, stream_of(dev)) == 0
    for q in range(nb):
        e1 = ci.energy(X[q], torch.zeros(1, dtype=torch.complex128, device=dev)).item()
        assert abs(energies[q].item() - e1) < 1e-13 * max(1.0, abs(e1))


@pytest.mark.parametrize("singles", [True, False])
def test_response_set_matches_single_solves(dev, singles):
    """Three perturbations iterated as one ResponseSet against three
    ResponseCI solves: same projected derivatives, amplitudes and iteration
    counts."""
    from apyib_amd.spatial_ci import ResponseSet
    o, v, seed = 4, 9, 5
    E4, F, eps = _system(o, v, seed)
    params = dict(PARAMS)
    ints = SPIntegrals.from_mo(_d(E4, dev), o, singles=singles)
    ci = SpatialCI(ints, _d(F, dev), eps, singles=singles)
    E, _, _, _ = ci.solve(params)
    t_ref = ci.t.clone()
    perts = [_perturbation(o, v, seed + 20 + k, scale=0.02) for k in range(3)]
    rs = ResponseSet(ints, _d(F, dev), eps, singles, t_ref, E)
    for dE4, dF in perts:
        rs.add(SPIntegrals.from_mo(_d(dE4, dev), o, singles=singles), _d(dF, dev))
    dEs, g1, g2, its = rs.solve(params)
    for k, (dE4, dF) in enumerate(perts):
        r = ResponseCI(ints, _d(F, dev), eps, singles, t_ref, E,
                       SPIntegrals.from_mo(_d(dE4, dev), o, singles=singles), _d(dF, dev))
        dE, w1, w2, it = r.solve(params)
        assert its[k] == it
        assert abs(dEs[k] - dE) < 1e-11 * max(1.0, abs(dE))
        assert (g2[k] - w2).abs().max().item() < 1e-10 * max(1.0, w2.abs().max().item())
        if singles:
            assert (g1[k] - w1).abs().max().item() < 1e-10 * max(1.0, w1.abs().max().item())


def test_segmented_reductions_are_bitwise_the_single_calls(dev):
    """apy_zdot_many_seg, apy_zdiffsq_batch and apy_zcombine_seg on 5
    segments against apy_zdot_many, apy_zdiffsq_segments and apy_zcombine
    on each segment (bit-identical), and the dots against NumPy."""
    import ctypes
    from apyib_amd import _lib
    from apyib_amd.kernels import stream_of, workspace
    lib = _lib.load()
    nseg, seglen, split, k = 5, 70001, 1234, 3
    g = torch.Generator(device=dev)
    g.manual_seed(11)
    xs = [torch.randn(nseg * seglen, generator=g, device=dev, dtype=torch.complex128) for _ in range(k)]
    y = torch.randn(nseg * seglen, generator=g, device=dev, dtype=torch.complex128)
    st = stream_of(dev)
    P = lambda t: ctypes.c_void_p(t.data_ptr())
    red = int(lib.apy_reduce_seg_workspace(2 * nseg, k))
    ws = workspace(red, dev)
    arr = (ctypes.c_void_p * k)(*[x.data_ptr() for x in xs])
    out = torch.zeros((nseg, k), dtype=torch.complex128, device=dev)
    assert lib.apy_zdot_many_seg(k, arr, P(y), nseg, seglen, 1, P(out), P(ws), red, st) == 0
    dsq = torch.zeros((nseg, 2), dtype=torch.complex128, device=dev)
    assert lib.apy_zdiffsq_batch(P(xs[0]), P(y), nseg, seglen, split, P(dsq), P(ws), red, st) == 0
    coef = torch.randn((nseg, k), generator=g, device=dev, dtype=torch.complex128)
    comb = torch.zeros(nseg * seglen, dtype=torch.complex128, device=dev)
    assert lib.apy_zcombine_seg(k, arr, P(coef), nseg, seglen, P(comb), st) == 0
    out, dsq, comb = out.cpu().numpy(), dsq.cpu().numpy(), comb.cpu()
    ch = coef.cpu().numpy()
    one = int(lib.apy_reduce_workspace(seglen, k))
    for q in range(nseg):
        sl = slice(q * seglen, (q + 1) * seglen)
        segs = [x[sl] for x in xs]
        a1 = (ctypes.c_void_p * k)(*[x.data_ptr() for x in segs])
        o1 = torch.zeros(k, dtype=torch.complex128, device=dev)
        w1 = workspace(max(one, red), dev)
        assert lib.apy_zdot_many(k, a1, P(y[sl]), seglen, 1, P(o1), P(w1), one, st) == 0
        assert np.array_equal(o1.cpu().numpy(), out[q])
        want = np.array([np.vdot(x.cpu().numpy(), y[sl].cpu().numpy()) for x in segs])
        assert np.abs(out[q] - want).max() < 1e-10 * np.abs(want).max()
        off = (ctypes.c_int64 * 3)(0, split, seglen)
        d1 = torch.zeros(2, dtype=torch.complex128, device=dev)
        assert lib.apy_zdiffsq_segments(P(xs[0][sl]), P(y[sl]), 2, off, P(d1), P(w1), one, st) == 0
        assert np.array_equal(d1.cpu().numpy(), dsq[q])
        c1 = np.ascontiguousarray(ch[q]).view(np.float64)
        y1 = torch.zeros(seglen, dtype=torch.complex128, device=dev)
        assert lib.apy_zcombine(k, a1, c1.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), seglen, P(y1), st) == 0
        assert torch.equal(y1.cpu(), comb[sl])


@pytest.mark.parametrize("singles", [True, False])
def test_gauss3m_response_set_matches_four_product_set(dev, singles):
    """ResponseSet(gauss3m=True) against the four-product set: same
    projected derivatives (1e-10 relative) and amplitudes (1e-9)."""
    from apyib_amd.spatial_ci import ResponseSet
    o, v, seed = 5, 14, 6
    E4, F, eps = _system(o, v, seed)
    params = dict(PARAMS)
    ints = SPIntegrals.from_mo(_d(E4, dev), o, singles=singles)
    ci = SpatialCI(ints, _d(F, dev), eps, singles=singles)
    E, _, _, _ = ci.solve(params)
    t_ref = ci.t.clone()
    perts = [_perturbation(o, v, seed + 30 + k, scale=0.02) for k in range(4)]
    out = []
    for g3 in (False, True):
        rs = ResponseSet(ints, _d(F, dev), eps, singles, t_ref, E, gauss3m=g3)
        for dE4, dF in perts:
            rs.add(SPIntegrals.from_mo(_d(dE4, dev), o, singles=singles), _d(dF, dev))
        dEs, g1, g2, its = rs.solve(params)
        out.append((list(dEs), [x.clone() for x in g1] if singles else None, [x.clone() for x in g2]))
    (d4, a1, a2), (d3, b1, b2) = out
    for k in range(len(perts)):
        assert abs(d3[k] - d4[k]) < 1e-10 * max(1.0, abs(d4[k]))
        assert (b2[k] - a2[k]).abs().max().item() < 1e-9
        if singles:
            assert (b1[k] - a1[k]).abs().max().item() < 1e-9
