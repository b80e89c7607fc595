"""Device-controlled Davidson iteration (CISDDavidson.iterate_device /
solve_device): the projected eigenproblem on the GPU (apy_heev_lowest,
parallel Jacobi in one wavefront) against numpy.linalg.eigh, the device
convergence test (apy_davidson_check), and whole solves, launched directly
and replayed from captured HIP graphs, against the host-controlled Davidson,
the NumPy Davidson oracle and the explicit CISD matrix."""
import ctypes

import numpy as np
import pytest
import torch

from apyib_amd import _lib, synthetic
from apyib_amd.davidson import CISDDavidson
from apyib_amd.kernels import stream_of
from apyib_amd.so_ci import SOCI, SOIntegrals
from oracle import brute, davidson, so

pytestmark = pytest.mark.gpu

CDT = torch.complex128
MV = _lib.APY_MAXVEC


def _vp(t):
    return ctypes.c_void_p(t.data_ptr())


def _heev(H, dev, ld=MV):
    k = H.shape[0]
    G = np.zeros((ld, ld), complex)
    for j in range(k):
        G[j, :j + 1] = H[:j + 1, j]                     # column j of the upper triangle, contiguous
        G[j, j + 1:] = np.nan                           # never read
    g = torch.from_numpy(G).to(dev)
    eig = torch.full((2 * MV + 2,), complex(np.nan, np.nan), dtype=CDT, device=dev)
    lib = _lib.load()
    _lib.check(lib.apy_heev_lowest(k, _vp(g), ld, _vp(eig), stream_of(dev)), "apy_heev_lowest")
    return eig.cpu().numpy()


@pytest.mark.parametrize("k", list(range(1, MV + 1)))
def test_heev_lowest_matches_eigh(dev, k):
    rng = np.random.default_rng(100 + k)
    X = rng.standard_normal((k, k)) + 1j * rng.standard_normal((k, k))
    H = 0.5 * (X + X.conj().T)
    e = _heev(H, dev)
    w = np.linalg.eigvalsh(H)
    theta, y = e[0].real, e[1:1 + k]
    scale = np.abs(w).max()
    assert abs(e[0].imag) == 0.0
    assert abs(theta - w[0]) <= 1e-14 * scale
    assert abs(np.linalg.norm(y) - 1.0) < 1e-14
    assert np.abs(H @ y - theta * y).max() <= 1e-13 * scale
    assert np.all(e[1 + k:1 + MV] == 0)
    assert np.allclose(e[1 + MV:1 + MV + k], -theta * y, rtol=0, atol=1e-15 * scale)
    assert e[1 + MV + k] == 1.0


def test_heev_lowest_special_matrices(dev):
    # diagonal (no rotation needed), zero diagonal, a degenerate lowest pair, a near-zero coupling
    cases = [np.diag([3.0, -1.0, 2.0, -1.5]).astype(complex),
             np.array([[0.0, 2.0 - 1.0j], [2.0 + 1.0j, 0.0]]),
             np.diag([-2.0, -2.0, 1.0]).astype(complex),
             np.array([[-25.0, 1e-290j], [-1e-290j, 3.0]])]
    for H in cases:
        e = _heev(H, dev)
        k = H.shape[0]
        w = np.linalg.eigvalsh(H)
        y = e[1:1 + k]
        assert abs(e[0].real - w[0]) <= 1e-14 * np.abs(w).max()
        assert np.abs(H @ y - e[0].real * y).max() <= 1e-13 * np.abs(w).max()


def test_davidson_check_semantics(dev):
    lib = _lib.load()
    eig = torch.zeros(2 * MV + 2, dtype=CDT, device=dev)
    rr = torch.zeros(1, dtype=CDT, device=dev)
    ctl = torch.tensor([0.0, float("nan"), 0.0, 0.0], dtype=torch.float64, device=dev)
    log = torch.zeros(2 * 3, dtype=torch.float64, device=dev)
    seq = [(-1.0, 1e-2), (-1.0 - 1e-14, 1e-6), (-1.0 - 1e-14, 1e-10), (-1.0 - 2e-14, 1e-11)]
    status = []
    for theta, rn in seq:
        eig[0] = theta
        rr[0] = rn * rn
        _lib.check(lib.apy_davidson_check(_vp(eig), _vp(rr), 1e-12, 1e-8, _vp(ctl), _vp(log), 3, stream_of(dev)),
                   "apy_davidson_check")
        status.append(float(ctl[2].cpu()))
    assert status == [0.0, 0.0, 1.0, 1.0]
    c = ctl.cpu().numpy()
    assert c[0] == 4 and c[1] == seq[-1][0] and abs(c[3] - 1e-11) < 1e-24
    assert np.allclose(log.cpu().numpy().reshape(3, 2), np.array(seq[:3]), rtol=1e-15, atol=0)


@pytest.mark.parametrize("O,V,k", [(2, 3, 1), (4, 6, 5), (6, 14, 15)])
def test_fused_residual_and_gram_schmidt_passes(dev, O, V, k):
    """apy_davidson_residual / apy_davidson_combine_dots against the unfused
    calls (apy_zcombine_seg, apy_ci_precondition_packed, torch dots)."""
    lib = _lib.load()
    st = stream_of(dev)
    Lv = int(lib.apy_ci_packed_length(O, V))
    g = torch.Generator(device="cpu").manual_seed(O * 100 + k)
    rnd = lambda: torch.complex(torch.randn(Lv, generator=g, dtype=torch.float64),
                                torch.randn(Lv, generator=g, dtype=torch.float64)).to(dev)
    S = [rnd() for _ in range(k)]
    B = [rnd() for _ in range(k)]
    eps = torch.from_numpy(synthetic.fock(O, V, 3)[1]).to(dev)
    theta = -0.37
    y = torch.complex(torch.randn(k, generator=g, dtype=torch.float64), torch.randn(k, generator=g, dtype=torch.float64))
    eig = torch.zeros(2 * MV + 2, dtype=CDT)
    eig[0] = theta
    eig[1:1 + k] = y
    eig = eig.to(dev)
    r_ref = sum(y[j].item() * (S[j] - theta * B[j]) for j in range(k))
    d_ref = torch.zeros_like(r_ref)
    _lib.check(lib.apy_ci_precondition_packed(O, V, _vp(eps), theta, _vp(r_ref), _vp(d_ref), st), "precondition")
    d = torch.full_like(r_ref, complex(np.nan, np.nan))
    dots = torch.zeros(MV + 2, dtype=CDT, device=dev)
    nb = int(lib.apy_davidson_fused_workspace())
    ws = torch.empty(nb, dtype=torch.uint8, device=dev)
    arr = lambda xs: (ctypes.c_void_p * len(xs))(*[x.data_ptr() for x in xs])
    _lib.check(lib.apy_davidson_residual(k, arr(S), arr(B), _vp(eig), _vp(eps), O, V, _vp(d), _vp(dots), _vp(ws), nb,
                                         st), "apy_davidson_residual")
    scale = float(d_ref.abs().max().cpu())
    assert float((d - d_ref).abs().max().cpu()) <= 1e-13 * scale
    got = dots.cpu().numpy()
    want = [complex(torch.vdot(B[j], d_ref).cpu()) for j in range(k)] + [complex(torch.vdot(d_ref, d_ref).cpu()),
                                                                          complex(torch.vdot(r_ref, r_ref).cpu())]
    assert np.allclose(got[:k + 2], want, rtol=1e-12, atol=1e-12 * abs(want[k]))
    coef = torch.complex(torch.randn(k + 1, generator=g, dtype=torch.float64),
                         torch.randn(k + 1, generator=g, dtype=torch.float64)).to(dev)
    v_ref = sum(coef[j].item() * B[j] for j in range(k)) + coef[k].item() * d
    _lib.check(lib.apy_davidson_combine_dots(k, arr(B), _vp(coef), Lv, _vp(d), _vp(dots), _vp(ws), nb, st),
               "apy_davidson_combine_dots")
    assert float((d - v_ref).abs().max().cpu()) <= 1e-13 * float(v_ref.abs().max().cpu())
    want = [complex(torch.vdot(B[j], v_ref).cpu()) for j in range(k)] + [complex(torch.vdot(v_ref, v_ref).cpu())]
    assert np.allclose(dots.cpu().numpy()[:k + 1], want, rtol=1e-12, atol=1e-12 * abs(want[k]))


def _system(O, V, seed, dev):
    F, A, eps = synthetic.system(O, V, seed)
    ints = SOIntegrals.synthetic(O, V, seed, device=dev)
    return F, A, eps, ints


def _dav(F, eps, ints, dev, max_space):
    return CISDDavidson(SOCI(ints, torch.from_numpy(F).to(dev), eps, singles=True, fov_coupling=True),
                        max_space=max_space)


@pytest.mark.parametrize("graph,fused", [(False, True), (True, True), (False, False), (True, False)])
@pytest.mark.parametrize("O,V,seed,max_space", [(6, 14, 4, 15), (8, 20, 5, 6), (4, 33, 6, 4)])
def test_device_davidson_matches_host_control(dev, O, V, seed, max_space, graph, fused):
    F, A, eps, ints = _system(O, V, seed, dev)
    host = _dav(F, eps, ints, dev, max_space)
    Eh = host.solve(e_conv=1e-13, r_conv=1e-9)
    t1h, t2h = host.amplitudes()
    d = _dav(F, eps, ints, dev, max_space)
    Ed = d.solve_device(e_conv=1e-13, r_conv=1e-9, chunk=3, graph=graph, fused=fused)
    t1d, t2d = d.amplitudes()                            # intermediate normalisation: phase-free
    print("host", host.iterations, "device", d.iterations, Eh, Ed)
    assert abs(Ed - Eh) < 1e-11
    assert float((t1d - t1h).abs().max().cpu()) < 1e-8
    assert float((t2d - t2h).abs().max().cpu()) < 1e-8
    assert d.iterations <= host.iterations + 3
    hist = np.array(d.history)
    assert hist.shape[0] == d.iterations and hist[-1, 1] < 1e-9


def test_device_davidson_matches_oracle_and_explicit_matrix(dev):
    O, V, seed = 3, 4, 2
    F, A, eps, ints = _system(O, V, seed, dev)
    E_ref, t1_ref, t2_ref = brute.lowest_root(F, A, O)
    W = so.blocks_from_eri(A / 2, O)
    E_dav, _, _ = davidson.davidson(F, W, O, tol=1e-13, rtol=1e-10)
    for graph in (False, True):
        d = _dav(F, eps, ints, dev, 8)
        E = d.solve_device(e_conv=1e-13, r_conv=1e-9, graph=graph)
        t1, t2 = d.amplitudes()
        assert abs(E - E_ref) < 1e-11 and abs(E - E_dav) < 1e-11
        assert np.abs(t1.cpu().numpy() - t1_ref).max() < 1e-8
        assert np.abs(t2.cpu().numpy() - t2_ref).max() < 1e-8


@pytest.mark.parametrize("device_control,graph", [(False, False), (True, False), (True, True)])
def test_combined_restart_matches_fresh_sigma_restart(dev, device_control, graph):
    """restart="combine" (the Ritz vector's sigma formed as sum_j y_j S_j at
    a restart) against the fresh-sigma restart with a small subspace (many
    restarts): the same root to 1e-11 Eh and amplitudes to 1e-8, one sigma
    fewer per restart."""
    O, V, seed, M = 8, 20, 5, 4
    F, A, eps, ints = _system(O, V, seed, dev)
    res = {}
    for mode in ("sigma", "combine"):
        d = CISDDavidson(SOCI(ints, torch.from_numpy(F).to(dev), eps, singles=True, fov_coupling=True),
                         max_space=M, restart=mode)
        if device_control:
            E = d.solve_device(e_conv=1e-13, r_conv=1e-9, chunk=3, graph=graph)
        else:
            E = d.solve(e_conv=1e-13, r_conv=1e-9)
        res[mode] = (E, d.amplitudes(), d.nsigma, d.iterations)
    (Es, (t1s, t2s), ns, its), (Ec, (t1c, t2c), nc, itc) = res["sigma"], res["combine"]
    print("iterations", its, itc, "sigma", ns, nc)
    assert abs(Ec - Es) < 1e-11
    assert float((t1c - t1s).abs().max().cpu()) < 1e-8
    assert float((t2c - t2s).abs().max().cpu()) < 1e-8
    assert itc <= its + 3
    assert nc < ns
