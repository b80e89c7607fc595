"""Davidson lowest root of the Hermitian CISD matrix on the GPU and its
adjoint gradient dE/dF: against the explicit CISD matrix, the NumPy
Davidson oracle, the reference's Jacobi iteration (canonical case) and
central finite differences."""
import numpy as np
import pytest
import torch

from apyib_amd import synthetic
from apyib_amd.davidson import CISDDavidson, cisd_energy
from apyib_amd.so_ci import SOCI, SOIntegrals
from oracle import brute, davidson, so
from tests.helpers import PARAMS, brillouin

pytestmark = pytest.mark.gpu


def _d(x, dev):
    return torch.from_numpy(np.ascontiguousarray(x)).to(dev)


def _system(O, V, seed, dev, noise=0.05):
    F, A, eps = synthetic.system(O, V, seed, noise=noise)
    ints = SOIntegrals.synthetic(O, V, seed, device=dev)
    return F, A, eps, ints


def _dav(F, eps, ints, dev, max_space=16):
    return CISDDavidson(SOCI(ints, _d(F, dev), eps, singles=True, fov_coupling=True), max_space=max_space)


@pytest.mark.parametrize("O,V,seed", [(2, 3, 1), (3, 4, 2), (4, 4, 3)])
def test_lowest_root_matches_explicit_matrix(dev, O, V, seed):
    F, A, eps, ints = _system(O, V, seed, dev)
    E_ref, t1_ref, t2_ref = brute.lowest_root(F, A, O)
    dav = _dav(F, eps, ints, dev)
    E = dav.solve(e_conv=1e-13, r_conv=1e-8)
    t1, t2 = dav.amplitudes()
    assert abs(E - E_ref) < 1e-11
    assert np.abs(t1.cpu().numpy() - t1_ref).max() < 1e-8
    assert np.abs(t2.cpu().numpy() - t2_ref).max() < 1e-8


@pytest.mark.parametrize("O,V,seed,max_space", [(6, 14, 4, 16), (8, 20, 5, 6)])
def test_matches_numpy_davidson(dev, O, V, seed, max_space):
    """Medium sizes (with restarts for max_space=6) against the oracle."""
    F, A, eps, ints = _system(O, V, seed, dev)
    W = so.blocks_from_eri(A / 2, O)
    E_ref, (c0, c1, c2), _ = davidson.davidson(F, W, O, tol=1e-13, rtol=1e-10)
    dav = _dav(F, eps, ints, dev, max_space)
    E = dav.solve(e_conv=1e-13, r_conv=1e-8)
    t1, t2 = dav.amplitudes()
    print("history", dav.iterations, dav.history[-5:], "oracle", E_ref)
    assert abs(E - E_ref) < 1e-10
    assert np.abs(t1.cpu().numpy() - c1 / c0).max() < 1e-8
    assert np.abs(t2.cpu().numpy() - c2 / c0).max() < 1e-8


def test_canonical_case_equals_reference_iteration(dev):
    """F_ov = 0: the Hermitian root is the reference's Jacobi+DIIS solution."""
    O, V, seed = 4, 8, 6
    F, A, eps, ints = _system(O, V, seed, dev, noise=1e-3)
    F = brillouin(F, O)
    E_ref, _, _, _ = so.solve_cisd_so(F, A / 2, eps[:O], eps[O:], O, PARAMS)
    dav = _dav(F, eps, ints, dev)
    assert abs(dav.solve(e_conv=1e-13, r_conv=1e-8) - E_ref) < 1e-10


def test_fock_gradient_matches_oracle_density(dev):
    O, V, seed = 5, 9, 7
    F, A, eps, ints = _system(O, V, seed, dev)
    W = so.blocks_from_eri(A / 2, O)
    _, c, _ = davidson.davidson(F, W, O, tol=1e-13, rtol=1e-11)
    D_ref = davidson.one_particle_density(c, O)
    dav = _dav(F, eps, ints, dev)
    dav.solve(e_conv=1e-13, r_conv=1e-9)
    D = dav.gradient_fock().cpu().numpy()
    # D is invariant to the phase of the vector and first order in its error
    # (the residual of either solve stays above ~1e-10)
    assert np.abs(D - D_ref).max() < 2e-8


def test_fock_gradient_matches_central_differences(dev):
    """dE along Hermitian one-electron directions (real symmetric and
    i x antisymmetric, as in the nuclear and field perturbations), h = 1e-4."""
    O, V, seed, h = 6, 12, 8, 1e-4
    F, A, eps, ints = _system(O, V, seed, dev)
    dav = _dav(F, eps, ints, dev)
    dav.solve(e_conv=1e-13, r_conv=1e-8)
    D = dav.gradient_fock()
    n = O + V
    for k, kind in enumerate(('nuclear', 'field', 'nuclear', 'field')):
        dF = synthetic.fock_perturbation(n, 50 + k, kind)
        an = complex(torch.sum(D * _d(dF, dev)).cpu())
        Ep = _dav(F + h * dF, eps, ints, dev).solve(e_conv=1e-13, r_conv=1e-8)
        Em = _dav(F - h * dF, eps, ints, dev).solve(e_conv=1e-13, r_conv=1e-8)
        fd = (Ep - Em) / (2 * h)
        scale = float(torch.sum((D * _d(dF, dev)).abs()).cpu())
        assert abs(an.imag) < 1e-10 * scale
        assert abs(fd - an.real) < 1e-7 * scale, (kind, fd, an)


def test_autograd_function(dev):
    O, V, seed = 4, 7, 9
    F, A, eps, ints = _system(O, V, seed, dev)
    Ft = _d(F, dev).requires_grad_(True)
    E = cisd_energy(Ft, ints, eps)
    (g,) = torch.autograd.grad(E, [Ft])
    dav = _dav(F, eps, ints, dev)
    dav.solve(e_conv=1e-12, r_conv=1e-9)
    D = dav.gradient_fock()
    assert torch.abs(g - D.conj()).max().item() < 1e-8


def _batch_sigma(md, entries):
    md._sigma_batch(entries)


def test_batched_sigma_equals_single_instance(dev):
    """apy_cisd_so_sigma_batch (stacked PPL/ring/vvvo GEMMs) against the
    single-instance sigma, three instances with different Fock matrices."""
    from apyib_amd.davidson import MultiDavidson
    O, V, seed = 4, 9, 11
    F, A, eps, ints = _system(O, V, seed, dev)
    n = O + V
    focks = [_d(F + 1e-2 * synthetic.fock_perturbation(n, 60 + k, 'field'), dev) for k in range(3)]
    md = MultiDavidson(ints, focks, eps, max_space=4)
    rng = np.random.default_rng(3)
    L = md.L
    for k, d in enumerate(md.dav):
        v = rng.standard_normal(L + 1) + 1j * rng.standard_normal(L + 1)
        t2 = v[O * V:L].reshape(O, O, V, V)
        t2 = t2 - t2.transpose(1, 0, 2, 3)
        t2 = t2 - t2.transpose(0, 1, 3, 2)
        v[O * V:L] = t2.ravel()
        d.B[1].copy_(_d(v, dev))
    _batch_sigma(md, [(d, 1) for d in md.dav])
    for d in md.dav:
        ref = torch.zeros(L, dtype=torch.complex128, device=dev)
        d.soci.sigma(d.B[1][:L], ref, c0=complex(d.B[1][L].cpu()))
        assert torch.abs(d.S[1][:L] - ref).max().item() < 1e-12


def test_multi_davidson_equals_separate_solves(dev):
    from apyib_amd.davidson import MultiDavidson
    O, V, seed = 5, 12, 12
    F, A, eps, ints = _system(O, V, seed, dev)
    n = O + V
    focks = [_d(F + 1e-3 * synthetic.fock_perturbation(n, 70 + k, kind), dev)
             for k, kind in enumerate(('field', 'nuclear', 'field'))]
    md = MultiDavidson(ints, focks, eps, max_space=6)
    E = md.solve(e_conv=1e-13, r_conv=1e-8)
    for k, Fk in enumerate(focks):
        dav = CISDDavidson(SOCI(ints, Fk, eps, singles=True, fov_coupling=True), max_space=6)
        assert abs(dav.solve(e_conv=1e-13, r_conv=1e-8) - E[k]) < 1e-11


def test_field_set_batched_fd_matches_adjoint(dev):
    """A +/- field-displaced instance set (3 components, h = 1e-4) solved as
    one batch: its central differences against the adjoint gradient of the
    unperturbed root."""
    from apyib_amd.davidson import MultiDavidson
    O, V, seed, h = 6, 14, 13, 1e-4
    F, A, eps, ints = _system(O, V, seed, dev)
    n = O + V
    dirs = [synthetic.fock_perturbation(n, 80 + k, 'field') for k in range(3)]
    focks = [_d(F + sgn * h * dF, dev) for dF in dirs for sgn in (+1, -1)]
    E = MultiDavidson(ints, focks, eps).solve(e_conv=1e-13, r_conv=1e-8)
    dav = _dav(F, eps, ints, dev)
    dav.solve(e_conv=1e-13, r_conv=1e-8)
    D = dav.gradient_fock()
    for k, dF in enumerate(dirs):
        fd = (E[2 * k] - E[2 * k + 1]) / (2 * h)
        an = complex(torch.sum(D * _d(dF, dev)).cpu())
        scale = float(torch.sum((D * _d(dF, dev)).abs()).cpu())
        assert abs(fd - an.real) < 1e-7 * scale


BLOCKS = ("vvvv_p", "ring", "oooo_p", "ovoo_p", "vvvo_p", "vovv_p", "ooov", "vvoo_t", "oovv")


def _gpu_vector(c, dev):
    c0, c1, c2 = c
    return _d(np.concatenate([c1.ravel(), c2.ravel(), [c0]]).astype(np.complex128), dev)


@pytest.mark.parametrize("O,V,seed", [(2, 3, 5), (4, 6, 9), (6, 11, 10)])
def test_tpd_blocks_match_oracle_on_same_vector(dev, O, V, seed):
    """apy_cisd_so_tpd and apy_so_ints_grad_to_full on the oracle's own
    normalised vector, block by block and as the whole tensor."""
    import ctypes
    from apyib_amd import _lib
    from apyib_amd.davidson import full_gradient
    from apyib_amd.kernels import stream_of, workspace
    F, A, eps, ints = _system(O, V, seed, dev)
    _, c, _ = davidson.davidson(F, so.blocks_from_eri(A / 2, O), O, tol=1e-13, rtol=1e-11)
    G_ref = davidson.two_particle_density(c, O)
    ref = davidson.tpd_blocks(G_ref, O)
    lib = _lib.load()
    G = SOIntegrals(O, V, dev)
    for blk in BLOCKS:                     # every entry must be written
        G[blk].fill_(float("nan"))
    cv = _gpu_vector(c, dev)
    nb = int(lib.apy_cisd_so_tpd_workspace(O, V))
    ws = workspace(nb, dev)
    _lib.check(lib.apy_cisd_so_tpd(O, V, ctypes.c_void_p(cv.data_ptr()), ctypes.byref(G.struct),
                                   ctypes.c_void_p(ws.data_ptr()), nb, stream_of(dev)), "apy_cisd_so_tpd")
    for blk in BLOCKS:
        got = G[blk].cpu().numpy()
        want = ref[blk].ravel()
        assert np.abs(got - want).max() < 1e-13 * max(1.0, np.abs(want).max()), blk
    Gf = full_gradient(G).cpu().numpy()
    assert np.abs(Gf - G_ref).max() < 1e-13


def test_tpd_energy_identity_and_oracle(dev):
    """GPU Davidson root: E = sum F D + sum_blocks ints . grad (the
    reference sketch's check, ci_wfn.py:411-414) and G against the oracle
    density of the oracle's root."""
    O, V, seed = 5, 9, 7
    F, A, eps, ints = _system(O, V, seed, dev)
    _, c, _ = davidson.davidson(F, so.blocks_from_eri(A / 2, O), O, tol=1e-13, rtol=1e-11)
    G_ref = davidson.two_particle_density(c, O)
    dav = _dav(F, eps, ints, dev)
    E = dav.solve(e_conv=1e-13, r_conv=1e-9)
    D = dav.gradient_fock()
    G = dav.gradient_integrals()
    two = sum(complex((ints[b] * G[b]).sum().cpu()) for b in BLOCKS)
    one = complex((_d(F, dev) * D).sum().cpu())
    assert abs(one + two - E) < 1e-11
    from apyib_amd.davidson import full_gradient
    # first order in the vector error (residual floor ~1e-10)
    assert np.abs(full_gradient(G).cpu().numpy() - G_ref).max() < 2e-8


def test_integral_gradient_matches_central_differences(dev):
    """dE along a Hermitian, antisymmetric direction of <pq||rs> (the real
    displacement draw of the synthetic generator): adjoint sum G dA against
    central differences of the GPU Davidson energy."""
    O, V, seed = 4, 8, 12
    F, A, eps, _ = _system(O, V, seed, dev)
    n = O + V
    dA = synthetic.eri_perturbation_full(n, seed + 100)
    Ft = _d(F, dev)

    def energy(X):
        ints = SOIntegrals.from_full(_d(X, dev), O)
        return CISDDavidson(SOCI(ints, Ft, eps, singles=True, fov_coupling=True)).solve(e_conv=1e-13, r_conv=1e-9)

    h = 1e-4
    fd = (energy(A + h * dA) - energy(A - h * dA)) / (2 * h)
    ints = SOIntegrals.from_full(_d(A, dev), O)
    dav = CISDDavidson(SOCI(ints, Ft, eps, singles=True, fov_coupling=True))
    dav.solve(e_conv=1e-13, r_conv=1e-9)
    from apyib_amd.davidson import full_gradient
    G = full_gradient(dav.gradient_integrals()).cpu().numpy()
    adj = np.sum(G * dA)
    assert abs(adj.imag) < 1e-10 * max(1.0, abs(adj))
    assert abs(adj.real - fd) < 1e-7 * max(1.0, abs(fd))


def test_autograd_full_tensor(dev):
    """cisd_energy(F, A, eps, O): backward gives conj(D) and conj(G), and the
    loss gradient reproduces the central difference along (dF, dA)."""
    O, V, seed = 3, 6, 13
    F, A, eps, _ = _system(O, V, seed, dev)
    n = O + V
    Ft = _d(F, dev).requires_grad_(True)
    At = _d(A, dev).requires_grad_(True)
    E = cisd_energy(Ft, At, eps, O)
    E.backward()
    dF = synthetic.fock_perturbation(n, seed, "field")
    dA = synthetic.eri_perturbation_full(n, seed + 7)
    lin = np.sum(np.conj(Ft.grad.cpu().numpy()) * dF) + np.sum(np.conj(At.grad.cpu().numpy()) * dA)
    h = 1e-4
    Ep = float(cisd_energy(_d(F + h * dF, dev), _d(A + h * dA, dev), eps, O))
    Em = float(cisd_energy(_d(F - h * dF, dev), _d(A - h * dA, dev), eps, O))
    assert abs(lin.real - (Ep - Em) / (2 * h)) < 1e-7 * max(1.0, abs(lin))


def test_packed_vectors_round_trip_and_metric(dev):
    """apy_ci_pack / apy_ci_unpack on an antisymmetric [c1 | c2 | c0]: the
    round trip is exact, the Euclidean product of packed vectors equals the
    CI-metric product of the dense ones, and the packed preconditioner is
    the dense one packed."""
    import ctypes
    from apyib_amd import _lib
    from apyib_amd.kernels import stream_of
    lib = _lib.load()
    O, V = 5, 7
    L = O * V + O * O * V * V
    rng = np.random.default_rng(2)

    def antisym():
        v = rng.standard_normal(L + 1) + 1j * rng.standard_normal(L + 1)
        t2 = v[O * V:L].reshape(O, O, V, V)
        t2 = t2 - t2.transpose(1, 0, 2, 3)
        t2 = t2 - t2.transpose(0, 1, 3, 2)
        v[O * V:L] = t2.ravel()
        return v
    P = lambda t: ctypes.c_void_p(t.data_ptr())
    st = stream_of(dev)
    Lp = int(lib.apy_ci_packed_length(O, V))
    assert Lp == O * V + (O * (O - 1) // 2) * (V * (V - 1) // 2) + 1
    x, y = antisym(), antisym()
    xd, yd = _d(x, dev), _d(y, dev)
    xp, yp = (torch.empty(Lp, dtype=torch.complex128, device=dev) for _ in range(2))
    assert lib.apy_ci_pack(O, V, P(xd), P(xp), st) == 0 and lib.apy_ci_pack(O, V, P(yd), P(yp), st) == 0
    back = torch.empty(L + 1, dtype=torch.complex128, device=dev)
    assert lib.apy_ci_unpack(O, V, P(xp), P(back), st) == 0
    assert torch.equal(back, xd)
    w = np.ones(L + 1)
    w[O * V:L] = 0.25
    ref = np.sum(w * np.conj(x) * y)
    got = torch.sum(torch.conj(xp) * yp).item()
    assert abs(got - ref) < 1e-12 * abs(ref)
    eps = np.sort(rng.uniform(-2, 2, O + V))
    ed = torch.from_numpy(eps).to(dev)
    rd, rp = torch.empty_like(xd), torch.empty_like(xp)
    assert lib.apy_ci_precondition(O, V, P(ed), -0.3, P(xd), P(rd), st) == 0
    assert lib.apy_ci_precondition_packed(O, V, P(ed), -0.3, P(xp), P(rp), st) == 0
    packed_ref = torch.empty_like(xp)
    assert lib.apy_ci_pack(O, V, P(rd), P(packed_ref), st) == 0
    assert torch.equal(rp, packed_ref)
