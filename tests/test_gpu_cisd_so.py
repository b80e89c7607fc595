"""GPU parity of the spin-orbital CISD / CID path against the oracle."""
import numpy as np
import pytest
import torch

from apyib_amd import synthetic
from apyib_amd.so_ci import SOCI, SOIntegrals, npair, so_fock_from_mo
from oracle import brute, so, transforms
from tests.gpu_layouts import expected_blocks
from tests.helpers import PARAMS, brillouin, rand_amplitudes

pytestmark = pytest.mark.gpu

SIZES = [(4, 6, 3), (6, 8, 5), (5, 7, 9), (2, 3, 1)]


def _dev(x, dev):
    return torch.from_numpy(np.ascontiguousarray(x)).to(dev)


@pytest.mark.parametrize("O,V,seed", SIZES)
def test_synthetic_blocks_match_host_generator(dev, O, V, seed):
    A = synthetic.eri_full(O + V, seed)
    exp = expected_blocks(A, O)
    ints = SOIntegrals.synthetic(O, V, seed, device=dev)
    for name, want in exp.items():
        got = ints[name].cpu().numpy()
        assert np.abs(got - want).max() < 1e-15, name


@pytest.mark.parametrize("O,V,seed", SIZES)
def test_blocks_from_full_are_exact_copies(dev, O, V, seed):
    A = synthetic.eri_full(O + V, seed)
    ints = SOIntegrals.from_full(_dev(A, dev), O)
    for name, want in expected_blocks(A, O).items():
        assert np.array_equal(ints[name].cpu().numpy(), want), name


@pytest.mark.parametrize("m,Osp", [(5, 4), (6, 6), (4, 2)])
def test_blocks_from_mo_spin_expansion(dev, m, Osp):
    rng = np.random.default_rng(m)
    eri_mo = rng.standard_normal((m,) * 4) + 1j * rng.standard_normal((m,) * 4)
    eri_so = transforms.compute_ERI_SO(eri_mo).swapaxes(1, 2)        # Dirac <pq|rs>
    A = eri_so - eri_so.swapaxes(2, 3)                                 # <pq||rs> per term
    ints = SOIntegrals.from_mo(_dev(eri_mo, dev), Osp)
    for name, want in expected_blocks(A, Osp).items():
        assert np.abs(ints[name].cpu().numpy() - want).max() < 1e-14, name
    f_mo = rng.standard_normal((m, m)) + 1j * rng.standard_normal((m, m))
    f_so = so_fock_from_mo(_dev(f_mo, dev)).cpu().numpy()
    assert np.array_equal(f_so, transforms.compute_F_SO_loop(f_mo))


def _system(O, V, seed, dev, brill=False):
    F, A, eps = synthetic.system(O, V, seed)
    if brill:
        F = brillouin(F, O)
    ints = SOIntegrals.synthetic(O, V, seed, device=dev)
    return F, A, eps, ints


@pytest.mark.parametrize("O,V,seed", SIZES)
@pytest.mark.parametrize("singles", [True, False])
def test_residual_step_matches_oracle(dev, O, V, seed, singles):
    F, A, eps, ints = _system(O, V, seed, dev)
    t1, t2 = rand_amplitudes(O, V, seed + 1)
    if not singles:
        t1 = np.zeros_like(t1)
    E = 0.013 - 0.002j
    solver = SOCI(ints, _dev(F, dev), eps, singles=singles)
    t = _dev(np.concatenate([t1.ravel(), t2.ravel()]), dev)
    tn = solver.new_vector()
    r = solver.new_vector()
    Ed = torch.tensor([E], dtype=torch.complex128, device=dev)
    solver.step(t, tn, r, energy=Ed)
    D1, D2 = so.denominators(eps[:O], eps[O:])
    if singles:
        r1, r2 = so.cisd_so_residual(F, A / 2, t1, t2, E, O)
    else:
        r1, r2 = np.zeros_like(t1), so.cid_so_residual(F, A / 2, t2, E, O)
    g1, g2 = solver.split(r)
    n1, n2 = solver.split(tn)
    scale = max(1.0, np.abs(r2).max())
    assert np.abs(g2.cpu().numpy() - r2).max() < 1e-13 * scale
    assert np.abs(n2.cpu().numpy() - (t2 + r2 / D2)).max() < 1e-12
    if singles:
        assert np.abs(g1.cpu().numpy() - r1).max() < 1e-13 * scale
        assert np.abs(n1.cpu().numpy() - (t1 + r1 / D1)).max() < 1e-12


@pytest.mark.parametrize("O,V,seed", [(4, 5, 3), (2, 4, 5)])
def test_sigma_mode_hermitian_matches_determinant_ci(dev, O, V, seed):
    """sigma(c0, x) with the f_vo coupling switched on is the CI matrix-vector
    product, including the reference weight c0."""
    F, A, eps, ints = _system(O, V, seed, dev)
    t1, t2 = rand_amplitudes(O, V, seed + 3)
    c0 = 0.7 - 0.2j
    n = O + V
    sp = brute.CISpace(n, O)
    H, Eref = brute.hamiltonian(F, A, sp)
    c = brute.vector_from_amplitudes(sp, t1, t2, c0=c0)
    _, s1, s2 = brute.amplitudes_from_vector(sp, H @ c - Eref * c, V)
    solver = SOCI(ints, _dev(F, dev), eps, singles=True, fov_coupling=True)
    x = _dev(np.concatenate([t1.ravel(), t2.ravel()]), dev)
    out = solver.new_vector()
    solver.sigma(x, out, c0=c0)
    g1, g2 = solver.split(out)
    assert np.abs(g1.cpu().numpy() - s1).max() < 1e-13
    assert np.abs(g2.cpu().numpy() - s2).max() < 1e-13


@pytest.mark.parametrize("O,V,seed", [(4, 6, 3), (6, 8, 5), (2, 5, 2)])
def test_solve_cisd_so_matches_oracle(dev, O, V, seed):
    F, A, eps, ints = _system(O, V, seed, dev)
    E_ref, t1_ref, t2_ref, it_ref = so.solve_cisd_so(F, A / 2, eps[:O], eps[O:], O, PARAMS)
    solver = SOCI(ints, _dev(F, dev), eps, singles=True)
    E, t1, t2, it = solver.solve(PARAMS)
    assert abs(E - E_ref) < 1e-10
    assert np.abs(t1.cpu().numpy() - t1_ref).max() < 1e-9
    assert np.abs(t2.cpu().numpy() - t2_ref).max() < 1e-9
    assert abs(it - it_ref) <= 1


@pytest.mark.parametrize("O,V,seed", [(4, 6, 3), (6, 7, 1)])
def test_solve_cid_so_matches_oracle(dev, O, V, seed):
    F, A, eps, ints = _system(O, V, seed, dev)
    E_ref, t2_ref, it_ref = so.solve_cid_so(F, A / 2, eps[:O], eps[O:], O, PARAMS)
    solver = SOCI(ints, _dev(F, dev), eps, singles=False)
    E, _, t2, it = solver.solve(PARAMS)
    assert abs(E - E_ref) < 1e-10
    assert np.abs(t2.cpu().numpy() - t2_ref).max() < 1e-9


def test_solve_without_diis_matches_oracle(dev):
    O, V, seed = 4, 6, 4
    F, A, eps, ints = _system(O, V, seed, dev)
    p = dict(PARAMS, DIIS=False, max_iterations=400)
    E_ref, t1_ref, t2_ref, _ = so.solve_cisd_so(F, A / 2, eps[:O], eps[O:], O, p)
    E, t1, t2, _ = SOCI(ints, _dev(F, dev), eps).solve(p)
    assert abs(E - E_ref) < 1e-10
    assert np.abs(t2.cpu().numpy() - t2_ref).max() < 1e-9


def test_max_iterations_reached_returns_last_iterate(dev, capsys):
    O, V, seed = 4, 6, 4
    F, A, eps, ints = _system(O, V, seed, dev)
    p = dict(PARAMS, max_iterations=2)
    E_ref, _, t2_ref, _ = so.solve_cisd_so(F, A / 2, eps[:O], eps[O:], O, p)
    E, _, t2, it = SOCI(ints, _dev(F, dev), eps).solve(p)
    assert it == 2
    assert abs(E - E_ref) < 1e-12
    assert "Not converged." in capsys.readouterr().out
