"""Spatial (spin-adapted) oracle against the spin-orbital oracle and the
reference's golden energies.  CPU only."""
import numpy as np
import pytest

from oracle import spatial, so, transforms
from tests.test_oracle_golden import GOLD_CID_H2O, GOLD_MP2_H2O, h2o, h22  # noqa: F401
from tests.helpers import PARAMS


def _inputs(sys_):
    mol, bs, S, T, V, G, En, E, C, eps = sys_
    nbf = C.shape[0]
    C_list = [slice(0, 0), slice(0, 0), slice(0, 0), slice(0, nbf)]
    eri_mo = transforms.compute_ERI_MO(G, C, C_list)
    F_mo = transforms.compute_F_MO(T + V, G, C, [slice(0, 0), slice(0, mol.Z.sum().astype(int) // 2),
                                                  slice(0, 0), slice(0, nbf)], False)[0]
    return E + En, eri_mo, F_mo, eps


def test_spatial_mp2_golden(h2o):
    Eref, eri_mo, F, eps = _inputs(h2o)
    E, _ = spatial.solve_mp2(eri_mo.swapaxes(1, 2), eps, 5)
    assert abs(Eref + E - GOLD_MP2_H2O) < 1e-11


def test_spatial_cid_golden(h2o):
    Eref, eri_mo, F, eps = _inputs(h2o)
    E, t2, _ = spatial.solve_cid(F, eri_mo.swapaxes(1, 2), eps, 5, PARAMS)
    assert abs(Eref + E - GOLD_CID_H2O) < 1e-11


@pytest.mark.parametrize("which", ["h2o", "h22"])
def test_spatial_cisd_equals_alpha_beta_block_of_spin_orbital(which, request):
    sys_ = request.getfixturevalue(which)
    Eref, eri_mo, F, eps = _inputs(sys_)
    O = int(sys_[0].Z.sum()) // 2
    E, t1, t2, _ = spatial.solve_cisd(F, eri_mo.swapaxes(1, 2), eps, O, PARAMS)
    eri_so = transforms.compute_ERI_SO(eri_mo).swapaxes(1, 2)
    F_so = transforms.compute_F_SO(F)
    e = np.repeat(eps, 2)
    Es, s1, s2, _ = so.solve_cisd_so(F_so, eri_so, e[:2 * O], e[2 * O:], 2 * O, PARAMS)
    assert abs(E - Es) < 1e-11
    assert np.abs(t1 - s1[0::2, 0::2]).max() < 1e-9
    assert np.abs(t2 - s2[0::2, 1::2, 0::2, 1::2]).max() < 1e-9


def _spin_expand(t1, t2):
    """Spin-orbital amplitudes (spin orbital 2p + spin) from closed-shell ones:
    the map apy_sp_amplitudes_to_so applies on the device."""
    o, v = t1.shape
    O, V = 2 * o, 2 * v
    s_o, s_v = np.arange(O) % 2, np.arange(V) % 2
    T1 = np.where(s_o[:, None] == s_v[None, :], np.repeat(np.repeat(t1, 2, 0), 2, 1), 0)
    x = np.repeat(np.repeat(np.repeat(np.repeat(t2, 2, 0), 2, 1), 2, 2), 2, 3)
    y = x.swapaxes(2, 3)
    same_ia = (s_o[:, None, None, None] == s_v[None, None, :, None]) & (s_o[None, :, None, None] == s_v[None, None, None, :])
    cross = (s_o[:, None, None, None] == s_v[None, None, None, :]) & (s_o[None, :, None, None] == s_v[None, None, :, None])
    # x[I,J,A,B] = t2[i,j,a,b]; y[I,J,A,B] = t2[i,j,b,a] (the repeat keeps the spatial pattern per spin)
    T2 = np.where(same_ia, x, 0) - np.where(cross, y, 0)
    return T1, T2


@pytest.mark.parametrize("singles", [True, False])
def test_spin_orbital_solution_is_the_spin_expanded_spatial_one(singles):
    """For an RHF-type (spin-free) Hamiltonian the spin-orbital CISD / CID
    solution (ci_wfn.py:171-416) is the spatial one (:420-574, :51-167)
    expanded to spin orbitals: T_IJAB = t_ijab for opposite spins matched
    (I,A),(J,B), -t_ijba crosswise, t_ijab - t_ijba for equal spins."""
    T = transforms
    o, v = 2, 3
    rng = np.random.default_rng(3)
    n = o + v
    E4 = 0.05 * (rng.standard_normal((n,) * 4) + 1j * rng.standard_normal((n,) * 4))
    E4 = 0.5 * (E4 + E4.transpose(2, 3, 0, 1))               # (pq|rs) = (rs|pq), complex orbitals
    eps = np.concatenate([np.sort(rng.uniform(-2, -0.5, o)), np.sort(rng.uniform(0.3, 2, v))])
    F = np.diag(eps).astype(complex) + 1e-3 * (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    eri = E4.swapaxes(1, 2)
    Fso, eri_so = T.compute_F_SO(F), T.compute_ERI_SO(E4).swapaxes(1, 2)
    eps_so = np.repeat(eps, 2)
    if singles:
        Es, t1, t2, _ = spatial.solve_cisd(F, eri, eps, o, PARAMS)
        Eso, T1, T2, _ = so.solve_cisd_so(Fso, eri_so, eps_so[:2 * o], eps_so[2 * o:], 2 * o, PARAMS)
    else:
        Es, t2, _ = spatial.solve_cid(F, eri, eps, o, PARAMS)
        t1 = np.zeros((o, v), dtype=complex)
        Eso, T2, _ = so.solve_cid_so(Fso, eri_so, eps_so[:2 * o], eps_so[2 * o:], 2 * o, PARAMS)
        T1 = np.zeros((2 * o, 2 * v), dtype=complex)
    X1, X2 = _spin_expand(t1, t2)
    assert abs(Es - Eso) < 1e-13
    assert np.abs(X1 - T1).max() < 1e-12
    assert np.abs(X2 - T2).max() < 1e-12
