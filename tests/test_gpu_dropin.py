"""The reference's API (energy, ci_wfn, mp2_wfn, utils) running on the GPU
path, against the reference's golden energies and the oracle."""
import numpy as np
import pytest
import torch

import apyib_amd.energy as energy_mod
from apyib_amd import transforms as tr
from apyib_amd import utils
from apyib_amd.ci_wfn import ci_wfn
from apyib_amd.hamiltonian import Hamiltonian
from apyib_amd.hf_wfn import hf_wfn
from apyib_amd.mp2_wfn import mp2_wfn
from oracle import scf, so, spatial, transforms
from tests.molecules import H2O, H2_2
from tests.test_oracle_golden import GOLD_CID_H2O, GOLD_CISD_H2_2, GOLD_MP2_H2O, GOLD_RHF_H2O

pytestmark = pytest.mark.gpu


def params(geom, method, basis='STO-3G'):
    return {'geom': geom, 'basis': basis, 'method': method, 'e_convergence': 1e-12, 'd_convergence': 1e-12,
            'DIIS': True, 'freeze_core': False, 'F_el': [0.0, 0.0, 0.0], 'F_mag': [0.0, 0.0, 0.0],
            'max_iterations': 120}


@pytest.mark.parametrize("method,gold", [("RHF", GOLD_RHF_H2O), ("MP2", GOLD_MP2_H2O), ("MP2_SO", GOLD_MP2_H2O),
                                         ("CID", GOLD_CID_H2O), ("CID_SO", GOLD_CID_H2O)])
def test_energy_h2o_sto3g_golden(dev, method, gold):
    E_list, T_list, C, basis = energy_mod.energy(params(H2O, method))
    assert abs((E_list[0] + E_list[1] + E_list[2]) - gold) < 1e-10


# H2O (apyib/data/molecules.py:2-10) in the split-valence bases: test_002_RHF.py:73 and :98 (RHF, 6-31G(d)),
# test_003_MP2.py:60 and :99 (MP2, 6-31G(d)), test_005_CISD.py:21 and test_008_CISD_SO.py:21 (CISD, 6-31G).
GOLD_RHF_H2O_631GD = -75.97474825544526
GOLD_MP2_H2O_631GD = -76.17533681889488
GOLD_CISD_H2O_631G = -76.09503651362023


@pytest.mark.parametrize("method,basis,gold", [("RHF", "6-31G(d)", GOLD_RHF_H2O_631GD),
                                               ("MP2", "6-31G(d)", GOLD_MP2_H2O_631GD),
                                               ("MP2_SO", "6-31G(d)", GOLD_MP2_H2O_631GD),
                                               ("CISD", "6-31G", GOLD_CISD_H2O_631G),
                                               ("CISD_SO", "6-31G", GOLD_CISD_H2O_631G)])
def test_energy_h2o_split_valence_golden(dev, method, basis, gold):
    E_list, T_list, C, b = energy_mod.energy(params(H2O, method, basis))
    assert abs((E_list[0] + E_list[1] + E_list[2]) - gold) < 1e-10


# H2O/cc-pVDZ (spherical d, general contractions): test_005_CISD.py:59 (CISD), :122 (CISD, frozen core),
# test_003_MP2.py:124 (MP2, frozen core), test_004_CID.py:58-60 (RHF and CID: CFOUR's split), :125 (CID, frozen core)
@pytest.mark.parametrize("method,fc,gold", [("RHF", False, -75.98979581991861),
                                            ("CID", False, -75.98979581991861 - 0.21279410950205),
                                            ("CISD", False, -76.20375874728018), ("CISD", True, -76.20195791202802),
                                            ("MP2", True, -76.20202577979143), ("CID", True, -76.20078548433332),
                                            ("CISD_SO", False, -76.20375874728018),
                                            ("MP2_SO", True, -76.20202577979143)])
def test_energy_h2o_cc_pvdz_golden(dev, method, fc, gold):
    p = dict(params(H2O, method, 'cc-pVDZ'), freeze_core=fc)
    E_list, T_list, C, b = energy_mod.energy(p)
    assert b.nbf() == 24
    assert abs((E_list[0] + E_list[1] + E_list[2]) - gold) < 1e-10


@pytest.mark.parametrize("method", ["CISD_SO", "CISD"])
def test_energy_h2_2_cisd_golden(dev, method):
    E_list, T_list, C, basis = energy_mod.energy(params(H2_2, method))
    assert abs((E_list[0] + E_list[1] + E_list[2]) - GOLD_CISD_H2_2) < 1e-10


@pytest.fixture(scope="module")
def h2o_wfn(dev):
    p = params(H2O, 'CISD_SO')
    H = Hamiltonian(p)
    wfn = hf_wfn(H)
    wfn.solve_SCF(p)
    return p, H, wfn


def test_hf_matches_oracle_scf(h2o_wfn):
    p, H, wfn = h2o_wfn
    E, C, eps, _ = scf.solve_scf(H.T, H.V, H.ERI, H.S, H.E_nuc, wfn.ndocc, p)
    assert abs(wfn.E_SCF - E) < 1e-11
    assert np.abs(wfn.eps - eps).max() < 1e-9


def test_utils_transforms_match_oracle(h2o_wfn):
    p, H, wfn = h2o_wfn
    C_list, I_list = utils.get_slices(p, wfn)
    F_mo, E_fc = utils.compute_F_MO(p, wfn, C_list)
    F_ref, _ = transforms.compute_F_MO(H.T + H.V, H.ERI, wfn.C, C_list, False)
    assert np.abs(F_mo - F_ref).max() < 1e-12
    eri = utils.compute_ERI_MO(p, wfn, C_list)
    eri_ref = transforms.compute_ERI_MO(H.ERI, wfn.C, C_list)
    assert np.abs(eri - eri_ref).max() < 1e-12
    assert np.array_equal(utils.compute_F_SO(wfn, F_ref), transforms.compute_F_SO_loop(F_ref))
    assert np.array_equal(utils.compute_ERI_SO(wfn, eri_ref[:3, :3, :3, :3]),
                          transforms.compute_ERI_SO_loop(eri_ref[:3, :3, :3, :3]))


def test_utils_frozen_core_fock(h2o_wfn):
    p, H, wfn = h2o_wfn
    pf = dict(p, freeze_core=True)
    C_list, _ = utils.get_slices(pf, wfn)
    assert C_list[0] == slice(0, 1)
    F_mo, E_fc = utils.compute_F_MO(pf, wfn, C_list)
    F_ref, E_ref = transforms.compute_F_MO(H.T + H.V, H.ERI, wfn.C, C_list, True)
    assert abs(E_fc - E_ref) < 1e-11
    assert np.abs(F_mo - F_ref).max() < 1e-12


def test_mo_overlap_and_phase_match_oracle(h2o_wfn):
    p, H, wfn = h2o_wfn
    rng = np.random.default_rng(5)
    ph = np.exp(1j * rng.uniform(0, 2 * np.pi, wfn.nbf))
    C2 = wfn.C * ph[None, :]
    S = utils.compute_mo_overlap(wfn.ndocc, wfn.nbf, H.basis_set, wfn.C, H.basis_set, C2)
    assert np.abs(S - transforms.compute_mo_overlap(H.S, wfn.C, C2)).max() < 1e-12
    Cn = utils.compute_phase(wfn.ndocc, wfn.nbf, H.basis_set, wfn.C, H.basis_set, C2)
    assert np.abs(Cn - transforms.compute_phase(H.S, wfn.C, C2)).max() < 1e-12
    assert np.abs(Cn - wfn.C).max() < 1e-10       # the phases are undone
    Sso = utils.compute_so_overlap(wfn.nbf, S)
    assert np.array_equal(Sso, transforms.compute_so_overlap_loop(S))


def test_solve_general_diis_matches_oracle(dev):
    rng = np.random.default_rng(3)
    L = 50
    e_iter = rng.standard_normal((L, 4)) + 1j * rng.standard_normal((L, 4))
    t_iter = rng.standard_normal((L, 4)) + 1j * rng.standard_normal((L, 4))
    r = rng.standard_normal(L) + 1j * rng.standard_normal(L)
    t = rng.standard_normal(L) + 1j * rng.standard_normal(L)
    got = utils.solve_general_DIIS({}, r, t, e_iter, t_iter, 3)
    ref = so.solve_general_diis(r, t, e_iter, t_iter, 3)
    assert np.abs(got[0] - ref[0]).max() < 1e-10
    assert np.array_equal(got[1], ref[1]) and np.array_equal(got[2], ref[2])


def test_spatial_cisd_amplitudes_match_spatial_oracle(h2o_wfn):
    p, H, wfn = h2o_wfn
    E, t1, t2 = ci_wfn(dict(p, method='CISD'), wfn).solve_CISD()
    C_list = [slice(0, 0), slice(0, 5), slice(5, 7), slice(0, 7)]
    eri = transforms.compute_ERI_MO(H.ERI, wfn.C, C_list).swapaxes(1, 2)
    F, _ = transforms.compute_F_MO(H.T + H.V, H.ERI, wfn.C, C_list, False)
    Er, r1, r2, _ = spatial.solve_cisd(F, eri, wfn.eps, 5, p)
    assert abs(E - Er) < 1e-10
    assert np.abs(t1 - r1).max() < 1e-9 and np.abs(t2 - r2).max() < 1e-9


def test_spatial_mp2_amplitudes_match_oracle(h2o_wfn):
    p, H, wfn = h2o_wfn
    E, t2 = mp2_wfn(dict(p, method='MP2'), wfn).solve_MP2()
    C_list = [slice(0, 0), slice(0, 5), slice(5, 7), slice(0, 7)]
    eri = transforms.compute_ERI_MO(H.ERI, wfn.C, C_list).swapaxes(1, 2)
    Er, r2 = spatial.solve_mp2(eri, wfn.eps, 5)
    assert abs(E - Er) < 1e-12 and np.abs(t2 - r2).max() < 1e-12


def test_batched_mp2_uneven_instances_match_oracle(dev):
    from apyib_amd import synthetic
    from apyib_amd.so_ci import MP2_BLOCKS, SOIntegrals
    sizes = [(4, 6), (6, 9), (2, 3), (5, 12), (8, 7)]
    inst, refs = [], []
    for k, (O, V) in enumerate(sizes):
        F, A, eps = synthetic.system(O, V, 100 + k)
        ints = SOIntegrals.synthetic(O, V, 100 + k, device=dev, blocks=MP2_BLOCKS)
        inst.append((O, V, torch.from_numpy(eps).to(dev), ints['vvoo_t'], ints['oovv']))
        refs.append(so.solve_mp2_so(A / 2, eps[:O], eps[O:], O))
    E, t2s = tr.mp2_so_batched(inst, dev)
    E = E.cpu().numpy()
    for k in range(len(sizes)):
        assert abs(E[k] - refs[k][0]) < 1e-14
        assert np.abs(t2s[k].cpu().numpy() - refs[k][1]).max() < 1e-14
