"""Batched displaced Hartree-Fock (hf_wfn.solve_SCF_batch: one
apy_fock_build_batch per iteration for every instance), the batched
phase-corrected energies behind the finite-difference driver (CISD/CID
through spatial_ci.SpatialCISet) and the set solver itself, against the
one-by-one solves they replace (hf_wfn.py:36-170, energy.py:84-150,
ci_wfn.py:420-574)."""
import copy

import numpy as np
import pytest

from apyib_amd.energy import energy, phase_corrected_energies, phase_corrected_energy
from apyib_amd.hamiltonian import Hamiltonian
from apyib_amd.hf_wfn import hf_wfn, solve_SCF_batch
from apyib_amd.integrals import prime_device_eri
from apyib_amd.molecule import Molecule
from apyib_amd.utils import device
from tests.molecules import H2O

pytestmark = pytest.mark.gpu

PARAMS = {'basis': 'STO-3G', 'method': 'MP2', 'e_convergence': 1e-12, 'd_convergence': 1e-12, 'DIIS': True,
          'freeze_core': False, 'F_el': [0.0, 0.0, 0.0], 'F_mag': [0.0, 0.0, 0.0], 'max_iterations': 120}


def _displaced(method='MP2'):
    p0 = dict(PARAMS, geom=H2O, method=method)
    mol = Molecule.from_string(H2O)
    out = []
    for alpha, sgn in ((0, 1), (4, -1), (8, 1)):
        g = mol.coords.copy()
        g[alpha // 3][alpha % 3] += sgn * 1e-3
        q = copy.deepcopy(p0)
        q['geom'] = mol.with_coords(g).geometry_string()
        out.append(q)
    for beta, sgn in ((0, 1), (2, -1)):
        q = copy.deepcopy(p0)
        q['F_mag'][beta] += sgn * 1e-4
        out.append(q)
    return p0, out


def test_batched_scf_equals_single_solves(dev):
    _, plist = _displaced()
    batch = [hf_wfn(Hamiltonian(p)) for p in plist]
    prime_device_eri([w.H for w in batch], device())
    res = solve_SCF_batch(batch, plist[0])
    for p, (E, C), w in zip(plist, res, batch):
        single = hf_wfn(Hamiltonian(p))
        E1, C1 = single.solve_SCF(p)
        assert E == E1
        assert np.array_equal(C, C1)
        assert np.array_equal(w.eps, single.eps)


@pytest.mark.parametrize("method", ["MP2", "CISD", "CID"])
def test_phase_corrected_energies_equal_one_by_one(dev, method):
    p0, plist = _displaced(method)
    _, _, C0, basis0 = energy(p0)
    got = phase_corrected_energies(plist, basis0, C0)
    for p, (E_list, T_list, C, basis) in zip(plist, got):
        E1, T1, Cs, b1 = phase_corrected_energy(p, basis0, C0)
        assert np.abs(np.asarray(E_list, dtype=complex) - np.asarray(E1, dtype=complex)).max() < 1e-13
        assert np.abs(C - Cs).max() < 1e-12
        assert np.abs(np.asarray(T_list[2]) - np.asarray(T1[2])).max() < 1e-12


@pytest.mark.parametrize("method", ["CISD_SO", "CID_SO"])
def test_spin_orbital_set_equals_one_by_one(dev, method):
    """A spin-orbital finite-difference set solved as one spin-adapted set and
    expanded to spin orbitals against the per-instance spin-orbital solves
    (ci_wfn.py:171-416 on the SO blocks): energies and amplitudes equal to
    the convergence threshold (1e-12 on energy and amplitude change)."""
    p0, plist = _displaced(method)
    _, _, C0, basis0 = energy(p0)
    got = phase_corrected_energies(plist, basis0, C0)
    for p, (E_list, T_list, C, basis) in zip(plist, got):
        E1, T1, Cs, b1 = phase_corrected_energy(p, basis0, C0)
        assert abs(complex(E_list[1]) - complex(E1[1])) < 1e-10
        t2, t2s = np.asarray(T_list[2]), np.asarray(T1[2])
        assert t2.shape == t2s.shape and np.abs(t2 - t2s).max() < 1e-9
        assert np.abs(t2 + t2.transpose(1, 0, 2, 3)).max() == 0     # antisymmetric, as the SO solver's
        if method == "CISD_SO":
            assert np.abs(np.asarray(T_list[1]) - np.asarray(T1[1])).max() < 1e-9
        else:
            assert T_list[1] == 0 and T1[1] == 0


def _spatial_system(o, v, seed, std=0.05, noise=1e-3):
    rng = np.random.default_rng(seed)
    n = o + v
    E = std * (rng.standard_normal((n,) * 4) + 1j * rng.standard_normal((n,) * 4))
    E = 0.5 * (E + E.transpose(2, 3, 0, 1))
    eps = np.concatenate([np.sort(rng.uniform(-2.0, -0.5, o)), np.sort(rng.uniform(0.3, 2.0, v))])
    F = np.diag(eps).astype(complex) + noise * (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    return E, F, eps


@pytest.mark.parametrize("singles", [True, False], ids=["CISD", "CID"])
@pytest.mark.parametrize("o,v,nb", [(3, 5, 4), (4, 9, 7), (1, 4, 3)])
def test_spatial_set_equals_single_solves(dev, singles, o, v, nb):
    """nb systems with their own integrals, Fock matrices and orbital
    energies iterated together (apy_cisd_sp_step_multi): every system's
    energy, amplitudes and iteration count equal its own SpatialCI solve
    (host-controlled, the same DIIS arithmetic); systems converge at
    different iterations."""
    import torch
    from apyib_amd.spatial_ci import SPIntegrals, SpatialCI, SpatialCISet
    from tests.helpers import PARAMS as P
    systems = [_spatial_system(o, v, 40 + q, std=0.03 + 0.01 * q) for q in range(nb)]
    d = lambda x: torch.from_numpy(np.ascontiguousarray(x)).to(dev)
    cset = SpatialCISet([d(E4) for E4, _, _ in systems], [d(F) for _, F, _ in systems],
                        [eps for _, _, eps in systems], o, singles=singles)
    E, t1, t2, iters = cset.solve(P)
    if o > 1:
        assert len(set(iters)) > 1         # systems of different coupling stop at different iterations
    for q, (E4, F, eps) in enumerate(systems):
        ci = SpatialCI(SPIntegrals.from_mo(d(E4), o, singles=singles), d(F), eps, singles=singles)
        Eq, t1q, t2q, itq = ci.solve(P, control="host")
        assert iters[q] == itq
        assert abs(E[q] - Eq) < 1e-13 * max(1.0, abs(Eq))
        assert (t2[q] - t2q).abs().max().item() < 1e-12
        if singles:
            assert (t1[q] - t1q).abs().max().item() < 1e-12
