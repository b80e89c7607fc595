"""H2O at O-H 0.96 A and H-O-H 104.5 deg, STO-3G, zero field (14 spin
orbitals: 10 occupied, 4 virtual): RHF, MP2 and CISD (and their CID /
spin-orbital forms) through the reference's API on the GPU path
(energy(), energy.py:17-81), against the oracle chain on the same integrals
at 1e-10 Eh, with amplitudes at 1e-9.

The reference holds no golden value at this geometry (its goldens are at
the geometry of data/molecules.py:2-10); the integral engine and the chain
are pinned there (tests/test_oracle_golden.py, tests/test_gpu_dropin.py)."""
import numpy as np
import pytest

import apyib_amd.energy as energy_mod
from apyib_amd.hamiltonian import Hamiltonian
from oracle import scf, so, spatial, transforms
from tests.helpers import PARAMS

pytestmark = pytest.mark.gpu

H2O_C0 = """
O  0.000000000000  0.000000000000  0.000000000000
H  0.000000000000  0.759061990794  0.587728588833
H  0.000000000000 -0.759061990794  0.587728588833
symmetry c1
no_com
no_reorient
units angstrom
"""


def params(method):
    return {'geom': H2O_C0, 'basis': 'STO-3G', 'method': method, 'e_convergence': 1e-12, 'd_convergence': 1e-12,
            'DIIS': True, 'freeze_core': False, 'F_el': [0.0, 0.0, 0.0], 'F_mag': [0.0, 0.0, 0.0],
            'max_iterations': 120}


@pytest.fixture(scope="module")
def chain():
    """Oracle chain: host integrals -> RHF -> AO->MO(->SO) transforms."""
    p = params('RHF')
    H = Hamiltonian(p)
    E, C, eps, _ = scf.solve_scf(H.T, H.V, H.ERI, H.S, H.E_nuc, 5, PARAMS)
    return H, E, C, eps


def test_c0_sizes_and_geometry(chain):
    H, *_ = chain
    assert H.basis_set.nbf() == 7
    R = H.molecule.coords
    bohr = np.linalg.norm(R[1] - R[0])
    ang = np.degrees(np.arccos((R[1] - R[0]) @ (R[2] - R[0]) / (np.linalg.norm(R[1] - R[0])
                                                               * np.linalg.norm(R[2] - R[0]))))
    assert abs(np.linalg.norm(R[2] - R[0]) - bohr) < 1e-12
    assert abs(ang - 104.5) < 1e-9


@pytest.mark.parametrize("method", ["RHF", "MP2", "MP2_SO", "CID", "CID_SO", "CISD", "CISD_SO"])
def test_c0_energy_matches_oracle_chain(dev, chain, method):
    H, E_scf, C, eps = chain
    E_list, T_list, C_gpu, _ = energy_mod.energy(params(method))
    assert abs(E_list[0] - E_scf) < 1e-10
    assert abs(E_list[2] - H.E_nuc) < 1e-12
    if method == 'RHF':
        return
    nd, nbf = 5, 7
    C_list, _ = transforms.get_slices(nbf, nd, 0, method)
    # the GPU and host SCF orbitals may differ by a phase per orbital: the
    # correlated energy is invariant, the amplitudes are compared through
    # the oracle run on the GPU's own orbitals
    eri_mo = transforms.compute_ERI_MO(H.ERI, C_gpu, C_list)
    F_mo = transforms.compute_F_MO(H.T + H.V, H.ERI, C_gpu, C_list, False)[0]
    t1, t2 = T_list[1], T_list[2]
    if method.endswith('_SO'):
        eri = transforms.compute_ERI_SO(eri_mo).swapaxes(1, 2)
        F = transforms.compute_F_SO(F_mo)
        e = np.repeat(eps, 2)
        if method == 'MP2_SO':
            Er, r2 = so.solve_mp2_so(eri, e[:10], e[10:], 10)
            r1 = None
        elif method == 'CID_SO':
            Er, r2, _ = so.solve_cid_so(F, eri, e[:10], e[10:], 10, PARAMS)
            r1 = None
        else:
            Er, r1, r2, _ = so.solve_cisd_so(F, eri, e[:10], e[10:], 10, PARAMS)
    else:
        eri = eri_mo.swapaxes(1, 2)
        if method == 'MP2':
            Er, r2 = spatial.solve_mp2(eri, eps, nd)
            r1 = None
        elif method == 'CID':
            Er, r2, _ = spatial.solve_cid(F_mo, eri, eps, nd, PARAMS)
            r1 = None
        else:
            Er, r1, r2, _ = spatial.solve_cisd(F_mo, eri, eps, nd, PARAMS)
    assert abs(E_list[1] - Er) < 1e-10
    assert np.abs(np.asarray(t2) - r2).max() < 1e-9
    if r1 is not None:
        assert np.abs(np.asarray(t1) - r1).max() < 1e-9
