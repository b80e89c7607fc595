"""Uniform electric field on the GPU drop-in (energy(), hamiltonian.py:41-63,
with the dipole integrals from the device integral kernels): the RHF field
derivative equals minus the dipole moment of the zero-field density
(Hellmann-Feynman; the dipole itself is pinned by the published H2O/STO-3G
value in tests/test_oracle_field.py), E_nuc carries exactly the nuclear
dipole term, and correlated energies in a field match the oracle chain."""
import numpy as np
import pytest

import apyib_amd.energy as energy_mod
from apyib_amd import gaussian
from apyib_amd.hamiltonian import Hamiltonian
from oracle import scf, so, spatial, transforms
from tests.helpers import PARAMS
from tests.molecules import H2O
from tests.test_oracle_field import CRAWFORD_MU_Y

pytestmark = pytest.mark.gpu


def params(method, F_el=(0.0, 0.0, 0.0)):
    return {'geom': H2O, 'basis': 'STO-3G', 'method': method, 'e_convergence': 1e-13, 'd_convergence': 1e-13,
            'DIIS': True, 'freeze_core': False, 'F_el': list(F_el), 'F_mag': [0.0, 0.0, 0.0],
            'max_iterations': 200}


def total(E_list):
    return (E_list[0] + E_list[1] + E_list[2]).real


def test_rhf_field_derivative_is_minus_the_dipole(dev):
    E_list, _, C, basis, H = energy_mod.energy(params('RHF'), return_H=True)
    nd = 5
    D = 2.0 * C[:, :nd] @ C[:, :nd].conj().T
    r = gaussian.build_dipole(basis.functions)
    nuc = H.molecule.Z @ H.molecule.coords
    mu = np.array([nuc[a] - np.einsum('mn,nm->', D, r[a]).real for a in range(3)])
    assert abs(mu[1] - CRAWFORD_MU_Y) < 1e-9
    h = 1e-4
    for axis in range(3):
        F = np.zeros(3)
        F[axis] = h
        Ep, _, _, _, Hp = energy_mod.energy(params('RHF', F), return_H=True)
        Em, _, _, _, Hm = energy_mod.energy(params('RHF', -F), return_H=True)
        assert abs(Hp.E_nuc - H.E_nuc + h * nuc[axis]) < 1e-13
        assert abs(Hm.E_nuc - H.E_nuc - h * nuc[axis]) < 1e-13
        fd = (total(Ep) - total(Em)) / (2 * h)
        assert abs(fd + mu[axis]) < 1e-6, axis


@pytest.mark.parametrize("method", ["RHF", "MP2", "MP2_SO", "CID", "CISD", "CISD_SO"])
def test_energy_in_field_matches_oracle_chain(dev, method):
    F_el = (0.011, -0.017, 0.005)
    p = params(method, F_el)
    E_list, _, _, _ = energy_mod.energy(p)
    H = Hamiltonian(p)                                   # host integrals, same field terms
    nd, nbf = 5, 7
    E, C, eps, _ = scf.solve_scf(H.T, H.V, H.ERI, H.S, H.E_nuc, nd, PARAMS)
    want = E + H.E_nuc
    if method != 'RHF':
        C_list, _ = transforms.get_slices(nbf, nd, 0, method)
        eri_mo = transforms.compute_ERI_MO(H.ERI, C, C_list)
        F_mo = transforms.compute_F_MO(H.T + H.V, H.ERI, C, C_list, False)[0]
        if method.endswith('_SO'):
            eri = transforms.compute_ERI_SO(eri_mo).swapaxes(1, 2)
            F = transforms.compute_F_SO(F_mo)
            e = np.repeat(eps, 2)
            if method == 'MP2_SO':
                want += so.solve_mp2_so(eri, e[:10], e[10:], 10)[0]
            else:
                want += so.solve_cisd_so(F, eri, e[:10], e[10:], 10, PARAMS)[0]
        else:
            eri = eri_mo.swapaxes(1, 2)
            if method == 'MP2':
                want += spatial.solve_mp2(eri, eps, nd)[0]
            elif method == 'CID':
                want += spatial.solve_cid(F_mo, eri, eps, nd, PARAMS)[0]
            else:
                want += spatial.solve_cisd(F_mo, eri, eps, nd, PARAMS)[0]
    assert abs(total(E_list) - want.real) < 1e-10
