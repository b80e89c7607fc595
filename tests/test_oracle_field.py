"""Uniform electric field terms of the Hamiltonian (hamiltonian.py:41-63):
V -= F_el . mu_el with mu_el = -<m|r|n> and E_nuc(-F_el).  CPU only
(host integral engine + the oracle RHF).

Pins, independent of how the terms are written:
  * Hellmann-Feynman: RHF is variational, so the central difference of the
    total energy E(+-h F_el) equals dE/dF = Tr(D r) - sum_A Z_A R_A = -mu
    evaluated with the zero-field density D;
  * E_nuc(F_el) - E_nuc(0) = -F_el . sum_A Z_A R_A exactly (the nuclear
    dipole in the field);
  * the RHF dipole moment of H2O/STO-3G at the reference's geometry
    (data/molecules.py:2-10) equals the value published with Crawford's
    programming project #3 for the same geometry and basis, 0.603521296525
    au along +y, which pins the dipole integrals and their origin."""
import numpy as np
import pytest

from apyib_amd import gaussian
from apyib_amd.hamiltonian import Hamiltonian
from oracle import scf
from tests.helpers import PARAMS
from tests.molecules import H2O

CRAWFORD_MU_Y = 0.603521296525


def params(F_el=(0.0, 0.0, 0.0)):
    return {'geom': H2O, 'basis': 'STO-3G', 'method': 'RHF', 'e_convergence': 1e-13, 'd_convergence': 1e-13,
            'DIIS': True, 'freeze_core': False, 'F_el': list(F_el), 'F_mag': [0.0, 0.0, 0.0], 'max_iterations': 200}


def rhf_total(p, ndocc=5):
    H = Hamiltonian(p)
    E, C, eps, _ = scf.solve_scf(H.T, H.V, H.ERI, H.S, H.E_nuc, ndocc, PARAMS)
    return E + H.E_nuc, C, H


@pytest.fixture(scope="module")
def zero_field():
    return rhf_total(params())


def _dipole(C, H, ndocc=5):
    D = 2.0 * C[:, :ndocc] @ C[:, :ndocc].conj().T
    r = gaussian.build_dipole(H.basis_set.functions)
    nuc = H.molecule.Z @ H.molecule.coords
    return np.array([nuc[a] - np.einsum('mn,nm->', D, r[a]).real for a in range(3)])


def test_nuclear_energy_in_field_is_the_nuclear_dipole_term(zero_field):
    _, _, H0 = zero_field
    F = np.array([0.013, -0.021, 0.007])
    H = Hamiltonian(params(F))
    nuc = H0.molecule.Z @ H0.molecule.coords
    assert abs((H.E_nuc - H0.E_nuc) - (-F @ nuc)) < 1e-14


def test_field_adds_the_dipole_operator_to_V(zero_field):
    _, _, H0 = zero_field
    F = np.array([0.013, -0.021, 0.007])
    H = Hamiltonian(params(F))
    r = gaussian.build_dipole(H0.basis_set.functions)
    assert np.abs(H.V - (H0.V + np.einsum('a,amn->mn', F, r))).max() < 1e-15
    assert np.array_equal(H.T, H0.T) and np.array_equal(H.S, H0.S)


def test_rhf_dipole_matches_published_value(zero_field):
    _, C, H = zero_field
    mu = _dipole(C, H)
    assert abs(mu[0]) < 1e-10 and abs(mu[2]) < 1e-10
    assert abs(mu[1] - CRAWFORD_MU_Y) < 1e-9


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_hellmann_feynman_field_derivative(zero_field, axis):
    _, C, H = zero_field
    h = 1e-4
    Fp, Fm = np.zeros(3), np.zeros(3)
    Fp[axis], Fm[axis] = h, -h
    Ep, _, _ = rhf_total(params(Fp))
    Em, _, _ = rhf_total(params(Fm))
    fd = (Ep - Em).real / (2 * h)
    assert abs(fd - (-_dipole(C, H)[axis])) < 1e-6
