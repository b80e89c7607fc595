"""AO electron-repulsion integrals on the GPU (apy_eri_gto,
apy_eri_deriv_assemble) against the host McMurchie-Davidson engine
(gaussian.build_eri, gaussian.build_deriv1), whose integrals are pinned by
the reference's energy and AAT goldens and by central differences."""
import ctypes

import numpy as np
import pytest
import torch

from apyib_amd import _lib, gaussian
from apyib_amd.integrals import EriDerivatives, ao_eri, basis_table, eri
from apyib_amd.molecule import BasisSet, Molecule
from tests.molecules import H2O

pytestmark = pytest.mark.gpu

H4 = """
H 0.0 0.0 0.0
H 0.0 0.0 1.4
H 1.5 0.3 -0.2
H 1.6 0.5 1.3
units bohr
"""


@pytest.mark.parametrize("geom,basis", [(H2O, "STO-3G"), (H4, "STO-6G"), (H4, "STO-3G")])
def test_ao_eri_matches_host(dev, geom, basis):
    mol = Molecule.from_string(geom)
    bs = BasisSet(mol, basis)
    want = gaussian.build_eri(bs.functions)
    got = ao_eri(bs, mol.coords, dev).cpu().numpy()
    assert np.abs(got.imag).max() == 0.0
    assert np.abs(got.real - want).max() < 1e-12


def test_ao_eri_batched_geometries(dev):
    """Three displaced H2O geometries in one launch, each against the host
    integrals of its own basis (the functions follow their atoms)."""
    mol = Molecule.from_string(H2O)
    bs = BasisSet(mol, "STO-3G")
    rng = np.random.default_rng(4)
    geoms = np.stack([mol.coords + 0.05 * rng.standard_normal(mol.coords.shape) for _ in range(3)])
    got = ao_eri(bs, geoms, dev).cpu().numpy().real
    for g in range(3):
        want = gaussian.build_eri(BasisSet(mol.with_coords(geoms[g]), "STO-3G").functions)
        assert np.abs(got[g] - want).max() < 1e-12


def test_eri_derivatives_match_host(dev):
    mol = Molecule.from_string(H2O)
    bs = BasisSet(mol, "STO-3G")
    d = gaussian.build_deriv1(bs.functions, bs.atom_of, list(zip(mol.Z, mol.coords)))
    der = EriDerivatives(bs, mol.coords, dev)
    for A in range(mol.natom()):
        for x in range(3):
            got = der(A, x).cpu().numpy()
            assert np.abs(got.real - d['ERI'][A][x]).max() < 1e-12, (A, x)
            assert np.abs(got.imag).max() == 0.0


def test_entry_outside_limits_gives_nan(dev):
    """A table entry with an f-type term (l = 3) is rejected inside the
    kernel before any indexing: that element comes back NaN, the others
    are unaffected."""
    mol = Molecule.from_string(H2O)
    bs = BasisSet(mol, "STO-3G")
    good = basis_table(bs.functions, bs.atom_of, dev)
    g = _lib.Gto.from_buffer_copy(bytes(good[:ctypes.sizeof(_lib.Gto)].cpu().numpy()))
    g.lmn[0][0] = 3
    bad = torch.from_numpy(np.frombuffer(bytes(g), dtype=np.uint8).copy()).to(dev)
    out = eri(bad, good, good, good, (1, 1, 1, 1), mol.coords[None], dev).cpu().numpy()
    assert np.isnan(out.real).all()
    ok = eri(good, good, good, good, (1, 1, 1, 1), mol.coords[None], dev).cpu().numpy()
    assert np.isfinite(ok.real).all()
