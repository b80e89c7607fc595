"""AO integrals on the GPU (apy_eri_gto, apy_eri_deriv_assemble,
apy_oei_gto) against the host McMurchie-Davidson engine (gaussian.build_eri,
build_one_electron, build_dipole, build_angular_momentum,
build_mixed_overlap, build_deriv1), whose integrals are pinned by the
reference's energy and AAT goldens and by central differences."""
import ctypes

import numpy as np
import pytest
import torch

from apyib_amd import _lib, gaussian, integrals
from apyib_amd.integrals import (EriDerivatives, ao_eri, basis_table, derivative_table, eri, hamiltonian_integrals,
                                 mixed_overlaps, one_electron)
from apyib_amd.molecule import BasisSet, Molecule
from tests.molecules import H2O, H2O2

pytestmark = pytest.mark.gpu

H4 = """
H 0.0 0.0 0.0
H 0.0 0.0 1.4
H 1.5 0.3 -0.2
H 1.6 0.5 1.3
units bohr
"""


@pytest.mark.parametrize("geom,basis", [(H2O, "STO-3G"), (H4, "STO-6G"), (H4, "STO-3G")],
                         ids=["H2O-STO-3G", "H4-STO-6G", "H4-STO-3G"])
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
    """A table entry with a g-type term (l = 4, beyond APY_GTO_MAXL) is
    rejected inside the kernel before any indexing: that element comes back
    NaN, the others are unaffected."""
    mol = Molecule.from_string(H2O)
    bs = BasisSet(mol, "STO-3G")
    good = basis_table(bs.functions, bs.atom_of, dev)
    g = _lib.Gto.from_buffer_copy(bytes(good[:ctypes.sizeof(_lib.Gto)].cpu().numpy()))
    g.lmn[0][0] = _lib.APY_GTO_MAXL + 1
    bad = torch.from_numpy(np.frombuffer(bytes(g), dtype=np.uint8).copy()).to(dev)
    out = eri(bad, good, good, good, (1, 1, 1, 1), mol.coords[None], dev).cpu().numpy()
    assert np.isnan(out.real).all()
    ok = eri(good, good, good, good, (1, 1, 1, 1), mol.coords[None], dev).cpu().numpy()
    assert np.isfinite(ok.real).all()


# ------------------------------------------------------------------ one-electron integrals

def _host_one_electron(bs, mol):
    fns = bs.functions
    S, T, V = gaussian.build_one_electron(fns, list(zip(mol.Z, mol.coords)))
    return S, T, V, gaussian.build_dipole(fns), gaussian.build_angular_momentum(fns)


@pytest.mark.parametrize("geom,basis", [(H2O, "STO-3G"), (H4, "STO-6G"), (H2O2, "STO-3G")],
                         ids=["H2O-STO-3G", "H4-STO-6G", "H2O2-STO-3G"])
def test_one_electron_matches_host(dev, geom, basis):
    """S, T, V, dipole and angular momentum from one apy_oei_gto launch
    against the host engine (which the reference's RHF energy and AAT
    goldens pin)."""
    mol = Molecule.from_string(geom)
    bs = BasisSet(mol, basis)
    got = hamiltonian_integrals(bs, mol.coords[None], mol.Z, dev).cpu().numpy()[0]
    S, T, V, D, L = _host_one_electron(bs, mol)
    for k, want in enumerate([S, T, V, D[0], D[1], D[2], L[0], L[1], L[2]]):
        scale = max(1.0, np.abs(want).max())
        assert np.abs(got[k] - want).max() < 1e-12 * scale, k
    assert np.abs(got[6:] + got[6:].swapaxes(1, 2)).max() < 1e-12   # L antisymmetric


def test_one_electron_component_selection_and_batch(dev):
    """Selected components come out in the fixed order; geometries of one
    launch equal separate launches bit for bit."""
    mol = Molecule.from_string(H2O)
    bs = BasisSet(mol, "STO-3G")
    rng = np.random.default_rng(7)
    geoms = np.stack([mol.coords + 0.05 * rng.standard_normal(mol.coords.shape) for _ in range(4)])
    full = hamiltonian_integrals(bs, geoms, mol.Z, dev).cpu().numpy()
    sel = hamiltonian_integrals(bs, geoms, mol.Z, dev,
                                which=_lib.APY_OEI_KINETIC | _lib.APY_OEI_ANGMOM).cpu().numpy()
    assert sel.shape == (4, 4) + full.shape[2:]
    assert np.array_equal(sel[:, 0], full[:, 1]) and np.array_equal(sel[:, 1:], full[:, 6:])
    for g in range(4):
        one = hamiltonian_integrals(bs, geoms[g:g + 1], mol.Z, dev).cpu().numpy()[0]
        assert np.array_equal(one, full[g])
        S, T, V, D, L = _host_one_electron(BasisSet(mol.with_coords(geoms[g]), "STO-3G"), mol.with_coords(geoms[g]))
        assert np.abs(full[g, 2] - V).max() < 1e-12 * np.abs(V).max()


def test_mixed_overlaps_match_host(dev):
    """<bra|ket> between displaced geometries (utils.py:371-376 for every
    pair of an AAT overlap set) against the host overlaps."""
    mol = Molecule.from_string(H2O2)
    bs0 = BasisSet(mol, "STO-3G")
    rng = np.random.default_rng(11)
    others = [BasisSet(mol.with_coords(mol.coords + 1e-2 * rng.standard_normal(mol.coords.shape)), "STO-3G")
              for _ in range(3)]
    pairs = [(bs0.functions, b.functions) for b in others] + [(others[0].functions, others[1].functions),
                                                              (bs0.functions, bs0.functions)]
    got = mixed_overlaps(pairs, dev)
    for (bra, ket), S in zip(pairs, got):
        want = gaussian.build_mixed_overlap(bra, ket)
        assert np.abs(S.cpu().numpy() - want).max() < 1e-13


def test_one_electron_derivative_functions(dev):
    """Two-term (differentiated) functions: dS/dR and dT/dR assembled from
    (d m/dx | n) against the host derivative integrals."""
    mol = Molecule.from_string(H2O)
    bs = BasisSet(mol, "STO-3G")
    d = gaussian.build_deriv1(bs.functions, bs.atom_of, list(zip(mol.Z, mol.coords)), eri_derivatives=False)
    tab = basis_table(bs.functions, bs.atom_of, dev)
    c = mol.coords[None]
    on = np.asarray(bs.atom_of)
    for x in range(3):
        dtab = derivative_table(bs.functions, bs.atom_of, x, dev)
        DS, DT = one_electron(dtab, c, tab, c, _lib.APY_OEI_OVERLAP | _lib.APY_OEI_KINETIC,
                              dev).cpu().numpy()[0]
        for A in range(mol.natom()):
            m = (on == A).astype(float)
            for got, want in ((DS, d['S'][A][x]), (DT, d['T'][A][x])):
                asm = m[:, None] * got + m[None, :] * got.T
                assert np.abs(asm - want).max() < 1e-12, (A, x)


def test_one_electron_entry_outside_limits_gives_nan(dev):
    mol = Molecule.from_string(H2O)
    bs = BasisSet(mol, "STO-3G")
    good = basis_table(bs.functions, bs.atom_of, dev)
    g = _lib.Gto.from_buffer_copy(bytes(good[:ctypes.sizeof(_lib.Gto)].cpu().numpy()))
    g.atom = 7                                       # no such centre
    bad = torch.from_numpy(np.frombuffer(bytes(g), dtype=np.uint8).copy()).to(dev)
    c = mol.coords[None]
    out = one_electron(bad, c, good, c, integrals.OEI_ALL, dev, nuclei=c, charges=mol.Z).cpu().numpy()
    assert out.shape == (1, 9, 1, len(bs.functions)) and np.isnan(out).all()


@pytest.mark.parametrize("geom", [H2O, H2O2], ids=["H2O", "H2O2"])
def test_deriv1_one_electron_matches_host(dev, geom):
    """d/dR of S, T, V (with the operator term of the moving nucleus) and
    the half-derivative overlaps from the device against the host engine's
    build_deriv1 (which central differences of S, T, V pin)."""
    mol = Molecule.from_string(geom)
    bs = BasisSet(mol, "STO-3G")
    want = gaussian.build_deriv1(bs.functions, bs.atom_of, list(zip(mol.Z, mol.coords)), eri_derivatives=False)
    got = integrals.deriv1_one_electron(bs, mol.coords, mol.Z, dev)
    for key in ('S', 'T', 'V', 'half_S'):
        for A in range(mol.natom()):
            for x in range(3):
                w = want[key][A][x]
                assert np.abs(got[key][A][x] - w).max() < 1e-12 * max(1.0, np.abs(w).max()), (key, A, x)


def test_d_functions_match_host(dev):
    """6-31G* on H2O (Cartesian d on O): every one-electron integral, and
    300 seeded ERI elements plus 99 derivative-function ERI elements (f
    terms) against the host engine's single-quartet routine."""
    mol = Molecule.from_string(H2O)
    bs = BasisSet(mol, "6-31G*")
    fns = bs.functions
    n = len(fns)
    got = hamiltonian_integrals(bs, mol.coords[None], mol.Z, dev).cpu().numpy()[0]
    S, T, V, D, L = _host_one_electron(bs, mol)
    for k, want in enumerate([S, T, V, D[0], D[1], D[2], L[0], L[1], L[2]]):
        assert np.abs(got[k] - want).max() < 1e-12 * max(1.0, np.abs(want).max()), k
    G = ao_eri(bs, mol.coords, dev).cpu().numpy().real
    rng = np.random.default_rng(3)
    dd = [i for i, f in enumerate(fns) if sum(f.lmn) == 2]
    picks = [tuple(rng.integers(0, n, 4)) for _ in range(200)] + [tuple(rng.choice(dd, 4)) for _ in range(100)]
    for (i, j, k, l) in picks:
        want = gaussian.eri(fns[i], fns[j], fns[k], fns[l])
        assert abs(G[i, j, k, l] - want) < 1e-12 * max(1.0, abs(want)), (i, j, k, l)
    der = EriDerivatives(bs, mol.coords, dev)
    for x in range(3):
        Dx = der.D[x].cpu().numpy().real
        for _ in range(33):
            i = int(rng.choice(dd))
            j, k, l = (int(q) for q in rng.integers(0, n, 3))
            want = sum(gaussian.eri(t, fns[j], fns[k], fns[l]) for t in gaussian.center_derivative(fns[i], x))
            assert abs(Dx[i, j, k, l] - want) < 1e-12 * max(1.0, abs(want)), (x, i, j, k, l)


@pytest.mark.parametrize("geom,basis", [(H2O, "STO-3G"), (H2O, "6-31G*"), (H2O2, "STO-3G")],
                         ids=["H2O-STO-3G", "H2O-6-31G*", "H2O2-STO-3G"])
def test_eri_symmetric_launches_match_general_launch(dev, geom, basis):
    """apy_eri_gto computes one quartet per permutational image class when
    the four tables are one table (8 images) or the two ket tables are one
    table (2 images); the same tables passed as distinct copies take the
    general path, one quartet per element.  The three launches agree to
    rounding (each image keeps the summation order of the quartet that
    was computed), and the 8-fold result is exactly symmetric."""
    mol = Molecule.from_string(geom)
    bs = BasisSet(mol, basis)
    n = len(bs.functions)
    tab = basis_table(bs.functions, bs.atom_of, dev)
    copies = [tab.clone() for _ in range(3)]
    cnt = (n, n, n, n)
    G8 = eri(tab, tab, tab, tab, cnt, mol.coords[None], dev).cpu().numpy()[0].real
    G2 = eri(copies[0], copies[1], tab, tab, cnt, mol.coords[None], dev).cpu().numpy()[0].real
    G1 = eri(copies[0], copies[1], copies[2], tab, cnt, mol.coords[None], dev).cpu().numpy()[0].real
    scale = np.abs(G1).max()
    assert np.abs(G8 - G1).max() <= 1e-14 * scale
    assert np.abs(G2 - G1).max() <= 1e-14 * scale
    for perm in ((1, 0, 2, 3), (0, 1, 3, 2), (2, 3, 0, 1), (3, 2, 1, 0)):
        assert np.array_equal(G8, G8.transpose(perm))
    assert np.array_equal(G2, G2.transpose(0, 1, 3, 2))
