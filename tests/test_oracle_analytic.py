"""Derivative integrals and the Hartree-Fock analytic AAT oracle (CPU).

Goldens: tests/golden/analytic_aat_goldens.json (written by
make_analytic_goldens.py from apyib/tests/test_018..021_*_Analytic_AAT.py).
The H2O2/STO-3G entries run here; the H2O/6-31G* ones (d functions) would
take the host integral engine minutes and run on the GPU drop-in instead
(tests/test_gpu_analytic.py)."""
import copy
import json
import os

import numpy as np
import pytest

from apyib_amd import gaussian
from apyib_amd.molecule import BasisSet, Molecule
from oracle import analytic, spatial
from tests.molecules import H2O

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "analytic_aat_goldens.json")))


def golden_id(e):
    return f"{e['file'][:8]}:{e['line']}:{e['test']}"


def runnable(e):
    """Goldens the CPU oracle runs (the host integral engine at STO-3G)."""
    return e["parameters"]["basis"] == "STO-3G"


def gpu_runnable(e):
    """Goldens the GPU drop-in runs: every basis in the package's table."""
    return e["parameters"]["basis"] in ("STO-3G", "6-31G*")


def golden_params(e):
    p = copy.deepcopy(e["parameters"])
    p["geom"] = p["geom"]["geometry"]
    return p


def test_derivative_integrals_match_central_differences():
    """S, T, V (with the operator term), ERI and the bra half-derivative
    overlap against central differences of the integrals at displaced
    geometries (the displaced basis moves with its atom)."""
    mol = Molecule.from_string(H2O)
    bs = BasisSet(mol, 'STO-3G')
    d = gaussian.build_deriv1(bs.functions, bs.atom_of, list(zip(mol.Z, mol.coords)))
    h = 1e-5
    for A, x in ((0, 0), (1, 2), (2, 1)):
        vals = []
        for sgn in (+1, -1):
            c = mol.coords.copy()
            c[A, x] += sgn * h
            m2 = mol.with_coords(c)
            b2 = BasisSet(m2, 'STO-3G')
            S, T, V = gaussian.build_one_electron(b2.functions, list(zip(m2.Z, m2.coords)))
            vals.append((S, T, V, gaussian.build_eri(b2.functions),
                         gaussian.build_mixed_overlap(b2.functions, bs.functions)))
        for k, key in enumerate(('S', 'T', 'V', 'ERI', 'half_S')):
            fd = (vals[0][k] - vals[1][k]) / (2 * h)
            assert np.abs(fd - d[key][A][x]).max() < 5e-9, (A, x, key)


def test_one_rhf_aat_golden_set_is_orbital_independent():
    """The canonical and non-canonical tests hold the same array (only the
    vo blocks of U enter the Hartree-Fock AAT)."""
    rhf = [e for e in GOLDEN if e["call"]["function"] == "compute_RHF_AATs" and runnable(e)]
    assert len(rhf) == 2
    assert np.abs(np.array(rhf[0]["aat"]) - np.array(rhf[1]["aat"])).max() < 1e-9


def _oracle(e):
    c = e["call"]
    if c["function"] == "compute_RHF_AATs":
        return analytic.rhf_aat(golden_params(e), c["orbitals"])
    if c["function"] == "compute_MP2_AATs":
        return analytic.mp2_aat(golden_params(e), c["normalization"], c["orbitals"])
    method = "CISD" if c["function"] == "compute_CISD_AATs" else "CID"
    return analytic.ci_aat(golden_params(e), method, c["normalization"], c["orbitals"])


@pytest.mark.parametrize("e", [e for e in GOLDEN if runnable(e)], ids=golden_id)
def test_oracle_analytic_aat_golden(e):
    """Hartree-Fock, MP2, CID and CISD analytic AATs (canonical /
    non-canonical, with and without frozen core) against the reference
    tests' arrays."""
    I = _oracle(e)
    assert np.abs(I - np.array(e["aat"])).max() < e["tolerance"]


def _h2o2(method, freeze_core):
    e = next(e for e in GOLDEN if runnable(e) and e["parameters"]["method"] == method
             and e["parameters"]["freeze_core"] == freeze_core)
    return golden_params(e)


@pytest.mark.parametrize("method", ["CISD", "CID"])
def test_ci_response_is_the_derivative_of_the_amplitudes(method):
    """The first-order amplitudes and projected energy derivative of one
    field and one nuclear perturbation equal central differences of the
    amplitudes and energy re-solved with F + h dF, <pq|rs> + h d<pq|rs>
    (the response equations are the derivative of the amplitude equations;
    this does not use the reference's residual formulas)."""
    params = _h2o2(method, True)
    ctx = analytic.Context(params)
    _, _, _, t_ = ctx.C_list
    O = ctx.no - ctx.nfzc
    Fa, Ea, ea = ctx.F[t_, t_], ctx.ERI[t_, t_, t_, t_], ctx.eps[t_]
    singles = method == "CISD"

    def solve(F, A):
        if singles:
            E, t1, t2, _ = spatial.solve_cisd(F, A, ea, O, params)
            return E, t1, t2
        E, t2, _ = spatial.solve_cid(F, A, ea, O, params)
        return E, None, t2

    E, t1, t2 = solve(Fa, Ea)
    U, h_core = ctx.magnetic(1, 'non-canonical')
    pert = [(analytic._df_magnetic(ctx, U, h_core), analytic._dERI(U, ctx.ERI, (-1, -1, 1, 1)))]
    U_R, F1, S1, E1, _ = ctx.nuclear(2, 0, 'non-canonical')
    pert.append((analytic._df_nuclear(ctx, U_R, F1, S1), E1 + analytic._dERI(U_R, ctx.ERI, (1, 1, 1, 1))))
    h = 1e-4
    for df, dA in pert:
        dF, dE4 = df[t_, t_], dA[t_, t_, t_, t_]
        dE, dt1, dt2, _ = analytic.ci_response(Fa, Ea, ea, O, E, t1, t2, dF, dE4, 0.0, params)
        Ep, t1p, t2p = solve(Fa + h * dF, Ea + h * dE4)
        Em, t1m, t2m = solve(Fa - h * dF, Ea - h * dE4)
        scale = np.abs(dt2).max()
        assert scale > 1e-4
        assert np.abs((t2p - t2m) / (2 * h) - dt2).max() < 1e-6 * scale
        if singles:
            assert np.abs((t1p - t1m) / (2 * h) - dt1).max() < 1e-6 * max(scale, np.abs(dt1).max())
        assert abs((Ep - Em) / (2 * h) - dE) < 1e-7 * max(1.0, abs(dE))


@pytest.mark.parametrize("method", ["CISD", "CID"])
def test_density_gradient_equals_projected_derivative(method):
    """For every perturbation the one- and two-particle density gradient
    (the reference's guess, sum dF.D + sum d<pq|rs>.G) equals the converged
    projected energy derivative (Hellmann-Feynman for the variational CI
    root), which pins the restated densities and the response solution."""
    rec = {}
    analytic.ci_aat(_h2o2(method, True), method, 'full', 'canonical', record=rec)
    assert len(rec['dE']) == 3 + 3 * 4
    for g, d in zip(rec['guess'], rec['dE']):
        assert abs(g - d) < 1e-10 * max(1.0, abs(d))


@pytest.mark.parametrize("e", [e for e in GOLDEN if not runnable(e)], ids=golden_id)
def test_analytic_golden_on_the_host(e):
    pytest.skip("19 functions with d shells: the host integral engine takes minutes; this golden runs on the "
                "GPU drop-in (tests/test_gpu_analytic.py)")


def test_631gs_basis_water():
    """6-31G* on H2O: 19 Cartesian functions (O 1s, 2sp, 3sp, six d; H two
    s each), every contracted function normalised, one frozen core orbital."""
    mol = Molecule.from_string(H2O)
    bs = BasisSet(mol, '6-31G*')
    assert bs.nbf() == 19 and bs.n_frozen_core() == 1
    assert sum(1 for f in bs.functions if sum(f.lmn) == 2) == 6
    S = gaussian.build_mixed_overlap(bs.functions, bs.functions)
    assert np.abs(np.diag(S) - 1.0).max() < 1e-12
    assert np.linalg.eigvalsh(S).min() > 1e-3
