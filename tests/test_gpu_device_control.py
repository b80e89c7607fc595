"""The Jacobi + DIIS iteration controlled from the device (SOCI.solve_device:
apy_diis_slot, apy_zcombine_guarded, apy_iter_check; chunks of queued
iterations, the steady-state chunk replayed from a captured HIP graph)
against the same iteration controlled from the host (solve_host: the
scalars of every iteration read back, the DIIS system solved by NumPy),
which the oracle parity tests pin iterate for iterate.  Same iteration
count, energies, amplitudes and printed rows, for the spin-orbital and
spin-adapted solvers and the first-order response solver."""
import io
from contextlib import redirect_stdout

import numpy as np
import pytest
import torch

from apyib_amd import synthetic
from apyib_amd.so_ci import SOCI, SOIntegrals
from apyib_amd.spatial_ci import ResponseCI, SPIntegrals, SpatialCI
from tests.helpers import PARAMS

pytestmark = pytest.mark.gpu

MODES = [dict(graph=True), dict(graph=False), dict(heavy=True)]
MODE_IDS = ["graph", "queued", "per-iteration"]


def _d(x, dev):
    return torch.from_numpy(np.ascontiguousarray(x)).to(dev)


def _spatial_system(o, v, seed, std=0.05, noise=1e-3):
    rng = np.random.default_rng(seed)
    n = o + v
    E = std * (rng.standard_normal((n,) * 4) + 1j * rng.standard_normal((n,) * 4))
    E = 0.5 * (E + E.transpose(2, 3, 0, 1))
    eps = np.concatenate([np.sort(rng.uniform(-2.0, -0.5, o)), np.sort(rng.uniform(0.3, 2.0, v))])
    F = np.diag(eps).astype(complex) + noise * (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    return E, F, eps


def _run(solver, params, mode, print_level=0):
    if mode.get("heavy"):
        solver.CHUNK_MAX_ITERATION_S = 0.0
    buf = io.StringIO()
    with redirect_stdout(buf):
        if mode.get("host"):
            out = solver.solve(params, print_level, control="host")
        else:
            out = solver.solve(params, print_level, control="device", graph=mode.get("graph"))
    return out, buf.getvalue(), solver.t.clone()


def _rows(text):
    rows = []
    for line in text.splitlines():
        f = line.split()
        if len(f) == 10 and f[0].isdigit():
            rows.append([int(f[0])] + [float(x) for x in f[1:]])
    return np.array(rows)


def _compare(make, params, mode, tol_t=1e-10):
    (E_h, _, _, it_h), out_h, t_h = _run(make(), params, dict(host=True), print_level=1)
    (E_d, _, _, it_d), out_d, t_d = _run(make(), params, mode, print_level=1)
    assert it_d == it_h
    assert abs(E_d - E_h) < 1e-12 * max(1.0, abs(E_h))
    assert (t_d - t_h).abs().max().item() < tol_t
    rh, rd = _rows(out_h), _rows(out_d)
    assert rh.shape == rd.shape and np.array_equal(rh[:, 0], rd[:, 0])
    assert np.abs(rh[:, 1:4] - rd[:, 1:4]).max() < 1e-10                         # E, E_tot
    assert ("Not converged." in out_h) == ("Not converged." in out_d)
    return it_d


@pytest.mark.parametrize("mode", MODES, ids=MODE_IDS)
@pytest.mark.parametrize("singles", [True, False], ids=["CISD", "CID"])
def test_spin_orbital_device_control_equals_host(dev, mode, singles):
    O, V, seed = 6, 8, 5
    F, A, eps = synthetic.system(O, V, seed)
    ints = SOIntegrals.synthetic(O, V, seed, device=dev)
    make = lambda: SOCI(ints, _d(F, dev), eps, singles=singles)
    _compare(make, PARAMS, mode)


@pytest.mark.parametrize("mode", MODES, ids=MODE_IDS)
@pytest.mark.parametrize("o,v,seed", [(3, 6, 2), (5, 11, 7)])
def test_spatial_device_control_equals_host(dev, mode, o, v, seed):
    E4, F, eps = _spatial_system(o, v, seed)
    ints = SPIntegrals.from_mo(_d(E4, dev), o, singles=True)
    make = lambda: SpatialCI(ints, _d(F, dev), eps, singles=True)
    _compare(make, PARAMS, mode)


@pytest.mark.parametrize("mode", MODES, ids=MODE_IDS)
@pytest.mark.parametrize("max_it", [1, 2, 5, 9, 17])
def test_max_iterations_stop(dev, mode, max_it):
    """Stopping at max_iterations inside a queued chunk keeps that iterate
    (later queued iterations change nothing) and prints the same warning;
    17 iterations replay the captured chunk twice."""
    E4, F, eps = _spatial_system(4, 9, 3)
    ints = SPIntegrals.from_mo(_d(E4, dev), 4, singles=True)
    make = lambda: SpatialCI(ints, _d(F, dev), eps, singles=True)
    it = _compare(make, dict(PARAMS, max_iterations=max_it), mode)
    assert it == max_it


@pytest.mark.parametrize("mode", MODES, ids=MODE_IDS)
def test_response_device_control_equals_host(dev, mode):
    o, v, seed = 4, 9, 5
    E4, F, eps = _spatial_system(o, v, seed)
    rng = np.random.default_rng(seed + 20)
    n = o + v
    dE4 = 0.02 * (rng.standard_normal((n,) * 4) + 1j * rng.standard_normal((n,) * 4))
    dE4 = 0.5 * (dE4 + dE4.transpose(2, 3, 0, 1))
    dF = 0.02 * (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    ints = SPIntegrals.from_mo(_d(E4, dev), o, singles=True)
    base = SpatialCI(ints, _d(F, dev), eps, singles=True)
    E, _, _, _ = base.solve(PARAMS, control="host")
    t_ref = base.t.clone()
    dints = SPIntegrals.from_mo(_d(dE4, dev), o, singles=True)
    make = lambda: ResponseCI(ints, _d(F, dev), eps, True, t_ref, E, dints, _d(dF, dev))
    _compare(make, PARAMS, mode)


def test_graph_solve_repeated(dev):
    """Solving twice with fresh solvers (new graphs) and once more with the
    same solver object gives bit-identical amplitudes."""
    E4, F, eps = _spatial_system(5, 11, 7)
    ints = SPIntegrals.from_mo(_d(E4, dev), 5, singles=True)
    a = SpatialCI(ints, _d(F, dev), eps, singles=True)
    b = SpatialCI(ints, _d(F, dev), eps, singles=True)
    ra = a.solve(PARAMS, graph=True)
    ta = a.t.clone()
    rb = b.solve(PARAMS, graph=True)
    assert ra[3] == rb[3] and ra[0] == rb[0] and torch.equal(ta, b.t)
    rc = a.solve(PARAMS, graph=True)
    assert rc[3] == ra[3] and rc[0] == ra[0] and torch.equal(ta, a.t)
