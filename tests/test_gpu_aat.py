"""Finite-difference AATs on the GPU: the determinant-lemma records and
amplitude-determinant sums (apy_aat_gamma / apy_aat_terms) against the
oracle's explicit determinants, the AAT class against the oracle on the
same wavefunctions, and the drop-in driver (compute_parallel_aats) against
the AAT arrays held by the reference's tests."""
import ctypes

import numpy as np
import pytest
import torch

from apyib_amd import _lib
from apyib_amd.aats import AAT, NT, so_overlap, spatial_terms
from apyib_amd.hamiltonian import Hamiltonian
from apyib_amd.hf_wfn import hf_wfn
from apyib_amd.kernels import stream_of, workspace
from apyib_amd.parallel import compute_parallel_aats
from oracle import aats as oaats
from oracle import fd
from tests.test_oracle_aat import GOLDEN, golden_id, golden_params, golden_tolerance

pytestmark = pytest.mark.gpu


def _c(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def _d(x, dev):
    return torch.from_numpy(np.ascontiguousarray(x, dtype=np.complex128)).to(dev)


def _vp(t):
    return ctypes.c_void_p(t.data_ptr() if t is not None else 0)


def _gamma(S_list, no, nf, dev):
    lib = _lib.load()
    n = S_list[0].shape[0]
    S = _d(np.stack(S_list), dev)
    rec = int(lib.apy_aat_gamma_record(no, nf, n - no))
    g = torch.empty((len(S_list), rec), dtype=torch.complex128, device=dev)
    nb = int(lib.apy_aat_gamma_workspace(n, no, len(S_list)))
    ws = workspace(nb, dev)
    _lib.check(lib.apy_aat_gamma(n, no, nf, len(S_list), _vp(S), n * n, _vp(g), _vp(ws), nb, stream_of(dev)),
               "apy_aat_gamma")
    return g, rec


def _terms(g, rec, oa, nv, jobs, dev):
    lib = _lib.load()
    table = (_lib.AatJob * len(jobs))()
    for q, (gi, b1, b2, k1, k2) in enumerate(jobs):
        table[q] = _lib.AatJob(gi, *[t.data_ptr() if t is not None else None for t in (b1, b2, k1, k2)])
    out = torch.empty((len(jobs), NT), dtype=torch.complex128, device=dev)
    nb = int(lib.apy_aat_terms_workspace(oa, nv, len(jobs)))
    ws = workspace(nb, dev)
    _lib.check(lib.apy_aat_terms(oa, nv, _vp(g), rec, g.shape[0], len(jobs), table, _vp(out), _vp(ws), nb, stream_of(dev)),
               "apy_aat_terms")
    return out.cpu().numpy()


def test_gamma_records_match_numpy(dev):
    rng = np.random.default_rng(0)
    no, nv, nf = 5, 4, 1
    n = no + nv
    mats = [np.eye(n) + 0.3 * _c(rng, n, n) for _ in range(3)]
    P = _c(rng, n, n)
    P[0, 0] = 0.0                       # the first pivot needs a row swap
    mats.append(P)
    g, rec = _gamma(mats, no, nf, dev)
    g = g.cpu().numpy()
    oa = no - nf
    for S, r in zip(mats, g):
        Pi = np.linalg.inv(S[:no, :no])
        X = S[no:, :no] @ Pi
        Y = Pi @ S[:no, no:]
        W = S[no:, no:] - S[no:, :no] @ Y
        scale = max(1.0, np.abs(Pi).max())
        assert abs(r[0] - np.linalg.det(S[:no, :no])) < 1e-12 * abs(r[0])
        assert np.abs(r[16:16 + oa * nv].reshape(oa, nv) - X[:, nf:].T).max() < 1e-12 * scale
        assert np.abs(r[16 + oa * nv:16 + 2 * oa * nv].reshape(oa, nv) - Y[nf:, :]).max() < 1e-12 * scale
        o = 16 + 2 * oa * nv
        assert np.abs(r[o:o + oa * oa].reshape(oa, oa) - Pi[nf:, nf:]).max() < 1e-12 * scale
        assert np.abs(r[o + oa * oa:o + oa * oa + nv * nv].reshape(nv, nv) - W).max() < 1e-11 * scale ** 2


def test_gamma_singular_block_gives_zero_record(dev):
    rng = np.random.default_rng(1)
    n, no = 6, 3
    S = _c(rng, n, n)
    S[:no, 1] = 0.0                     # occupied block with a zero column
    g, _ = _gamma([S], no, 0, dev)
    assert torch.count_nonzero(g).item() == 0


@pytest.mark.parametrize("no,nv,nf,seed", [(3, 3, 0, 0), (4, 3, 1, 1), (2, 5, 0, 2), (5, 2, 2, 3)])
def test_spatial_terms_match_explicit_determinants(dev, no, nv, nf, seed):
    """Every term of one overlap class (aats.py:713-1047) against the
    oracle's explicit determinant tensors; amplitudes without symmetry."""
    rng = np.random.default_rng(seed)
    n, oa = no + nv, no - nf
    S = np.eye(n) + 0.3 * _c(rng, n, n)
    B1, K1 = _c(rng, oa, nv), _c(rng, oa, nv)
    B2, K2 = _c(rng, oa, oa, nv, nv), _c(rng, oa, oa, nv, nv)
    g, rec = _gamma([S], no, nf, dev)
    r = _terms(g, rec, oa, nv, [(0, _d(B1, dev), _d(B2, dev), _d(K1, dev), _d(K2, dev))], dev)[0]
    got = {k: v * r[0] ** 2 for k, v in spatial_terms(r).items()}
    dets = oaats.compute_all_dets(S, no, nv, nf)
    ref = oaats.spatial_class_terms(dets, (B1, B2), (K1, K2), 1.0, 1.0)
    ref.update({k: v for k, v in oaats.spatial_class_terms(dets, (B1, B2), None, 1.0, 1.0).items() if k in ('S0', 'D0')})
    ref.update({k: v for k, v in oaats.spatial_class_terms(dets, None, (K1, K2), 1.0, 1.0).items() if k in ('0S', '0D')})
    for k in ref:
        assert abs(got[k] - ref[k]) < 1e-12 * max(1.0, abs(ref[k])), k


@pytest.mark.parametrize("no,nv,seed", [(4, 4, 0), (2, 6, 1), (6, 2, 2)])
def test_spin_orbital_overlap_matches_determinant_sum(dev, no, nv, seed):
    """<bra|ket> of two CISD vectors in different orbital bases (the sums of
    compute_SO_I_*, aats.py:161-516) against the explicit determinant matrix."""
    rng = np.random.default_rng(seed)
    n = no + nv
    S = np.eye(n) + 0.3 * _c(rng, n, n)
    A4 = lambda t: 0.25 * (t - t.swapaxes(2, 3) - t.swapaxes(0, 1) + t.swapaxes(0, 1).swapaxes(2, 3))
    B1, K1 = _c(rng, no, nv), _c(rng, no, nv)
    B2, K2 = A4(_c(rng, no, no, nv, nv)), A4(_c(rng, no, no, nv, nv))
    states = oaats._so_states(no, nv, True)
    Dm = oaats.so_det_matrix(S, no, states)
    g, rec = _gamma([S], no, 0, dev)
    r = _terms(g, rec, no, nv, [(0, _d(B1, dev), _d(B2, dev), _d(K1, dev), _d(K2, dev))], dev)[0]
    for A0, K0 in ((0, 0), (1, 0), (0, 1), (1, 1)):
        ref = oaats._so_coeffs(states, A0, B1, B2) @ Dm @ oaats._so_coeffs(states, K0, K1, K2)
        assert abs(r[0] * so_overlap(r, A0, K0) - ref) < 1e-12 * abs(ref)


def _oracle_inputs(params, dR, dH):
    wfns, ndocc, nbf, nfzc = fd.compute_AAT(params, dR, dH)
    w = lambda k: [x[0] for x in wfns[k]]
    b = lambda k: [x[1] for x in wfns[k]]
    t = lambda k: [x[2] for x in wfns[k]]
    C0, f0, T0 = wfns['u']
    wfn = hf_wfn(Hamiltonian(params))
    args = (params, wfn, C0, f0, T0, w('np'), w('nn'), b('np'), b('nn'), t('np'), t('nn'),
            w('mp'), w('mn'), b('mp'), b('mn'), t('mp'), t('mn'), dR, dH)
    return wfns, ndocc, nbf, nfzc, args


@pytest.mark.parametrize("method", ["CISD", "MP2", "CID", "RHF", "CISD_SO", "CID_SO"])
def test_aat_class_matches_oracle_on_same_wavefunctions(dev, method):
    """(H2)2/STO-3G wavefunctions from the oracle chain; the GPU AAT class and
    the oracle's explicit-determinant AATs on them, both normalisations."""
    from tests.molecules import H2_2
    params = {'geom': H2_2, 'basis': 'STO-3G', 'method': method, 'e_convergence': 1e-12, 'd_convergence': 1e-12,
              'DIIS': True, 'freeze_core': False, 'F_el': [0.0] * 3, 'F_mag': [0.0] * 3, 'max_iterations': 120}
    dR = dH = 1e-4
    wfns, ndocc, nbf, nfzc, args = _oracle_inputs(params, dR, dH)
    so_form = method.endswith('_SO')
    ov = fd.overlap_sets(wfns, so_form)
    T = fd.amplitude_sets(wfns)
    gpu = AAT(*args)
    for norm in ('full', 'intermediate'):
        I = gpu.compute_aats(norm)
        for a in range(12):
            for b in range(3):
                if so_form:
                    ref = oaats.so_aat(method, 2 * ndocc, 2 * (nbf - ndocc), ov, T, dR, dH, a, b, norm)
                else:
                    ref = oaats.spatial_aat(method, ndocc, nbf - ndocc, nfzc, ov, T, dR, dH, a, b, norm)
                assert abs(I[a, b] - ref) < 1e-9, (norm, a, b)
        single = gpu.compute_SO_aats(4, 1, norm) if so_form else gpu.compute_spatial_aats(4, 1, norm)
        assert abs(single - I[4, 1]) < 1e-12
        # the reference's I_00, I_0S, ... parts, element by element
        P = gpu.compute_aat_terms(norm)
        for a, b in ((0, 0), (4, 1), (11, 2)):
            if so_form:
                ref = oaats.so_aat(method, 2 * ndocc, 2 * (nbf - ndocc), ov, T, dR, dH, a, b, norm, split=True)
            else:
                ref = oaats.spatial_aat(method, ndocc, nbf - ndocc, nfzc, ov, T, dR, dH, a, b, norm, split=True)
            for k in oaats.SPLIT_TERMS:
                assert abs(P[k][a, b] - ref[k]) < 1e-9, (norm, a, b, k)
        if so_form:
            for k in oaats.SPLIT_TERMS:
                assert abs(getattr(gpu, f"compute_SO_I_{k}")(4, 1, norm) - P[k][4, 1]) < 1e-12, k


@pytest.mark.parametrize("e", GOLDEN, ids=golden_id)
def test_compute_parallel_aats_golden(dev, e):
    """The drop-in driver (energy -> finite_difference.compute_AAT -> AAT) on
    the GPU against the reference test's AAT array, at its own tolerance."""
    I = compute_parallel_aats(golden_params(e), e["dR"], e["dH"], e["normalization"])
    assert np.abs(I - np.array(e["aat"])).max() < golden_tolerance(e)


@pytest.mark.parametrize("e", [e for e in GOLDEN if e["components"]], ids=golden_id)
def test_aat_terms_golden(dev, e):
    """The reference tests' I_00, I_0D, I_D0 and I_DD arrays against the
    drop-in chain's parts (energy -> finite_difference.compute_AAT -> AAT).
    Spatial MP2/CID have no 0D/D0 parts in the reference's code; the stored
    0D and D0 arrays cancel."""
    from apyib_amd.energy import energy
    from apyib_amd.fin_diff import finite_difference
    params = golden_params(e)
    E_list, T_list, C, basis = energy(params)
    wfn = hf_wfn(Hamiltonian(params))
    fdo = finite_difference(params, basis, C)
    r = fdo.compute_AAT(e["dR"], e["dH"])
    gpu = AAT(params, wfn, C, basis, T_list, *r, e["dR"], e["dH"])
    P = gpu.compute_aat_terms(e["normalization"])
    c = {k[2:4]: np.array(v) for k, v in e["components"].items()}
    tol = golden_tolerance(e)
    for k in (("00", "0D", "D0", "DD") if e["form"] == "SO" else ("00", "DD")):
        assert np.abs(P[k] - c[k]).max() < tol, k
    assert np.abs(P["0D"] + P["D0"] - c["0D"] - c["D0"]).max() < tol
