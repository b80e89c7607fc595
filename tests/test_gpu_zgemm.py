"""The complex GEMM on FP64 MFMA (apy_zgemm) against NumPy products: every
tile configuration, operand layout, conjugation flag, split-K count, alpha /
beta, batching and batch reduction, at shapes that are not multiples of the
tiles (edge masking)."""
import itertools

import numpy as np
import pytest
import torch

from apyib_amd import _lib, kernels

pytestmark = pytest.mark.gpu


def _c(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def _d(x, dev):
    return torch.from_numpy(np.ascontiguousarray(x)).to(dev)


def _op(X, conj):
    return np.conj(X) if conj else X


@pytest.mark.parametrize("tile", range(1, 10))
def test_tiles_layouts_conj_split(dev, tile):
    rng = np.random.default_rng(tile)
    m, n, k = 77, 45, 53
    for al, bl, ca, cb, split in itertools.product((_lib.APY_MK, _lib.APY_KM), (_lib.APY_KN, _lib.APY_NK),
                                                   (0, 1), (0, 1), (1, 3)):
        Am = _c(rng, m, k)
        Bm = _c(rng, k, n)
        Cm = _c(rng, m, n)
        A = _d(Am if al == _lib.APY_MK else Am.T, dev)
        B = _d(Bm if bl == _lib.APY_KN else Bm.T, dev)
        C = _d(Cm, dev)
        alpha, beta = 0.7 - 0.2j, 0.3 + 0.5j
        kernels.zgemm_raw(m=m, n=n, k=k, a=A, a_layout=al, lda=k if al == _lib.APY_MK else m, b=B, b_layout=bl,
                          ldb=n if bl == _lib.APY_KN else k, c=C, ldc=n, alpha=alpha, beta=beta, conj_a=ca,
                          conj_b=cb, split_k=split, tile=tile)
        ref = alpha * _op(Am, ca) @ _op(Bm, cb) + beta * Cm
        err = np.abs(C.cpu().numpy() - ref).max()
        assert err < 1e-12 * np.sqrt(k) * 4, (al, bl, ca, cb, split, err)


@pytest.mark.parametrize("tile,split", [(1, 1), (2, 4), (7, 2), (5, 3)])
def test_batched_and_reduced(dev, tile, split):
    rng = np.random.default_rng(10 + tile)
    m, n, k, nb = 40, 70, 36, 5
    Am, Bm = _c(rng, nb, m, k), _c(rng, nb, k, n)
    A, B = _d(Am, dev), _d(Bm, dev)
    C = torch.zeros((nb, m, n), dtype=torch.complex128, device=dev)
    kernels.zgemm_raw(m=m, n=n, k=k, a=A, a_layout=_lib.APY_MK, lda=k, b=B, b_layout=_lib.APY_KN, ldb=n, c=C, ldc=n,
                      batch=nb, stride_a=m * k, stride_b=k * n, stride_c=m * n, split_k=split, tile=tile)
    assert np.abs(C.cpu().numpy() - Am @ Bm).max() < 1e-12 * k
    R = _d(_c(rng, m, n), dev)
    R0 = R.cpu().numpy()
    kernels.zgemm_raw(m=m, n=n, k=k, a=A, a_layout=_lib.APY_MK, lda=k, b=B, b_layout=_lib.APY_KN, ldb=n, c=R, ldc=n,
                      batch=nb, stride_a=m * k, stride_b=k * n, stride_c=0, reduce_batch=True, beta=1.0,
                      split_k=split, tile=tile)
    assert np.abs(R.cpu().numpy() - (R0 + (Am @ Bm).sum(0))).max() < 1e-12 * k * nb


def test_degenerate_shapes(dev):
    """m, n or k of 1 and a k shorter than one LDS stage."""
    rng = np.random.default_rng(3)
    for m, n, k in ((1, 1, 1), (1, 37, 5), (29, 1, 3), (17, 19, 1), (64, 64, 2)):
        Am, Bm = _c(rng, m, k), _c(rng, k, n)
        C = kernels.mm(_d(Am, dev), _d(Bm, dev))
        assert np.abs(C.cpu().numpy() - Am @ Bm).max() < 1e-13 * max(1, k)


@pytest.mark.parametrize("split", [32, 57, 200])
def test_skinny_many_slices(dev, split):
    """Small outputs split along a long K into many slices (the wide
    split-K reduction): plain, with beta, and batched with a shared A."""
    rng = np.random.default_rng(split)
    m, n, k, nb = 20, 37, 9001, 3
    Am, Bm, Cm = _c(rng, m, k), _c(rng, nb, k, n), _c(rng, nb, m, n)
    A, B, C = _d(Am, dev), _d(Bm, dev), _d(Cm, dev)
    kernels.zgemm_raw(m=m, n=n, k=k, a=A, a_layout=_lib.APY_MK, lda=k, b=B[0], b_layout=_lib.APY_KN, ldb=n,
                      c=C[0], ldc=n, alpha=0.5 + 0.25j, beta=-0.75j, split_k=split, tile=1)
    ref = (0.5 + 0.25j) * Am @ Bm[0] - 0.75j * Cm[0]
    assert np.abs(C[0].cpu().numpy() - ref).max() < 1e-12 * np.sqrt(k) * 4
    D = torch.zeros((nb, m, n), dtype=torch.complex128, device=dev)
    kernels.zgemm_raw(m=m, n=n, k=k, a=A, a_layout=_lib.APY_MK, lda=k, b=B, b_layout=_lib.APY_KN, ldb=n, c=D, ldc=n,
                      batch=nb, stride_a=0, stride_b=k * n, stride_c=m * n, split_k=split, tile=2)
    assert np.abs(D.cpu().numpy() - Am @ Bm).max() < 1e-12 * np.sqrt(k) * 4


@pytest.mark.parametrize("tile", range(1, 10))
def test_gauss3m_tiles_layouts_conj_split(dev, tile):
    """The three-product form (gauss3m) on every tile, layout, conjugation
    and split-K count, with alpha / beta, batching and batch reduction: its
    error bound is that of the four-product form scaled by the operand
    magnitudes (|re| + |im|), which these unit-variance operands keep at the
    same order."""
    rng = np.random.default_rng(50 + tile)
    m, n, k = 77, 45, 53
    for al, bl, ca, cb, split in itertools.product((_lib.APY_MK, _lib.APY_KM), (_lib.APY_KN, _lib.APY_NK),
                                                   (0, 1), (0, 1), (1, 3)):
        Am, Bm, Cm = _c(rng, m, k), _c(rng, k, n), _c(rng, m, n)
        A = _d(Am if al == _lib.APY_MK else Am.T, dev)
        B = _d(Bm if bl == _lib.APY_KN else Bm.T, dev)
        C = _d(Cm, dev)
        alpha, beta = 0.7 - 0.2j, 0.3 + 0.5j
        kernels.zgemm_raw(m=m, n=n, k=k, a=A, a_layout=al, lda=k if al == _lib.APY_MK else m, b=B, b_layout=bl,
                          ldb=n if bl == _lib.APY_KN else k, c=C, ldc=n, alpha=alpha, beta=beta, conj_a=ca,
                          conj_b=cb, split_k=split, tile=tile, gauss3m=True)
        ref = alpha * _op(Am, ca) @ _op(Bm, cb) + beta * Cm
        err = np.abs(C.cpu().numpy() - ref).max()
        assert err < 1e-12 * np.sqrt(k) * 4, (al, bl, ca, cb, split, err)
    nb = 4
    Am, Bm = _c(rng, nb, m, k), _c(rng, nb, k, n)
    R = _d(_c(rng, m, n), dev)
    R0 = R.cpu().numpy()
    kernels.zgemm_raw(m=m, n=n, k=k, a=_d(Am, dev), a_layout=_lib.APY_MK, lda=k, b=_d(Bm, dev), b_layout=_lib.APY_KN,
                      ldb=n, c=R, ldc=n, batch=nb, stride_a=m * k, stride_b=k * n, stride_c=0, reduce_batch=True,
                      beta=1.0, split_k=2, tile=tile, gauss3m=True)
    assert np.abs(R.cpu().numpy() - (R0 + (Am @ Bm).sum(0))).max() < 1e-12 * k * nb


def test_gauss3m_real_and_imaginary_operands(dev):
    """Purely real and purely imaginary operands through the three-product
    form: the sums ar + ai reduce to one part, so the results are those of
    real products (up to the cancellation T3 - T1 - T2 of equal terms)."""
    rng = np.random.default_rng(77)
    m, n, k = 64, 96, 40
    for fa, fb in ((1.0, 1.0), (1j, 1.0), (1.0, 1j), (1j, 1j)):
        Am, Bm = fa * rng.standard_normal((m, k)), fb * rng.standard_normal((k, n))
        C = torch.zeros((m, n), dtype=torch.complex128, device=dev)
        A, B = _d(Am.astype(complex), dev), _d(Bm.astype(complex), dev)
        kernels.zgemm_raw(m=m, n=n, k=k, a=A, a_layout=_lib.APY_MK, lda=k, b=B, b_layout=_lib.APY_KN, ldb=n, c=C,
                          ldc=n, gauss3m=True)
        ref = Am @ Bm
        assert np.abs(C.cpu().numpy() - ref).max() < 1e-13 * k
