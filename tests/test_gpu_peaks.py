"""The roofline microbenchmarks (apy_peak_mfma_f64, apy_peak_copy,
apy_peak_read): the copy copies, the read sums, the matrix-core chains
produce their closed-form sums, and the flop count matches the launch."""
import ctypes

import numpy as np
import pytest
import torch

from apyib_amd import _lib
from apyib_amd.kernels import stream_of

pytestmark = pytest.mark.gpu


def test_copy_copies(dev):
    lib = _lib.load()
    n = 3 * (1 << 20) + 17                                 # ragged tail
    g = torch.Generator(device="cpu").manual_seed(5)
    src = torch.complex(torch.randn(n, generator=g, dtype=torch.float64),
                        torch.randn(n, generator=g, dtype=torch.float64)).to(dev)
    dst = torch.full_like(src, complex(np.nan, np.nan))
    _lib.check(lib.apy_peak_copy(n, ctypes.c_void_p(src.data_ptr()), ctypes.c_void_p(dst.data_ptr()), 64,
                                 stream_of(dev)), "apy_peak_copy")
    assert torch.equal(src, dst)
    out = torch.zeros(64 * 256, dtype=torch.float64, device=dev)
    _lib.check(lib.apy_peak_read(n, ctypes.c_void_p(src.data_ptr()), ctypes.c_void_p(out.data_ptr()), 64,
                                 stream_of(dev)), "apy_peak_read")
    ref = float((src.real + src.imag).sum().cpu())
    assert abs(float(out.sum().cpu()) - ref) <= 1e-9 * float((src.real.abs() + src.imag.abs()).sum().cpu())
    assert lib.apy_peak_copy(n, ctypes.c_void_p(src.data_ptr()), ctypes.c_void_p(src.data_ptr()), 64,
                             stream_of(dev)) == _lib.APY_E_ARG


def test_mfma_chains_sum_and_flop_count(dev):
    """Each wave's A operand holds a_l = 1e-3 (l + 1) of its lanes (thread l of
    the workgroup), B holds b = 1 + 1e-6 block: after `iters` steps every entry
    of a chain's 16 x 16 block is iters * b * sum_k a(row + 16 k), so a lane's
    output (4 chains x its 4 entries) is known in closed form."""
    lib = _lib.load()
    blocks, iters = 3, 7
    out = torch.zeros(blocks * 256, dtype=torch.float64, device=dev)
    _lib.check(lib.apy_peak_mfma_f64(blocks, iters, ctypes.c_void_p(out.data_ptr()), stream_of(dev)),
               "apy_peak_mfma_f64")
    got = out.cpu().numpy().reshape(blocks, 256)
    for blk in range(blocks):
        b = 1.0 + 1e-6 * blk
        for t in range(256):
            w, lane = divmod(t, 64)
            a = lambda l: 1e-3 * (64 * w + l + 1)
            # lane holds D rows lane/16 + 4 r (r = 0..3), column lane % 16; A[i][k] = a(i + 16 k)
            s = sum(iters * b * sum(a(i + 16 * k) for k in range(4)) for i in (lane // 16 + 4 * r for r in range(4)))
            assert abs(got[blk, t] - 4 * s) <= 1e-12 * abs(4 * s)
    assert lib.apy_peak_mfma_f64_flops(blocks, iters) == blocks * 4 * iters * 4 * 16 * 16 * 4 * 2
