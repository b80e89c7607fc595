"""Thin torch-facing wrappers over the C ABI.

Tensors are complex128 CUDA (HIP) tensors; the wrappers pass raw device
pointers, element strides and the current HIP stream of the tensor's
device.  No computation happens in Python.
"""
import ctypes
import json
import os

import numpy as np
import torch

from . import _lib

_WS = {}


def stream_of(device):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def workspace(nbytes, device):
    """A per-device scratch buffer of at least ``nbytes`` bytes (grown, never
    shrunk).  Calls on one device and stream are ordered, so sharing it is
    safe for stream-ordered use."""
    key = torch.device(device)
    buf = _WS.get(key)
    if buf is None or buf.numel() < nbytes:
        buf = torch.empty(max(int(nbytes), 1 << 20), dtype=torch.uint8, device=key)
        _WS[key] = buf
    return buf


def _require(t, name):
    if not isinstance(t, torch.Tensor) or t.dtype != torch.complex128 or not t.is_cuda:
        raise TypeError(f"{name} must be a complex128 device tensor")


def ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def zgemm_raw(*, m, n, k, a, a_layout, lda, b, b_layout, ldb, c, ldc, alpha=1.0, beta=0.0,
              conj_a=False, conj_b=False, batch=1, stride_a=0, stride_b=0, stride_c=0,
              reduce_batch=False, split_k=0, tile=0, tune=False, reps=3, gauss3m=False):
    """C = alpha * op(A) op(B) + beta * C on raw strides (see apy_zgemm).
    With tune=True the candidate plans are timed first (apy_zgemm_tune) and
    the fastest is cached for the shape; returns (tile, split, ms) then.
    gauss3m: three real products per complex multiply-add instead of four."""
    lib = _lib.load()
    alpha = complex(alpha)
    beta = complex(beta)
    d = _lib.GemmDesc(
        m, n, k, a_layout, b_layout, int(conj_a), int(conj_b),
        alpha.real, alpha.imag, beta.real, beta.imag,
        a.data_ptr(), lda, stride_a,
        b.data_ptr(), ldb, stride_b,
        c.data_ptr(), ldc, stride_c,
        batch, int(reduce_batch), int(split_k), int(tile), int(gauss3m))
    nbytes = lib.apy_zgemm_workspace(ctypes.byref(d))
    ws = workspace(nbytes, c.device) if nbytes else None
    wsp = ctypes.c_void_p(ws.data_ptr() if ws is not None else 0)
    if tune:
        ti, sp, ms = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        st = lib.apy_zgemm_tune(ctypes.byref(d), wsp, ctypes.c_size_t(nbytes), int(reps), stream_of(c.device),
                                ctypes.byref(ti), ctypes.byref(sp), ctypes.byref(ms))
        _lib.check(st, "apy_zgemm_tune")
        return ti.value, sp.value, ms.value
    st = lib.apy_zgemm(ctypes.byref(d), wsp, ctypes.c_size_t(nbytes), stream_of(c.device))
    _lib.check(st, "apy_zgemm")
    return c


def zgemm_plan(**kw):
    """(tile, split) apy_zgemm would choose for the same arguments."""
    lib = _lib.load()
    alpha, beta = complex(kw.get('alpha', 1.0)), complex(kw.get('beta', 0.0))
    a, b, c = kw['a'], kw['b'], kw['c']
    d = _lib.GemmDesc(kw['m'], kw['n'], kw['k'], kw['a_layout'], kw['b_layout'], 0, 0, alpha.real, alpha.imag,
                      beta.real, beta.imag, a.data_ptr(), kw['lda'], kw.get('stride_a', 0), b.data_ptr(), kw['ldb'],
                      kw.get('stride_b', 0), c.data_ptr(), kw['ldc'], kw.get('stride_c', 0), kw.get('batch', 1),
                      int(kw.get('reduce_batch', False)), int(kw.get('split_k', 0)), int(kw.get('tile', 0)))
    ti, sp = ctypes.c_int(), ctypes.c_int()
    _lib.check(lib.apy_zgemm_plan(ctypes.byref(d), ctypes.byref(ti), ctypes.byref(sp)), "apy_zgemm_plan")
    return ti.value, sp.value


def _layout_a(A):
    if A.stride(1) == 1:
        return _lib.APY_MK, A.stride(0)
    if A.stride(0) == 1:
        return _lib.APY_KM, A.stride(1)
    raise ValueError("A must have a unit stride along one dimension")


def _layout_b(B):
    if B.stride(1) == 1:
        return _lib.APY_KN, B.stride(0)
    if B.stride(0) == 1:
        return _lib.APY_NK, B.stride(1)
    raise ValueError("B must have a unit stride along one dimension")


def mm(A, B, C=None, alpha=1.0, beta=0.0, conj_a=False, conj_b=False, split_k=0, tile=0, tune=False):
    """C = alpha*A@B + beta*C for 2-D complex128 device tensors (views with a
    unit stride along one dimension are accepted without copies)."""
    _require(A, "A")
    _require(B, "B")
    m, k = A.shape
    k2, n = B.shape
    if k != k2:
        raise ValueError("inner dimensions differ")
    if C is None:
        C = torch.empty((m, n), dtype=torch.complex128, device=A.device)
        beta = 0.0
    _require(C, "C")
    if C.shape != (m, n) or (C.stride(1) != 1 and n > 1):
        raise ValueError("C must be (m, n) with unit column stride")
    al, lda = _layout_a(A)
    bl, ldb = _layout_b(B)
    return zgemm_raw(m=m, n=n, k=k, a=A, a_layout=al, lda=lda, b=B, b_layout=bl, ldb=ldb,
                     c=C, ldc=C.stride(0), alpha=alpha, beta=beta, conj_a=conj_a, conj_b=conj_b,
                     split_k=split_k, tile=tile, tune=tune)


def zdot_many(xs, y, conj=False):
    """[sum_l op(x_m[l]) y[l] for x_m in xs] (apy_zdot_many, op = conj when
    conj=True; deterministic), as a host complex array."""
    lib = _lib.load()
    y = y.reshape(-1)
    out = np.zeros(len(xs), dtype=np.complex128)
    part = torch.empty(_lib.APY_MAXVEC, dtype=torch.complex128, device=y.device)
    nb = int(lib.apy_reduce_workspace(y.numel(), _lib.APY_MAXVEC))
    ws = workspace(nb, y.device)
    for s0 in range(0, len(xs), _lib.APY_MAXVEC):
        chunk = [x.reshape(-1) for x in xs[s0:s0 + _lib.APY_MAXVEC]]
        if any(x.numel() != y.numel() for x in chunk):
            raise ValueError("zdot_many: length mismatch")
        arr = (ctypes.c_void_p * len(chunk))(*[x.data_ptr() for x in chunk])
        _lib.check(lib.apy_zdot_many(len(chunk), arr, ctypes.c_void_p(y.data_ptr()), y.numel(), int(conj),
                                     ctypes.c_void_p(part.data_ptr()), ctypes.c_void_p(ws.data_ptr()), nb,
                                     stream_of(y.device)), "apy_zdot_many")
        out[s0:s0 + len(chunk)] = part[:len(chunk)].cpu().numpy()
    return out


def plans_export():
    """The GEMM plans cached on the current device (autotuner results or
    loaded), as a list of dicts."""
    lib = _lib.load()
    n = ctypes.c_int64(0)
    _lib.check(lib.apy_zgemm_tune_export(None, 0, ctypes.byref(n)), "apy_zgemm_tune_export")
    rows = (ctypes.c_int64 * (9 * max(n.value, 1)))()
    _lib.check(lib.apy_zgemm_tune_export(rows, n.value, ctypes.byref(n)), "apy_zgemm_tune_export")
    names = ("m", "n", "k", "a_layout", "b_layout", "batch", "reduce_batch", "tile", "split")
    return [dict(zip(names, rows[9 * q:9 * q + 9])) for q in range(n.value)]


_DEFAULT_PLANS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "plans_gfx950.json")
_DEFAULT_DONE = set()


def ensure_default_plans(device):
    """Install the saved gfx950 GEMM plans (plans_gfx950.json beside this
    file, or the file named by APYIB_AMD_PLANS; "none" disables) for
    ``device``, once per process and device.  Shapes without a saved plan use
    the cost model."""
    path = os.environ.get("APYIB_AMD_PLANS", _DEFAULT_PLANS)
    dev = torch.device(device)
    if dev.type != "cuda" or path.lower() == "none" or not os.path.exists(path):
        return
    idx = dev.index if dev.index is not None else torch.cuda.current_device()
    if (path, idx) in _DEFAULT_DONE:
        return
    with torch.cuda.device(idx):
        plans_load(json.load(open(path)))
    _DEFAULT_DONE.add((path, idx))


def plans_clear():
    """Forget every cached GEMM plan (saved or tuned) on every device."""
    _lib.check(_lib.load().apy_zgemm_tune_clear(), "apy_zgemm_tune_clear")
    _DEFAULT_DONE.clear()


def plans_load(entries):
    """Install saved plans (the output of plans_export) in the process-wide
    cache; later GEMMs of those shapes use them instead of the cost model."""
    lib = _lib.load()
    for e in entries:
        d = _lib.GemmDesc()
        d.m, d.n, d.k = e["m"], e["n"], e["k"]
        d.a_layout, d.b_layout = e["a_layout"], e["b_layout"]
        d.batch, d.reduce_batch = max(1, e["batch"]), e["reduce_batch"]
        d.lda = e["k"] if e["a_layout"] == _lib.APY_MK else e["m"]
        d.ldb = e["n"] if e["b_layout"] == _lib.APY_KN else e["k"]
        d.ldc = e["n"]
        d.stride_a = d.m * d.k
        d.stride_b = d.k * d.n
        d.stride_c = 0 if e["reduce_batch"] else d.m * d.n
        d.alpha_re = 1.0
        d.a = d.b = d.c = None                # shape-only descriptor (never launched)
        _lib.check(lib.apy_zgemm_tune_set(ctypes.byref(d), int(e["tile"]), int(e["split"])), "apy_zgemm_tune_set")
