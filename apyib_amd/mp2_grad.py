"""Spin-adapted MP2 energy derivatives on the device, in the reference's own
forms (compute_MP2_Hessian, analytic.py:307-335).

t2 [o, o, v, v] are closed-shell MP2 amplitudes (mp2_wfn.py:54, pair
symmetric); f_grad [nd, n, n] and eri_grad [nd, o, o, v, v] (<ij|ab>) are
nd directions (a derivative Fock matrix and the derivative integrals of one
perturbation each).  Plain products, no conjugation, as the reference
writes them.

    t2_response        t2_grad = (eri_grad + Fock terms) / D   analytic.py:326-331
    gradient_densities dE/dERI_grad, dE/df_oo, dE/df_vv         analytic.py:320-324
    energy_gradient    E_grad of every direction from the densities
"""
import ctypes

import numpy as np
import torch

from . import _lib
from .kernels import stream_of, workspace

CDT = torch.complex128


def _vp(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _dev_c(x, dev):
    return torch.as_tensor(x).to(dev, CDT).contiguous()


def t2_response(t2, f_grad, eri_grad, eps, form=0):
    """First-order amplitudes of nd directions at once (apy_mp2_sp_response).

    form 0: the index order of analytic.py:327-330 (and of the nuclear
    first-order amplitudes, analytic_aats.py:446-451); form 1: that of the
    magnetic-field amplitudes (analytic_aats.py:346-350).  f_grad [n, n] or
    [nd, n, n]; eri_grad of the same leading shape, or None.  Returns
    [nd, o, o, v, v] (or [o, o, v, v] for a single unbatched direction)."""
    lib = _lib.load()
    dev = t2.device
    o, _, v, _ = t2.shape
    single = f_grad.dim() == 2
    f = _dev_c(f_grad, dev).reshape(-1, o + v, o + v)
    nd = f.shape[0]
    e = _dev_c(eri_grad, dev).reshape(nd, o, o, v, v) if eri_grad is not None else None
    ep = torch.as_tensor(np.asarray(eps, dtype=np.float64) if not isinstance(eps, torch.Tensor) else eps,
                         dtype=torch.float64).to(dev).contiguous()
    out = torch.empty((nd, o, o, v, v), dtype=CDT, device=dev)
    nbytes = int(lib.apy_mp2_sp_response_workspace(o, v, nd))
    ws = workspace(nbytes, dev)
    _lib.check(lib.apy_mp2_sp_response(o, v, nd, int(form), _vp(t2.contiguous()), _vp(f), _vp(e), _vp(ep),
                                       _vp(out), _vp(ws), nbytes, stream_of(dev)), "apy_mp2_sp_response")
    return out[0] if single else out


def gradient_densities(t2):
    """(dE/dERI_grad [o,o,v,v] = 2(2t - t^T), dE/df_oo [o,o], dE/df_vv [v,v])
    of E_grad (analytic.py:320-324), apy_mp2_sp_densities."""
    lib = _lib.load()
    dev = t2.device
    o, _, v, _ = t2.shape
    d_eri = torch.empty_like(t2)
    d_oo = torch.empty((o, o), dtype=CDT, device=dev)
    d_vv = torch.empty((v, v), dtype=CDT, device=dev)
    nbytes = int(lib.apy_mp2_sp_densities_workspace(o, v))
    ws = workspace(nbytes, dev)
    _lib.check(lib.apy_mp2_sp_densities(o, v, _vp(t2.contiguous()), _vp(d_eri), _vp(d_oo), _vp(d_vv), _vp(ws),
                                        nbytes, stream_of(dev)), "apy_mp2_sp_densities")
    return d_eri, d_oo, d_vv


def energy_gradient(densities, f_grad=None, eri_grad=None):
    """E_grad of analytic.py:320-324 for nd directions (apy_mp2_sp_directional):
    densities from gradient_densities; f_grad [nd, n, n] and/or eri_grad
    [nd, o, o, v, v].  Returns [nd] complex (device)."""
    lib = _lib.load()
    d_eri, d_oo, d_vv = densities
    dev = d_oo.device
    o, v = d_oo.shape[0], d_vv.shape[0]
    f = _dev_c(f_grad, dev).reshape(-1, o + v, o + v) if f_grad is not None else None
    e = _dev_c(eri_grad, dev).reshape(-1, o * o * v * v) if eri_grad is not None else None
    if f is None and e is None:
        raise ValueError("energy_gradient needs f_grad, eri_grad or both")
    nd = (f if f is not None else e).shape[0]
    if f is not None and e is not None and e.shape[0] != nd:
        raise ValueError("f_grad and eri_grad must hold the same number of directions")
    out = torch.empty(nd, dtype=CDT, device=dev)
    nbytes = int(lib.apy_mp2_sp_directional_workspace(nd))
    ws = workspace(nbytes, dev)
    _lib.check(lib.apy_mp2_sp_directional(o, v, nd, _vp(d_eri), _vp(d_oo), _vp(d_vv), _vp(f), _vp(e), _vp(out),
                                          _vp(ws), nbytes, stream_of(dev)), "apy_mp2_sp_directional")
    return out
