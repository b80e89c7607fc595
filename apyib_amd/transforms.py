"""Device-level AO->MO transforms, Fock builds, overlaps, phase and MP2
(torch complex128 tensors in and out; all arithmetic in libapyib_amd.so)."""
import ctypes

import numpy as np
import torch

from . import _lib
from .kernels import mm, stream_of, workspace

CDT = torch.complex128


def _vp(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


def to_dev(x, device):
    """Host array (real or complex) -> contiguous complex128 device tensor."""
    if isinstance(x, torch.Tensor):
        return x.to(device=device, dtype=CDT).contiguous()
    return torch.from_numpy(np.ascontiguousarray(np.asarray(x, dtype=np.complex128))).to(device)


def eri_mo(eri_ao, C):
    """(pq|rs) = sum conj(C_mp) C_nq conj(C_lr) C_gs (mn|lg)  (utils.py:258-279).
    eri_ao: [nbf]^4 device complex; C: [nbf, nact] device complex."""
    lib = _lib.load()
    nbf, nact = C.shape
    C = C.contiguous()
    eri_ao = eri_ao.contiguous()
    out = torch.empty((nact,) * 4, dtype=CDT, device=C.device)
    nbytes = int(lib.apy_eri_mo_workspace(nbf, nact))
    ws = workspace(nbytes, C.device)
    _lib.check(lib.apy_eri_mo(nbf, nact, _vp(eri_ao), _vp(C), _vp(out), _vp(ws), nbytes, stream_of(C.device)),
               "apy_eri_mo")
    return out


def fock(h, eri_ao, D, scale=1.0):
    """F = h + scale * sum_ls D_ls (2 (mn|ls) - (ml|ns))."""
    lib = _lib.load()
    nbf = D.shape[0]
    F = torch.empty((nbf, nbf), dtype=CDT, device=D.device)
    _lib.check(lib.apy_fock_build(nbf, _vp(h.contiguous() if h is not None else None), _vp(eri_ao.contiguous()),
                                  _vp(D.contiguous()), float(scale), _vp(F), stream_of(D.device)), "apy_fock_build")
    return F


def fock_batch(h, eri_ao, D, scale=1.0):
    """F[b] = h[b] + scale * sum_ls D[b]_ls (2 (mn|ls)_b - (ml|ns)_b) for a
    batch: h, D [batch, n, n]; eri_ao [batch, n, n, n, n] or one [n, n, n, n]
    shared by all."""
    lib = _lib.load()
    batch, nbf = D.shape[0], D.shape[1]
    shared = eri_ao.dim() == 4
    F = torch.empty((batch, nbf, nbf), dtype=CDT, device=D.device)
    _lib.check(lib.apy_fock_build_batch(nbf, batch, _vp(h.contiguous()), _vp(eri_ao.contiguous()),
                                        0 if shared else nbf ** 4, _vp(D.contiguous()), float(scale), _vp(F),
                                        stream_of(D.device)), "apy_fock_build_batch")
    return F


def density(C_occ):
    """D = C_occ C_occ^H (utils.py:246)."""
    return mm(C_occ, C_occ.t(), conj_b=True)


def sandwich(A, M, B):
    """A^H M B (utils.py:252, 381)."""
    return mm(A.t(), mm(M, B), conj_a=True)


def mo_overlap(C_bra, S_ao, C_ket):
    """C_bra^H S_AO C_ket for one pair, or a batch when the inputs are 3-D
    ([batch, nbf, nmo]; S [batch, nbf, nbf] or shared [nbf, nbf])."""
    lib = _lib.load()
    batched = C_bra.dim() == 3
    Cb = C_bra if batched else C_bra.unsqueeze(0)
    Ck = C_ket if batched else C_ket.unsqueeze(0)
    S = S_ao if S_ao.dim() == 3 else S_ao.unsqueeze(0)
    nb, nbf, nmo = Cb.shape
    Cb, Ck, S = Cb.contiguous(), Ck.contiguous(), S.contiguous()
    out = torch.empty((nb, nmo, nmo), dtype=CDT, device=Cb.device)
    nbytes = int(lib.apy_mo_overlap_workspace(nbf, nmo, nb))
    ws = workspace(nbytes, Cb.device)
    _lib.check(lib.apy_mo_overlap(nbf, nmo, nb, _vp(Cb), nbf * nmo if Cb.shape[0] > 1 else 0, _vp(S),
                                  nbf * nbf if S.shape[0] > 1 else 0, _vp(Ck), nbf * nmo if Ck.shape[0] > 1 else 0,
                                  _vp(out), _vp(ws), nbytes, stream_of(Cb.device)), "apy_mo_overlap")
    return out if batched else out[0]


def phase_align(S_mo, C):
    """C[:, m] / (S_mm/|S_mm|) (utils.py:435-444)."""
    lib = _lib.load()
    nbf, ncol = C.shape
    C = C.contiguous()
    S_mo = S_mo.contiguous()
    out = torch.empty_like(C)
    _lib.check(lib.apy_phase_align(nbf, ncol, _vp(S_mo), S_mo.shape[1], _vp(C), _vp(out), stream_of(C.device)),
               "apy_phase_align")
    return out


def eri_so(eri_mo_t):
    """Chemists' spin-orbital tensor of utils.py:317-365 (not antisymmetrized)."""
    lib = _lib.load()
    m = eri_mo_t.shape[0]
    out = torch.empty((2 * m,) * 4, dtype=CDT, device=eri_mo_t.device)
    _lib.check(lib.apy_eri_so_from_mo(m, _vp(eri_mo_t.contiguous()), _vp(out), stream_of(eri_mo_t.device)),
               "apy_eri_so_from_mo")
    return out


def so_fock(f_mo):
    """utils.py:283-313."""
    lib = _lib.load()
    m = f_mo.shape[0]
    out = torch.empty((2 * m, 2 * m), dtype=CDT, device=f_mo.device)
    _lib.check(lib.apy_so_fock_from_mo(m, _vp(f_mo.contiguous()), _vp(out), stream_of(f_mo.device)),
               "apy_so_fock_from_mo")
    return out


def mp2_spatial(eri_mo_t, eps, O):
    """solve_MP2 (mp2_wfn.py:42-59) from chemists' MO integrals over the
    active space; returns (E device scalar, t2 [O,O,V,V])."""
    lib = _lib.load()
    m = eri_mo_t.shape[0]
    V = m - O
    dev = eri_mo_t.device
    e = torch.as_tensor(np.asarray(eps, dtype=np.float64)).to(dev)
    t2 = torch.empty((O, O, V, V), dtype=CDT, device=dev)
    E = torch.empty(1, dtype=CDT, device=dev)
    nbytes = int(lib.apy_mp2_spatial_workspace())
    ws = workspace(nbytes, dev)
    _lib.check(lib.apy_mp2_spatial(O, V, _vp(e), _vp(eri_mo_t.contiguous()), _vp(t2), _vp(E), _vp(ws), nbytes,
                                   stream_of(dev)), "apy_mp2_spatial")
    return E, t2


def mp2_so_batched(instances, device):
    """Canonical spin-orbital MP2 (mp2_wfn.py:64-87) for a list of
    (eps [n] float64 device, vvoo_t, oovv) instances of any (O, V) in one
    launch.  Returns (energies [n] device, [t2 per instance])."""
    lib = _lib.load()
    n = len(instances)
    arr = (_lib.Mp2Instance * n)()
    outs = []
    keep = []
    for k, (O, V, eps, vvoo_t, oovv) in enumerate(instances):
        t2 = torch.empty((O, O, V, V), dtype=CDT, device=device)
        outs.append(t2)
        keep.append(eps)
        arr[k] = _lib.Mp2Instance(O, V, eps.data_ptr(), vvoo_t.data_ptr(), oovv.data_ptr(), t2.data_ptr())
    E = torch.empty(n, dtype=CDT, device=device)
    nbytes = int(lib.apy_mp2_so_batched_workspace(n, arr))
    ws = workspace(nbytes, device)
    _lib.check(lib.apy_mp2_so_batched(n, arr, _vp(E), _vp(ws), nbytes, stream_of(device)), "apy_mp2_so_batched")
    return E, outs
