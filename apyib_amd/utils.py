"""apyib's utils API (utils.py:104-446) on the GPU.

Same names, arguments and return types as the reference (NumPy arrays in
and out); the contractions run in libapyib_amd.so.  `*_device` variants
keep their results on the device for the solvers.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from . import transforms as tr
from .integrals import mixed_overlaps
from .kernels import stream_of, workspace

CDT = torch.complex128


def device():
    """The current GPU; raises if none is available (no CPU fallback)."""
    if not torch.cuda.is_available():
        raise RuntimeError("apyib_amd needs a GPU (MI355X / gfx950): torch.cuda.is_available() is False")
    _lib.load()
    return torch.device("cuda", torch.cuda.current_device())


def get_slices(parameters, wfn):
    """utils.py:184-213."""
    nfzc = wfn.H.basis_set.n_frozen_core() if parameters.get('freeze_core', False) else 0
    nbf = wfn.nbf
    no = wfn.ndocc
    C_list = [slice(0, nfzc), slice(nfzc, no), slice(no, nbf), slice(nfzc, nbf)]
    m = parameters['method']
    if m in ('RHF', 'MP2', 'CID', 'CISD'):
        I_list = [slice(0, nfzc), slice(0, no - nfzc), slice(no - nfzc, nbf - nfzc), slice(0, nbf - nfzc)]
    elif m in ('MP2_SO', 'CID_SO', 'CISD_SO'):
        I_list = [slice(0, 2 * nfzc), slice(0, 2 * no - 2 * nfzc), slice(2 * no - 2 * nfzc, 2 * nbf - 2 * nfzc),
                  slice(0, 2 * nbf - 2 * nfzc)]
    else:
        I_list = [slice(0, nfzc), slice(0, no - nfzc), slice(no - nfzc, nbf - nfzc), slice(0, nbf - nfzc)]
    return C_list, I_list


def compute_F_MO_device(parameters, wfn, C_list):
    """utils.py:217-254 with the AO Fock builds and C^H F C on the GPU.
    Returns (F_MO device [nact,nact], E_fc complex)."""
    dev = device()
    f, o, v, t = C_list
    C = tr.to_dev(wfn.C, dev)
    eri = wfn.H.device_eri(dev)
    h = tr.to_dev(wfn.H.T + wfn.H.V, dev)
    E_fc = 0
    if parameters.get('freeze_core', False) and (f.stop - f.start) > 0:
        D_fc = tr.density(C[:, f])
        h_fc = tr.fock(h, eri, D_fc, 1.0)                                         # :241
        E_fc = complex(torch.sum(D_fc.t() * (h + h_fc)).cpu())                    # :242
        h = h_fc
    D = tr.density(C[:, o])                                                       # :246
    F_AO = tr.fock(h, eri, D, 1.0)                                                # :249
    Ct = C[:, t].contiguous()
    return tr.sandwich(Ct, F_AO, Ct), E_fc                                        # :252


def compute_F_MO(parameters, wfn, C_list):
    F, E_fc = compute_F_MO_device(parameters, wfn, C_list)
    return F.cpu().numpy(), E_fc


def compute_ERI_MO_device(parameters, wfn, C_list):
    """utils.py:258-279: four quarter transforms on the matrix cores."""
    dev = device()
    C = tr.to_dev(wfn.C, dev)[:, C_list[3]].contiguous()
    return tr.eri_mo(wfn.H.device_eri(dev), C)


def compute_ERI_MO(parameters, wfn, C_list):
    return compute_ERI_MO_device(parameters, wfn, C_list).cpu().numpy()


def compute_F_SO(wfn, F_MO):
    """utils.py:283-313."""
    dt = np.asarray(F_MO).dtype
    out = tr.so_fock(tr.to_dev(F_MO, device())).cpu().numpy()
    return out if np.iscomplexobj(np.zeros(1, dtype=dt)) else out.real.copy()


def compute_ERI_SO(wfn, ERI_MO):
    """utils.py:317-365 (the reference's four-index Python loop)."""
    dt = np.asarray(ERI_MO).dtype
    out = tr.eri_so(tr.to_dev(ERI_MO, device())).cpu().numpy()
    return out if np.iscomplexobj(np.zeros(1, dtype=dt)) else out.real.copy()


def solve_general_DIIS(parameters, res_vec, t_vec, e_iter, t_iter, iteration, min_DIIS=1, max_DIIS=7):
    """utils.py:104-140 with the error-matrix dots and the extrapolation on
    the GPU.  e_iter / t_iter are (length, nvec) arrays as in the reference."""
    if len(e_iter[0]) > max_DIIS:
        while len(e_iter[0]) > max_DIIS:
            e_iter = np.delete(e_iter, 0, axis=1)
            t_iter = np.delete(t_iter, 0, axis=1)
    if iteration != 1:
        e_iter = np.hstack((e_iter, np.atleast_2d(res_vec).T))
        t_iter = np.hstack((t_iter, np.atleast_2d(t_vec).T))
    dev = device()
    lib = _lib.load()
    nv = e_iter.shape[1]
    L = e_iter.shape[0]
    E = torch.from_numpy(np.ascontiguousarray(e_iter.T, dtype=np.complex128)).to(dev)
    T = torch.from_numpy(np.ascontiguousarray(t_iter.T, dtype=np.complex128)).to(dev)
    Bd = torch.empty((nv, nv), dtype=CDT, device=dev)
    nbytes = int(lib.apy_reduce_workspace(L, nv))
    ws = workspace(nbytes, dev)
    ptrs = (ctypes.c_void_p * nv)(*[E[k].data_ptr() for k in range(nv)])
    for n_ in range(nv):
        # column n of B[m, n] = sum_o conj(e[o, m]) e[o, n]  (utils.py:121), stored in row n of Bd
        _lib.check(lib.apy_zdot_many(nv, ptrs, ctypes.c_void_p(E[n_].data_ptr()), L, 1,
                                     ctypes.c_void_p(Bd[n_].data_ptr()), ctypes.c_void_p(ws.data_ptr()), nbytes,
                                     stream_of(dev)), "apy_zdot_many")
    Bt = Bd.cpu().numpy().T   # row n_ of Bd holds column n_ of B
    real = not (np.iscomplexobj(e_iter) or np.iscomplexobj(t_iter))
    B = np.zeros((nv + 1, nv + 1), dtype=np.float64 if real else np.complex128)
    B[:nv, :nv] = Bt.real if real else Bt
    B[-1, :] = -1
    B[:, -1] = -1
    B[-1, -1] = 0
    A = np.zeros(nv + 1)
    A[-1] = -1
    C_diis = np.linalg.solve(B, A)                                                # :135
    out = torch.empty(L, dtype=CDT, device=dev)
    coef = np.ascontiguousarray(np.asarray(C_diis[:-1], dtype=np.complex128)).view(np.float64)
    tp = (ctypes.c_void_p * nv)(*[T[k].data_ptr() for k in range(nv)])
    _lib.check(lib.apy_zcombine(nv, tp, coef.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), L,
                                ctypes.c_void_p(out.data_ptr()), stream_of(dev)), "apy_zcombine")
    t_new = out.cpu().numpy()                                                     # :138
    if real:
        t_new = t_new.real.copy()
    return t_new, e_iter, t_iter


def ao_overlap(bra_basis, ket_basis):
    """AO overlap <bra|ket> (mixed geometries allowed; utils.py:371-376),
    computed on the device (apy_oei_gto): float64 device tensor (a host
    array for bases with a function transform, X = T_bra^T S T_ket)."""
    S = mixed_overlaps([(bra_basis.functions, ket_basis.functions)], device())[0]
    if getattr(bra_basis, 'transform', None) is None and getattr(ket_basis, 'transform', None) is None:
        return S
    return mixed_to_basis(bra_basis, ket_basis, S.cpu().numpy())


def mixed_to_basis(bra_basis, ket_basis, S):
    """<bra fn|ket fn> [n_fn, n_fn] (host) -> over the two bases' functions."""
    Tb, Tk = getattr(bra_basis, 'transform', None), getattr(ket_basis, 'transform', None)
    Tb = np.eye(S.shape[0]) if Tb is None else Tb
    Tk = np.eye(S.shape[1]) if Tk is None else Tk
    return Tb.T @ S @ Tk


def compute_mo_overlap(ndocc, nbf, bra_basis, bra_wfn, ket_basis, ket_wfn):
    """utils.py:370-388: C_bra^H S_AO(bra, ket) C_ket."""
    dev = device()
    S = tr.to_dev(ao_overlap(bra_basis, ket_basis), dev)
    return tr.mo_overlap(tr.to_dev(bra_wfn, dev), S, tr.to_dev(ket_wfn, dev)).cpu().numpy()


def compute_so_overlap(nbf, mo_overlap):
    """utils.py:393-422."""
    return tr.so_fock(tr.to_dev(mo_overlap, device())).cpu().numpy()


def compute_phase(ndocc, nbf, unperturbed_basis, unperturbed_wfn, ket_basis, ket_wfn):
    """utils.py:427-446."""
    dev = device()
    S = tr.to_dev(ao_overlap(unperturbed_basis, ket_basis), dev)
    C0 = tr.to_dev(unperturbed_wfn, dev)
    C = tr.to_dev(ket_wfn, dev)
    S_mo = tr.mo_overlap(C0, S, C)
    return tr.phase_align(S_mo, C).cpu().numpy()
