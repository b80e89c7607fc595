"""AO electron-repulsion integrals on the GPU for the package's basis sets.

Replaces psi4's MintsHelper.ao_eri (hamiltonian.py:33) and the two-electron
part of mo_tei_deriv1 (analytic_aats.py:94) on the device: the basis (and
its nuclear derivatives) become apy_gto tables and apy_eri_gto evaluates
(mn|ls) for one or several geometries in one launch; apy_eri_deriv_assemble
forms d(mn|ls)/dR for one atom and axis from the derivative-function
integrals.  Results are complex128 device tensors, the form every consumer
on the GPU (Fock build, AO -> MO transforms) reads.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from .gaussian import center_derivative
from .kernels import stream_of


def _entry(atom, terms):
    """One apy_gto entry from (lmn, coefs) terms sharing exps."""
    g = _lib.Gto()
    exps = terms[0][2]
    if len(exps) > _lib.APY_GTO_MAXPRIM or not 1 <= len(terms) <= 2:
        raise ValueError("function outside the integral kernel's limits")
    g.atom, g.nprim, g.nterm = int(atom), len(exps), len(terms)
    for k, (lmn, coefs, e) in enumerate(terms):
        if sum(lmn) > 2 or not np.array_equal(e, exps):
            raise ValueError("function outside the integral kernel's limits")
        for x in range(3):
            g.lmn[k][x] = int(lmn[x])
        for p in range(len(exps)):
            g.coefs[k][p] = float(coefs[p])
    for p in range(len(exps)):
        g.exps[p] = float(exps[p])
    return g


def _upload(entries, device):
    arr = (_lib.Gto * len(entries))(*entries)
    buf = np.frombuffer(bytes(arr), dtype=np.uint8).copy()
    return torch.from_numpy(buf).to(device)


def basis_table(functions, atom_of, device):
    """Device apy_gto table of the basis functions."""
    return _upload([_entry(a, [(f.lmn, f.coefs, f.exps)]) for f, a in zip(functions, atom_of)], device)


def derivative_table(functions, atom_of, axis, device):
    """Device apy_gto table of d f / dA_axis for every basis function f
    (centred at A; gaussian.center_derivative's terms)."""
    out = []
    for f, a in zip(functions, atom_of):
        out.append(_entry(a, [(t.lmn, t.coefs, t.exps) for t in center_derivative(f, axis)]))
    return _upload(out, device)


def _centers(coords, device):
    c = np.ascontiguousarray(np.asarray(coords, dtype=np.float64))
    return torch.from_numpy(c).to(device)


def eri(fa, fb, fc, fd, counts, centers, device):
    """(fa fb | fc fd) for the device tables fa..fd with `counts` entries,
    at every geometry of `centers` [ngeom, natom, 3] (bohr).  Returns a
    complex128 device tensor [ngeom, na, nb, nc, nd]."""
    lib = _lib.load()
    cen = _centers(centers, device)
    ngeom, natom = cen.shape[0], cen.shape[1]
    na, nb, nc, nd = counts
    out = torch.empty((ngeom, na, nb, nc, nd), dtype=torch.complex128, device=device)
    P = lambda t: ctypes.c_void_p(t.data_ptr())
    _lib.check(lib.apy_eri_gto(ngeom, natom, P(cen), P(fa), na, P(fb), nb, P(fc), nc, P(fd), nd, P(out),
                               stream_of(device)), "apy_eri_gto")
    return out


def ao_eri(basis_set, coords, device):
    """(mn|ls) of a BasisSet at one geometry (natom x 3, bohr) or at several
    ([ngeom, natom, 3]; the functions follow their atoms): complex128 device
    tensor [n,n,n,n] (or [ngeom, n,n,n,n])."""
    fns = basis_set.functions
    n = len(fns)
    tab = basis_table(fns, basis_set.atom_of, device)
    c = np.asarray(coords, dtype=np.float64)
    one = c.ndim == 2
    out = eri(tab, tab, tab, tab, (n, n, n, n), c[None] if one else c, device)
    return out[0] if one else out


class EriDerivatives:
    """d(mn|ls)/dR_{A,x} for every atom A and axis x of a BasisSet at one
    geometry: the three (dm/dx n|l s) tensors stay on the device and each
    derivative is assembled on request (apy_eri_deriv_assemble)."""

    def __init__(self, basis_set, coords, device):
        fns = basis_set.functions
        self.n = n = len(fns)
        self.device = device
        tab = basis_table(fns, basis_set.atom_of, device)
        cen = np.asarray(coords, dtype=np.float64)[None]
        self.D = [eri(derivative_table(fns, basis_set.atom_of, x, device), tab, tab, tab, (n, n, n, n), cen,
                      device)[0] for x in range(3)]
        self.atom_of = torch.from_numpy(np.asarray(basis_set.atom_of, dtype=np.int32)).to(device)

    def __call__(self, atom, axis):
        out = torch.empty_like(self.D[axis])
        _lib.check(_lib.load().apy_eri_deriv_assemble(self.n, ctypes.c_void_p(self.atom_of.data_ptr()), int(atom),
                                                      ctypes.c_void_p(self.D[axis].data_ptr()),
                                                      ctypes.c_void_p(out.data_ptr()), stream_of(self.device)),
                   "apy_eri_deriv_assemble")
        return out


def prime_device_eri(hamiltonians, device):
    """Fill the device ERI cache of several Hamiltonians with one
    apy_eri_gto launch per basis (same functions on the same atoms): the
    distinct geometries are computed together, and Hamiltonians at the same
    geometry (field displacements) share one tensor."""
    key = str(device)
    groups = {}
    for H in hamiltonians:
        if key in H._device_cache:
            continue
        bs = H.basis_set
        sig = (bs.name, tuple(int(z) for z in H.molecule.Z))
        groups.setdefault(sig, []).append(H)
    for Hs in groups.values():
        geoms, index = [], []
        for H in Hs:
            c = np.asarray(H.molecule.coords, dtype=np.float64)
            for g, other in enumerate(geoms):
                if np.array_equal(other, c):
                    index.append(g)
                    break
            else:
                index.append(len(geoms))
                geoms.append(c)
        out = ao_eri(Hs[0].basis_set, np.stack(geoms), device)
        for H, g in zip(Hs, index):
            H._device_cache[key] = out[g]
