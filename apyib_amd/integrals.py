"""AO integrals on the GPU for the package's basis sets.

Replaces psi4's MintsHelper.ao_eri (hamiltonian.py:33) and the two-electron
part of mo_tei_deriv1 (analytic_aats.py:94) on the device: the basis (and
its nuclear derivatives) become apy_gto tables and apy_eri_gto evaluates
(mn|ls) for one or several geometries in one launch; apy_eri_deriv_assemble
forms d(mn|ls)/dR for one atom and axis from the derivative-function
integrals.  Results are complex128 device tensors, the form every consumer
on the GPU (Fock build, AO -> MO transforms) reads.

The one-electron integrals of hamiltonian.py:29-35, 49, 55 (overlap,
kinetic, nuclear attraction, dipole, angular momentum) and the
mixed-geometry overlaps of utils.py:371-376 come from apy_oei_gto, again
for many geometries (or bra/ket geometry pairs) per launch.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from .gaussian import assemble_deriv1, center_derivative
from .kernels import stream_of


def _entry(atom, terms):
    """One apy_gto entry from (lmn, coefs) terms sharing exps."""
    g = _lib.Gto()
    exps = terms[0][2]
    if len(exps) > _lib.APY_GTO_MAXPRIM or not 1 <= len(terms) <= 2:
        raise ValueError("function outside the integral kernel's limits")
    g.atom, g.nprim, g.nterm = int(atom), len(exps), len(terms)
    for k, (lmn, coefs, e) in enumerate(terms):
        if sum(lmn) > _lib.APY_GTO_MAXL or not np.array_equal(e, exps):
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


def function_table(functions, device):
    """Device apy_gto table in which every function is its own centre
    (entry i at centre i), and the [n, 3] host array of those centres: the
    form for function lists that carry their centres (mixed overlaps)."""
    tab = _upload([_entry(i, [(f.lmn, f.coefs, f.exps)]) for i, f in enumerate(functions)], device)
    return tab, np.stack([np.asarray(f.center, dtype=np.float64) for f in functions])


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


def eri_to_basis(basis_set, G):
    """(mn|ls) over basis_set.functions (device [n_fn]^4) -> over the basis
    functions when the basis has a transform (spherical d shells, split
    contractions): the four-index transform with C = transform
    (apy_eri_mo)."""
    if basis_set.transform is None:
        return G
    from .transforms import eri_mo
    C = torch.from_numpy(basis_set.transform.astype(np.complex128)).to(G.device)
    return eri_mo(G.contiguous(), C)


OEI_ALL = (_lib.APY_OEI_OVERLAP | _lib.APY_OEI_KINETIC | _lib.APY_OEI_POTENTIAL | _lib.APY_OEI_DIPOLE
           | _lib.APY_OEI_ANGMOM)


def one_electron(fa, ca, fb, cb, which, device, nuclei=None, charges=None, origin=(0.0, 0.0, 0.0)):
    """apy_oei_gto over the device tables fa (bra) and fb (ket) with host
    centres ca [nset, nca, 3] and cb [nset, ncb, 3] (bohr); nuclei
    [nset, nnuc, 3] and charges [nnuc] for the nuclear attraction.  Returns a
    float64 device tensor [nset, ncomp, na, nb], the components `which`
    selects in the order S, T, V, Dx, Dy, Dz, Lx, Ly, Lz."""
    lib = _lib.load()
    ncomp = int(lib.apy_oei_components(int(which)))
    if ncomp == 0:
        raise ValueError(f"invalid one-electron component selection {which}")
    na, nb = fa.numel() // ctypes.sizeof(_lib.Gto), fb.numel() // ctypes.sizeof(_lib.Gto)
    ca, cb = np.asarray(ca, dtype=np.float64), np.asarray(cb, dtype=np.float64)
    if ca.ndim != 3 or cb.ndim != 3 or ca.shape[0] != cb.shape[0] or ca.shape[2] != 3 or cb.shape[2] != 3:
        raise ValueError("centres must be [nset, ncentre, 3] with the same nset")
    nset = ca.shape[0]
    dca, dcb = _centers(ca, device), _centers(cb, device)
    dnuc = dchg = None
    nnuc = 0
    if which & _lib.APY_OEI_POTENTIAL:
        nuc = np.asarray(nuclei, dtype=np.float64)
        if nuc.ndim != 3 or nuc.shape[0] != nset or nuc.shape[2] != 3 or len(charges) != nuc.shape[1]:
            raise ValueError("nuclei must be [nset, nnuc, 3] with one charge per nucleus")
        dnuc, dchg, nnuc = _centers(nuc, device), _centers(charges, device), nuc.shape[1]
    out = torch.empty((nset, ncomp, na, nb), dtype=torch.float64, device=device)
    P = lambda t: t.data_ptr() if t is not None else None
    args = _lib.OeiArgs(nset, P(fa), na, P(dca), ca.shape[1], P(fb), nb, P(dcb), cb.shape[1], P(dnuc), P(dchg), nnuc,
                        (ctypes.c_double * 3)(*[float(x) for x in origin]), int(which), 0, P(out))
    _lib.check(lib.apy_oei_gto(ctypes.byref(args), stream_of(device)), "apy_oei_gto")
    return out


def hamiltonian_integrals(basis_set, coords, charges, device, which=OEI_ALL):
    """One-electron integrals of a BasisSet at several geometries
    ([ngeom, natom, 3], bohr; functions follow their atoms, nuclei of the
    given charges at the same positions): float64 device tensor
    [ngeom, ncomp, n, n] (components as in one_electron)."""
    tab = basis_table(basis_set.functions, basis_set.atom_of, device)
    c = np.asarray(coords, dtype=np.float64)
    return one_electron(tab, c, tab, c, which, device, nuclei=c, charges=np.asarray(charges, dtype=np.float64))


def deriv1_one_electron(basis_set, coords, charges, device):
    """The one-electron part of gaussian.build_deriv1 (per atom A and axis x:
    d/dR_Ax of S, T, V, and half_S = <d m/dR_Ax | n>; the AO integrals behind
    mo_oei_deriv1 and mo_overlap_half_deriv1 of analytic_aats.py:84-96) from
    two apy_oei_gto launches: the three derivative tables stacked as one bra
    table of 3n entries against the basis, giving <d m/dx | n> and
    <d m/dx | T | n>, then <d m/dx | V_C | n> for every nucleus C alone (one
    set per nucleus at unit charge, scaled by Z_C); assembled on the host."""
    fns, atom_of = basis_set.functions, basis_set.atom_of
    n = len(fns)
    c = np.asarray(coords, dtype=np.float64)
    natom = c.shape[0]
    Z = np.asarray(charges, dtype=np.float64)
    dtab = torch.cat([derivative_table(fns, atom_of, x, device) for x in range(3)])
    tab = basis_table(fns, atom_of, device)
    ST = one_electron(dtab, c[None], tab, c[None], _lib.APY_OEI_OVERLAP | _lib.APY_OEI_KINETIC, device)
    crep = np.ascontiguousarray(np.broadcast_to(c, (natom, natom, 3)))
    VC = one_electron(dtab, crep, tab, crep, _lib.APY_OEI_POTENTIAL, device, nuclei=c[:, None, :],
                      charges=np.ones(1))
    ST, VC = ST[0].cpu().numpy(), VC[:, 0].cpu().numpy()
    dVC = VC.reshape(natom, 3, n, n) * Z[:, None, None, None]
    return assemble_deriv1(ST[0].reshape(3, n, n), ST[1].reshape(3, n, n), dVC, atom_of)


def _signature(functions):
    return tuple((tuple(int(x) for x in f.lmn), np.asarray(f.exps).tobytes(), np.asarray(f.coefs).tobytes())
                 for f in functions)


def mixed_overlaps(pairs, device):
    """<bra_i|ket_j> for a list of (bra functions, ket functions) pairs, each
    function at its own centre (utils.py:371-376 for every pair): pairs with
    the same function shapes share one apy_oei_gto launch.  Returns a list of
    float64 device tensors [na, nb] in the order of `pairs`."""
    groups = {}
    for q, (bra, ket) in enumerate(pairs):
        groups.setdefault((_signature(bra), _signature(ket)), []).append(q)
    out = [None] * len(pairs)
    for qs in groups.values():
        bra0, ket0 = pairs[qs[0]]
        fa, _ = function_table(bra0, device)
        fb, _ = function_table(ket0, device)
        ca = np.stack([[np.asarray(f.center, dtype=np.float64) for f in pairs[q][0]] for q in qs])
        cb = np.stack([[np.asarray(f.center, dtype=np.float64) for f in pairs[q][1]] for q in qs])
        S = one_electron(fa, ca, fb, cb, _lib.APY_OEI_OVERLAP, device)
        for k, q in enumerate(qs):
            out[q] = S[k, 0]
    return out


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
        if Hs[0].basis_set.transform is not None:
            out = [eri_to_basis(Hs[0].basis_set, out[g]) for g in range(len(geoms))]
        for H, g in zip(Hs, index):
            H._device_cache[key] = out[g]
