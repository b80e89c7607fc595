"""Complex restricted Hartree-Fock with apyib's interface (hf_wfn.py:9-170).

The upstream producer of C and eps for the correlated path.  The O(nbf^4)
Fock build runs on the GPU (apy_fock_build_batch); the nbf x nbf algebra
(orthogonalisation, eigh, DIIS on Fock matrices) is the reference's small
host-side linear algebra.
"""
import numpy as np
import scipy.linalg as la
import torch

from . import transforms as tr
from .utils import device


def _fock_diis(res_vec, F_vec, e_iter, F_iter, iteration, max_DIIS=7):
    """solve_general_DIIS (utils.py:104-140, as hf_wfn.py:91 calls it) on the
    host: the SCF's vectors hold nbf^2 elements, far too few for a device
    round trip.  Keeps at most max_DIIS + 1 (error, Fock) pairs (the oldest
    dropped before appending), solves the bordered system
    [[E^H E, -1], [-1, 0]] c = [0, ..., 0, -1] and returns sum_j c_j F_j."""
    while e_iter.shape[1] > max_DIIS:
        e_iter, F_iter = e_iter[:, 1:], F_iter[:, 1:]
    if iteration != 1:
        e_iter = np.hstack((e_iter, np.atleast_2d(res_vec).T))
        F_iter = np.hstack((F_iter, np.atleast_2d(F_vec).T))
    nv = e_iter.shape[1]
    B = -np.ones((nv + 1, nv + 1), dtype=np.result_type(e_iter.dtype, np.float64))
    B[:nv, :nv] = np.conjugate(e_iter).T @ e_iter
    B[-1, -1] = 0
    A = np.zeros(nv + 1)
    A[-1] = -1
    c = np.linalg.solve(B, A)
    return F_iter @ c[:-1], e_iter, F_iter


class hf_wfn(object):
    def __init__(self, H, charge=0):
        self.H = H
        self.nelec = self.nelectron(charge)
        self.ndocc = self.nelec // 2
        self.nbf = H.basis_set.nbf()
        self.C = np.zeros((self.nbf, self.nbf))
        self.eps = np.zeros((self.nbf))
        self.E_SCF = 0

    def nelectron(self, charge):
        nelec = -charge
        for atom in range(self.H.molecule.natom()):
            nelec += self.H.molecule.true_atomic_number(atom)
        return nelec

    def solve_SCF(self, parameters, print_level=0):
        """hf_wfn.py:36-170.  Returns (E_SCF, C)."""
        return solve_SCF_batch([self], parameters, print_level)[0]


class _SCF:
    """One instance of the reference's SCF iteration (hf_wfn.py:36-170): the
    core guess, then per iteration (given the Fock matrix of the current
    density, hf_wfn.py:81) the DIIS extrapolation, eigh in the orthogonal
    basis, density, energy and convergence test, on the host."""

    def __init__(self, wfn, parameters, print_level):
        self.wfn, self.parameters, self.print_level = wfn, parameters, print_level
        H = wfn.H
        self.H_core = H.T + H.V
        self.X = np.linalg.inv(la.sqrtm(H.S))
        e, C_p = np.linalg.eigh(self.X @ self.H_core @ self.X)
        C = self.X @ C_p
        self.D = 2 * C[:, :wfn.ndocc] @ np.conjugate(C[:, :wfn.ndocc]).T
        self.E = np.einsum('mn,mn->', self.D, self.H_core)
        self.i = 0
        self.F_iter = self.e_iter = None
        self.done = parameters['max_iterations'] < 1
        if print_level > 0:
            print("\n Iter      E_elec(real)       E_elec(imaginary)        E(tot)           Delta_E(real)       "
                  "Delta_E(imaginary)      RMS_D(real)      RMS_D(imaginary)")
            print(" %02d %20.12f %20.12f %20.12f" % (0, self.E.real, self.E.imag, self.E.real + H.E_nuc))

    def update(self, F):
        wfn, H, X, parameters = self.wfn, self.wfn.H, self.X, self.parameters
        self.i += 1
        i = self.i
        E_old, D_old = self.E, self.D
        if not np.iscomplexobj(self.H_core) and not np.iscomplexobj(self.D):
            F = F.real
        if parameters['DIIS']:
            SDF = H.S @ self.D @ F
            res_vec = np.reshape(X @ (SDF - np.conjugate(SDF.T)) @ X, (-1))
            F_vec = np.reshape(F, (-1))
            if i == 1:
                self.F_iter = np.atleast_2d(F_vec).T
                self.e_iter = np.atleast_2d(res_vec).T
            F_vec, self.e_iter, self.F_iter = _fock_diis(res_vec, F_vec, self.e_iter, self.F_iter, i)
            F = np.reshape(F_vec, (wfn.nbf, wfn.nbf))
        wfn.eps, C_p = np.linalg.eigh(X @ F @ X)
        C = wfn.C = X @ C_p
        self.D = 2 * C[:, :wfn.ndocc] @ np.conjugate(C[:, :wfn.ndocc]).T
        self.E = 0.5 * np.einsum('nm,mn->', self.D, self.H_core + F)
        delta_E = self.E - E_old
        rms_D = np.sqrt(np.sum((D_old - self.D) ** 2))
        if self.print_level > 0:
            print(" %02d %20.12f %20.12f %20.12f %20.12f %20.12f %20.12f %20.12f" % (
                i, self.E.real, self.E.imag, (self.E + H.E_nuc).real, delta_E.real, delta_E.imag,
                rms_D.real, rms_D.imag))
        if i > 1 and abs(delta_E) < parameters['e_convergence'] and rms_D < parameters['d_convergence']:
            self.done = True
        elif i == parameters['max_iterations']:
            self.done = True
            if abs(delta_E) > parameters['e_convergence'] or rms_D > parameters['d_convergence']:
                if self.print_level > 0:
                    print("Not converged.")


def solve_SCF_batch(wfns, parameters, print_level=0):
    """hf_wfn.solve_SCF for several wavefunctions with the same number of
    basis functions (the displaced geometries and fields of a
    finite-difference set), iterated together: each iteration builds the
    Fock matrices of all of them in one apy_fock_build_batch call
    (hf_wfn.py:81); DIIS, eigh and the convergence test stay per instance,
    so every result equals its own solve_SCF.  Returns [(E_SCF, C)]."""
    dev = device()
    states = [_SCF(w, parameters, print_level) for w in wfns]
    eris = [w.H.device_eri(dev) for w in wfns]
    eri = eris[0] if all(e is eris[0] for e in eris) else torch.stack(eris)
    hcore = torch.stack([tr.to_dev(st.H_core, dev) for st in states])
    while not all(st.done for st in states):
        D = torch.from_numpy(np.stack([np.asarray(st.D, dtype=np.complex128) for st in states])).to(dev)
        F = tr.fock_batch(hcore, eri, D, 0.5).cpu().numpy()
        for q, st in enumerate(states):
            if not st.done:
                st.update(F[q])
    for w, st in zip(wfns, states):
        w.E_SCF = st.E
    return [(w.E_SCF, w.C) for w in wfns]
