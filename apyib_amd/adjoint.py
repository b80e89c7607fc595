"""MP2 energy with an adjoint (lambda) backward behind torch.autograd.

Orbital-invariant spin-orbital MP2 (the Fock part of ci_wfn.py:210-212
driven by <ab||ij>, energy of mp2_wfn.py:85; at diagonal F it is
solve_MP2_SO):
    forward : t from  <ab||ij> + P(ab) t F_vv^T - P(ij) F_oo^T t = 0,
              E = 1/4 sum <ij||ab> t
    backward: lambda from the transposed equations driven by <ij||ab>/4, then
              dE/dF (oo and vv blocks), dE/d<ij||ab> = t/4, dE/d<ab||ij> = lambda
No iteration history is kept: the backward pass needs only t and lambda.
All contractions run in libapyib_amd.so (apy_mp2_nc_step,
apy_mp2_nc_densities, apy_zdot_many).
"""
import ctypes

import numpy as np
import torch

from . import _lib
from .kernels import stream_of, workspace

CDT = torch.complex128


def _vp(t):
    return ctypes.c_void_p(t.data_ptr())


class MP2Adjoint:
    """Device state of one MP2 energy/gradient evaluation.

    fock [n,n], W = <ij||ab> [O,O,V,V], Vd = <ab||ij> at [i,j,a,b], eps [n]
    (denominators of the Jacobi preconditioner).  hermitian=True (the caller
    asserts F = F^H and Vd = conj(W)): the lambda equations are the
    conjugated amplitude equations scaled by 1/4, so gradients() takes
    lambda = conj(t)/4 instead of solving them."""

    def __init__(self, fock, W, Vd, eps, O, tol=1e-12, maxit=200, hermitian=False):
        self.lib = _lib.load()
        self.dev = fock.device
        self.O = int(O)
        self.V = fock.shape[0] - self.O
        self.F = fock.contiguous()
        self.W = W.contiguous()
        self.Vd = Vd.contiguous()
        self.eps = torch.as_tensor(np.asarray(eps, dtype=np.float64) if not isinstance(eps, torch.Tensor) else eps,
                                   dtype=torch.float64).to(self.dev).contiguous()
        self.tol, self.maxit = tol, maxit
        self.hermitian = bool(hermitian)
        self.ws_bytes = int(self.lib.apy_mp2_nc_workspace(self.O, self.V))
        self.rnorm = torch.zeros(1, dtype=CDT, device=self.dev)
        self.t = None
        self.iterations = {}

    def _solve(self, transpose, drv, scale):
        x = torch.zeros_like(drv)
        args = _lib.Mp2NcArgs(self.O, self.V, self.F.data_ptr(), self.eps.data_ptr(), int(transpose), float(scale),
                              drv.data_ptr(), x.data_ptr(), x.data_ptr(), self.rnorm.data_ptr())
        ws = workspace(self.ws_bytes, self.dev)
        it = 0
        for it in range(1, self.maxit + 1):
            _lib.check(self.lib.apy_mp2_nc_step(ctypes.byref(args), _vp(ws), self.ws_bytes, stream_of(self.dev)),
                       "apy_mp2_nc_step")
            if it > 1 and float(self.rnorm.real.cpu()[0]) ** 0.5 < self.tol:
                break
        self.iterations['lambda' if transpose else 't'] = it
        return x

    def energy(self):
        self.t = self._solve(False, self.Vd, 1.0)
        out = torch.zeros(1, dtype=CDT, device=self.dev)
        nb = int(self.lib.apy_reduce_workspace(self.t.numel(), 1))
        ws = workspace(nb, self.dev)
        arr = (ctypes.c_void_p * 1)(self.W.data_ptr())
        _lib.check(self.lib.apy_zdot_many(1, arr, _vp(self.t), self.t.numel(), 0, _vp(out), _vp(ws), nb,
                                          stream_of(self.dev)), "apy_zdot_many")
        return 0.25 * out[0]

    def gradients(self):
        """(dE/dF [n,n], dE/d<ij||ab>, dE/d<ab||ij>) at the current solution."""
        if self.t is None:
            self.energy()
        lam = 0.25 * self.t.conj() if self.hermitian else self._solve(True, self.W, 0.25)
        n = self.O + self.V
        G = torch.empty((n, n), dtype=CDT, device=self.dev)
        nb = int(self.lib.apy_mp2_nc_density_workspace(self.O, self.V))
        ws = workspace(nb, self.dev)
        _lib.check(self.lib.apy_mp2_nc_densities(self.O, self.V, _vp(self.t), _vp(lam), _vp(G), _vp(ws), nb,
                                                 stream_of(self.dev)), "apy_mp2_nc_densities")
        return G, 0.25 * self.t, lam


class MP2Energy(torch.autograd.Function):
    """E(F, <ij||ab>, <ab||ij>) with the lambda backward.  E is holomorphic in
    its inputs (no conjugation in the energy expression), so the
    vector-Jacobian product is grad_output * conj(dE/dx)."""

    @staticmethod
    def forward(ctx, fock, W, Vd, eps, O):
        solver = MP2Adjoint(fock, W, Vd, eps, O)
        E = solver.energy()
        ctx.solver = solver
        return E

    @staticmethod
    def backward(ctx, gE):
        G, gW, gV = ctx.solver.gradients()
        return gE * G.conj(), gE * gW.conj(), gE * gV.conj(), None, None


def mp2_energy(fock, W, Vd, eps, O):
    return MP2Energy.apply(fock, W, Vd, eps, O)
