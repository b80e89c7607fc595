"""Spin-adapted closed-shell CISD / CID on the GPU (ci_wfn.py:420-574 and
:51-167): the reference's Jacobi + DIIS iteration over spatial amplitudes
[t1 (o x v) | t2 (o x o x v x v)], with the residual of apy_cisd_sp_step
(complex GEMMs on the FP64 matrix cores) and the energy of
apy_cisd_sp_energy.  The iteration driver (DIIS, convergence test,
printing) is the spin-orbital solver's, so the two forms share one
restatement of utils.py:104-140 and ci_wfn.py:361-378.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from .kernels import stream_of, workspace
from .so_ci import SOCI

CDT = torch.complex128
CID_SP_BLOCKS = ("vvvs", "vvva", "oooo", "w21", "w3", "vvoo", "l")


def _vp(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


class SPIntegrals:
    """Spatial Dirac integral blocks resident on one device, in the layouts
    documented for apy_sp_ints in include/apyib_amd.h."""

    def __init__(self, o, v, device="cuda", blocks=_lib.SP_BLOCKS):
        lib = _lib.load()
        self.o, self.v = int(o), int(v)
        self.device = torch.device(device)
        sizes = (ctypes.c_int64 * len(_lib.SP_BLOCKS))()
        _lib.check(lib.apy_sp_ints_sizes(self.o, self.v, sizes), "apy_sp_ints_sizes")
        self.blocks = {name: torch.empty(int(sizes[k]), dtype=CDT, device=self.device)
                       for k, name in enumerate(_lib.SP_BLOCKS) if name in blocks}
        self.struct = _lib.SpInts(self.o, self.v, *[self.blocks[b].data_ptr() if b in self.blocks else None
                                                    for b in _lib.SP_BLOCKS])

    def __getitem__(self, name):
        return self.blocks[name]

    @classmethod
    def from_mo(cls, eri_mo, o, singles=True):
        """From the chemists' MO tensor (pq|rs) over the active orbitals
        (utils.py:258-279; <pq|rs> = (pr|qs), ci_wfn.py:433)."""
        m = eri_mo.shape[0]
        self = cls(o, m - o, eri_mo.device, _lib.SP_BLOCKS if singles else CID_SP_BLOCKS)
        eri_mo = eri_mo.to(CDT).contiguous()
        _lib.check(_lib.load().apy_sp_ints_from_mo(m, _vp(eri_mo), ctypes.byref(self.struct),
                                                   stream_of(eri_mo.device)), "apy_sp_ints_from_mo")
        return self


class SpatialCI(SOCI):
    """One spin-adapted CISD (singles=True) or CID problem on one device.

    ints: SPIntegrals; fock: [n,n] MO Fock matrix (device); eps: [n] orbital
    energies of the denominators (ci_wfn.py:41-43)."""

    def __init__(self, ints, fock, eps, singles=True):
        self.lib = _lib.load()
        self.ints = ints
        self.O, self.V = ints.o, ints.v
        self.n = self.O + self.V
        self.device = ints.device
        self.singles = bool(singles)
        self.fov = False
        self.fock = fock.to(self.device, CDT).contiguous()
        self.eps = torch.as_tensor(np.asarray(eps, dtype=np.float64) if not isinstance(eps, torch.Tensor) else eps,
                                   dtype=torch.float64).to(self.device).contiguous()
        o, v = self.O, self.V
        self.OV = o * v
        self.L = o * v + o * o * v * v
        self.hoo = torch.empty(o ** 4, dtype=CDT, device=self.device)
        self.g = torch.empty(2 * o * v, dtype=CDT, device=self.device)
        _lib.check(self.lib.apy_cisd_sp_prepare(o, v, _vp(self.fock), ctypes.byref(ints.struct), _vp(self.hoo),
                                                _vp(self.g), stream_of(self.device)), "apy_cisd_sp_prepare")
        self.ws_bytes = int(self.lib.apy_cisd_sp_workspace(o, v, int(self.singles)))
        self.red_bytes = int(self.lib.apy_reduce_workspace(self.L, _lib.APY_MAXVEC))
        self.en_bytes = int(self.lib.apy_cisd_sp_energy_workspace())
        self.energy_dev = torch.zeros(1, dtype=CDT, device=self.device)
        self.zero_energy = torch.zeros(1, dtype=CDT, device=self.device)
        self.scal = torch.zeros(_lib.APY_MAXVEC + 4, dtype=CDT, device=self.device)
        self.ev_dominant = None

    def step(self, t, t_new, out, energy=None, **_):
        """r = sigma(t) - E t, t_new = t + r / D (apy_cisd_sp_step)."""
        args = _lib.CisdSpArgs(self.O, self.V, self.fock.data_ptr(), self.eps.data_ptr(),
                               ctypes.pointer(self.ints.struct), self.hoo.data_ptr(), self.g.data_ptr(),
                               int(self.singles), (energy if energy is not None else self.energy_dev).data_ptr(),
                               t.data_ptr(), t_new.data_ptr(), out.data_ptr())
        ws = workspace(self.ws_bytes, self.device)
        _lib.check(self.lib.apy_cisd_sp_step(ctypes.byref(args), _vp(ws), self.ws_bytes, stream_of(self.device)),
                   "apy_cisd_sp_step")

    def energy(self, t, out=None):
        out = self.energy_dev if out is None else out
        ws = workspace(self.en_bytes, self.device)
        _lib.check(self.lib.apy_cisd_sp_energy(self.O, self.V, int(self.singles), _vp(self.g), _vp(self.ints["l"]),
                                               _vp(t), _vp(out), _vp(ws), self.en_bytes, stream_of(self.device)),
                   "apy_cisd_sp_energy")
        return out

    def sigma(self, *a, **k):
        raise NotImplementedError("the spin-adapted solver runs the reference's Jacobi iteration only")

    def tune(self, *a, **k):
        raise NotImplementedError("the spin-adapted solver uses the cost-model GEMM plans")


class ResponseCI(SpatialCI):
    """First-order CISD (singles=True) or CID amplitudes dt for one
    perturbation (dF, d<pq|rs>) of a converged spin-adapted solution t with
    correlation energy E (analytic_aats.py:733-890, :980-1143; CID
    :1540-1660): the reference's Jacobi + DIIS iteration on

        r = R_dH(1, t) - s t + [sigma_H(0, dt) - E dt],
        s = E_dH(t) + E_H(dt)           (projected energy derivative),

    run as apy_cisd_sp_step on dt with the response term resp_b =
    R_dH(1, t) - R_H(0) (the perturbed residual of t with s = 0, less this
    Hamiltonian's driver), resp_t = t, resp_energy = E_H(dt) of the
    iteration and resp_shift = E_dH(t).  The iteration's "energy" is E_H(dt);
    solve() returns it with E_dH(t) added.  The guess is one such step from
    dt = 0 (the reference seeds its guess with the density gradient instead;
    the converged amplitudes are the same).

    ints, fock, eps: the unperturbed problem (as for SpatialCI); t: its
    converged amplitude vector on the device; dints: SPIntegrals of the
    perturbed <pq|rs>; dfock: [n, n] derivative Fock matrix (device)."""

    def __init__(self, ints, fock, eps, singles, t, E, dints, dfock):
        super().__init__(ints, fock, eps, singles)
        self.t_ref = t.contiguous()
        self.E_dev = torch.full((1,), complex(E), dtype=CDT, device=self.device)
        pert = SpatialCI(dints, dfock, self.eps, singles)
        zero = self.new_vector()
        scratch = self.new_vector()
        r_pert = self.new_vector()
        drv = self.new_vector()
        pert.step(self.t_ref, scratch, r_pert, energy=self.zero_energy)
        SpatialCI.step(self, zero, scratch, drv, energy=self.zero_energy)
        self.b = self.new_vector()
        self.combine([r_pert, drv], [1.0, -1.0], self.b)
        self.shift = complex(pert.energy(self.t_ref, torch.zeros(1, dtype=CDT, device=self.device)).cpu()[0])

    def step(self, t, t_new, out, energy=None, **_):
        args = _lib.CisdSpArgs(self.O, self.V, self.fock.data_ptr(), self.eps.data_ptr(),
                               ctypes.pointer(self.ints.struct), self.hoo.data_ptr(), self.g.data_ptr(),
                               int(self.singles), self.E_dev.data_ptr(), t.data_ptr(), t_new.data_ptr(),
                               out.data_ptr(), self.b.data_ptr(), self.t_ref.data_ptr(),
                               (energy if energy is not None else self.energy_dev).data_ptr(),
                               self.shift.real, self.shift.imag)
        ws = workspace(self.ws_bytes, self.device)
        _lib.check(self.lib.apy_cisd_sp_step(ctypes.byref(args), _vp(ws), self.ws_bytes, stream_of(self.device)),
                   "apy_cisd_sp_step")

    def initial_guess(self, t):
        scratch = self.new_vector()
        self.step(self.new_vector(), t, scratch, energy=self.zero_energy)
        return t

    def solve(self, params, print_level=0, max_diis=7, e_offset=0.0):
        """Returns (dE, dt1, dt2, iterations): dE the converged projected
        energy derivative, dt1/dt2 views of the device amplitude vector."""
        E, dt1, dt2, it = super().solve(params, print_level, max_diis, e_offset)
        return E + self.shift, dt1, dt2, it
