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
from .kernels import ensure_default_plans, stream_of, workspace
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
    energies of the denominators (ci_wfn.py:41-43).  gauss3m: the ladder and
    ring products in the three-product complex form (3/4 of their
    matrix-core work; for integrals and amplitudes whose imaginary parts are
    not small against their real parts, or one of the two purely real or
    purely imaginary, where the form is as accurate as four products)."""

    def __init__(self, ints, fock, eps, singles=True, gauss3m=False):
        self.lib = _lib.load()
        self.ints = ints
        self.O, self.V = ints.o, ints.v
        self.n = self.O + self.V
        self.device = ints.device
        self.singles = bool(singles)
        self.fov = False
        self.gauss3m = bool(gauss3m)
        ensure_default_plans(self.device)
        self.fock = fock.to(self.device, CDT).contiguous()
        self.eps = torch.as_tensor(np.array(eps, dtype=np.float64) if not isinstance(eps, torch.Tensor) else eps,
                                   dtype=torch.float64).to(self.device).contiguous()
        o, v = self.O, self.V
        self.OV = o * v
        self.L = o * v + o * o * v * v
        self.hoo = torch.empty(o ** 4, dtype=CDT, device=self.device)
        self.g = torch.empty(2 * o * v, dtype=CDT, device=self.device)
        _lib.check(self.lib.apy_cisd_sp_prepare(o, v, _vp(self.fock), ctypes.byref(ints.struct), _vp(self.hoo),
                                                _vp(self.g), stream_of(self.device)), "apy_cisd_sp_prepare")
        # the ring operator with F_vv folded in: one GEMM fewer per step
        self.w21_f = torch.empty_like(ints["w21"])
        _lib.check(self.lib.apy_cisd_sp_fold_fock(o, v, 1, _vp(self.fock), _vp(ints["w21"]), _vp(self.w21_f),
                                                  stream_of(self.device)), "apy_cisd_sp_fold_fock")
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
        args.w21_f = self.w21_f.data_ptr()
        args.gauss3m = int(self.gauss3m)
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

    def tune(self, nb=1, reps=3):
        """Time the candidate GEMM plans of one step over nb vectors
        (apy_cisd_sp_tune; nb > 1 is the shared-integral batch of
        ResponseSet) and cache the fastest per shape, process-wide."""
        t = torch.zeros((nb, self.L), dtype=CDT, device=self.device)
        e = torch.zeros(nb, dtype=CDT, device=self.device)
        args = _lib.CisdSpArgs(self.O, self.V, self.fock.data_ptr(), self.eps.data_ptr(),
                               ctypes.pointer(self.ints.struct), self.hoo.data_ptr(), self.g.data_ptr(),
                               int(self.singles), e.data_ptr(), t.data_ptr(), t.data_ptr(), t.data_ptr())
        args.w21_f = self.w21_f.data_ptr()
        args.gauss3m = int(self.gauss3m)
        nbytes = int(self.lib.apy_cisd_sp_batch_workspace(self.O, self.V, int(self.singles), nb))
        ws = workspace(nbytes, self.device)
        _lib.check(self.lib.apy_cisd_sp_tune(ctypes.byref(args), nb, 0, _vp(ws), nbytes, int(reps),
                                             stream_of(self.device)), "apy_cisd_sp_tune")


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

    def __init__(self, ints, fock, eps, singles, t, E, dints, dfock, gauss3m=False):
        super().__init__(ints, fock, eps, singles, gauss3m=gauss3m)
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
        args.w21_f = self.w21_f.data_ptr()
        args.gauss3m = int(self.gauss3m)
        ws = workspace(self.ws_bytes, self.device)
        _lib.check(self.lib.apy_cisd_sp_step(ctypes.byref(args), _vp(ws), self.ws_bytes, stream_of(self.device)),
                   "apy_cisd_sp_step")

    def initial_guess(self, t):
        scratch = self.new_vector()
        self.step(self.new_vector(), t, scratch, energy=self.zero_energy)
        return t

    def solve(self, params, print_level=0, max_diis=7, e_offset=0.0, **control):
        """Returns (dE, dt1, dt2, iterations): dE the converged projected
        energy derivative, dt1/dt2 views of the device amplitude vector."""
        E, dt1, dt2, it = super().solve(params, print_level, max_diis, e_offset, **control)
        return E + self.shift, dt1, dt2, it


class ResponseSet:
    """First-order CISD / CID amplitudes of several perturbations of one
    converged spin-adapted solution, iterated together: each iteration is
    one apy_cisd_sp_step_batch over all perturbations (the ladder and ring
    GEMMs stacked along M), then per perturbation the reference's DIIS
    extrapolation (utils.py:104-140) and convergence test
    (analytic_aats.py:869-882); a perturbation whose test has passed keeps
    its amplitudes from then on, as the reference's loop over perturbations
    would leave them.  Each perturbation's equations and guess are those of
    ResponseCI.

    ints, fock, eps, singles, t, E: the unperturbed problem and solution;
    add(dints, dfock) registers one perturbation (its residual of t is
    formed at once, so dints can be released afterwards).  gauss3m: as for
    SpatialCI, in the batched steps (each perturbation's right-hand side is
    formed with four products)."""

    def __init__(self, ints, fock, eps, singles, t, E, gauss3m=False):
        self.base = SpatialCI(ints, fock, eps, singles)
        self.gauss3m = bool(gauss3m)
        self.lib = self.base.lib
        self.singles = bool(singles)
        self.device = self.base.device
        self.O, self.V, self.L, self.OV = self.base.O, self.base.V, self.base.L, self.base.OV
        self.t_ref = t.contiguous()
        self.E = complex(E)
        self.zero = torch.zeros(1, dtype=CDT, device=self.device)
        scratch = self.base.new_vector()
        self.drv = self.base.new_vector()
        SpatialCI.step(self.base, self.base.new_vector(), scratch, self.drv, energy=self.zero)
        self.b, self.shift = [], []

    def add(self, dints, dfock):
        pert = SpatialCI(dints, dfock, self.base.eps, self.singles)
        scratch, r = self.base.new_vector(), self.base.new_vector()
        pert.step(self.t_ref, scratch, r, energy=self.zero)
        b = self.base.new_vector()
        self.base.combine([r, self.drv], [1.0, -1.0], b)
        self.b.append(b)
        self.shift.append(complex(pert.energy(self.t_ref, torch.zeros(1, dtype=CDT, device=self.device)).cpu()[0]))

    def _step(self, x, x_new, out, s):
        nb = x.shape[0]
        base = self.base
        args = _lib.CisdSpArgs(self.O, self.V, base.fock.data_ptr(), base.eps.data_ptr(),
                               ctypes.pointer(base.ints.struct), base.hoo.data_ptr(), base.g.data_ptr(),
                               int(self.singles), self.E_arr.data_ptr(), x.data_ptr(), x_new.data_ptr(),
                               out.data_ptr(), self.B.data_ptr(), self.t_ref.data_ptr(), s.data_ptr(), 0.0, 0.0)
        args.w21_f = base.w21_f.data_ptr()
        args.gauss3m = int(self.gauss3m)
        nbytes = int(self.lib.apy_cisd_sp_batch_workspace(self.O, self.V, int(self.singles), nb))
        ws = workspace(nbytes, self.device)
        _lib.check(self.lib.apy_cisd_sp_step_batch(ctypes.byref(args), nb, _vp(self.shift_dev), _vp(ws), nbytes,
                                                   stream_of(self.device)), "apy_cisd_sp_step_batch")

    def _energies(self, x, out):
        nb = x.shape[0]
        nbytes = nb * int(self.lib.apy_cisd_sp_energy_workspace())
        ws = workspace(nbytes, self.device)
        _lib.check(self.lib.apy_cisd_sp_energy_batch(self.O, self.V, int(self.singles), nb, _vp(self.base.g),
                                                     _vp(self.base.ints["l"]), _vp(x), _vp(out), _vp(ws), nbytes,
                                                     stream_of(self.device)), "apy_cisd_sp_energy_batch")

    def solve(self, params, print_level=0, max_diis=7):
        """Returns (dE [nb] complex host array, dt1 [nb,o,v] or None,
        dt2 [nb,o,o,v,v] device views, iterations per perturbation)."""
        nb, dev = len(self.b), self.device
        self.B = torch.stack(self.b)
        self.b = []
        self.shift_dev = torch.tensor(self.shift, dtype=CDT, device=dev)
        self.E_arr = torch.full((nb,), self.E, dtype=CDT, device=dev)
        E, X, iters = _solve_set(self._step, self._energies, nb, self.L, self.OV, self.singles, dev, params,
                                 print_level, max_diis, shift=self.shift)
        dE = E + np.asarray(self.shift)
        O, V = self.O, self.V
        dt1 = X[:, :self.OV].view(nb, O, V) if self.singles else None
        return dE, dt1, X[:, self.OV:].view(nb, O, O, V, V), iters


def _solve_set(step, energies, nb, L, OV, singles, dev, params, print_level=0, max_diis=7, shift=None):
    """The reference's Jacobi + DIIS iteration (ci_wfn.py:302-378,
    utils.py:104-140) for nb problems iterated together.  step(x, x_new,
    out, s): one step of all problems (vectors [nb, L]) with the
    per-problem energies s ([nb] device); energies(x, out): their energies.
    The guess is one step from zero amplitudes with zero energies.  Every
    problem's DIIS history holds the same iterations (they start together
    and each keeps the newest max_diis+1), so one segmented call forms all
    new rows of the DIIS matrices, one the extrapolated amplitudes and one
    the convergence sums; each problem's sums are bit-identical to its own
    solve's, and a problem whose test has passed keeps its amplitudes from
    then on, as one solve per problem would leave them.  Returns (E [nb]
    host, X [nb, L] device, iterations per problem)."""
    from .so_ci import _converged
    lib = _lib.load()
    st = stream_of(dev)
    diis = bool(params['DIIS'])
    nvec = max_diis + 1
    X = torch.zeros((nb, L), dtype=CDT, device=dev)
    Xn = torch.zeros((nb, L), dtype=CDT, device=dev)
    s = torch.zeros(nb, dtype=CDT, device=dev)
    s_new = torch.zeros(nb, dtype=CDT, device=dev)
    R = torch.zeros((nb, L), dtype=CDT, device=dev)
    step(Xn, X, R, s)                              # guess: one step from zero amplitudes (s = 0)
    Xn.copy_(X)
    energies(X, s)
    E = s.cpu().numpy()
    if diis:
        Th = torch.zeros((nvec, nb, L), dtype=CDT, device=dev)
        Eh = torch.zeros((nvec, nb, L), dtype=CDT, device=dev)
    hist = []                                      # (iteration, slot), shared by the problems
    Bs = np.zeros((nb, nvec, nvec), dtype=np.complex128)
    scal = torch.zeros((nb, _lib.APY_MAXVEC), dtype=CDT, device=dev)
    dsq = torch.zeros((nb, 2), dtype=CDT, device=dev)
    active = np.ones(nb, dtype=bool)
    iters = [0] * nb
    red = int(lib.apy_reduce_seg_workspace(2 * nb, _lib.APY_MAXVEC))
    shift = np.zeros(nb) if shift is None else np.asarray(shift)
    it = 0
    while active.any() and it < params['max_iterations']:
        it += 1
        if diis:
            slot = (it - 1) % nvec
            step(X, Th[slot], Eh[slot], s)
            while len(hist) > max_diis:
                hist.pop(0)
            hist.append((it, slot))
            k = len(hist)
            arr = (ctypes.c_void_p * k)(*[Eh[h].data_ptr() for _, h in hist])
            ws = workspace(red, dev)
            _lib.check(lib.apy_zdot_many_seg(k, arr, _vp(Eh[slot]), nb, L, 1, _vp(scal), _vp(ws), red, st),
                       "apy_zdot_many_seg")
            rows = scal.view(-1)[:nb * k].cpu().numpy().reshape(nb, k)      # packed [nb][k]
            slots = [h for _, h in hist]
            Bs[:, slots, slot] = rows                                         # E^H E by slot (utils.py:121)
            Bs[:, slot, slots] = np.conj(rows)
            act = np.flatnonzero(active)
            Bm = np.full((len(act), k + 1, k + 1), -1.0, dtype=np.complex128)
            Bm[:, :k, :k] = Bs[act][:, slots][:, :, slots]
            Bm[:, k, k] = 0
            rhs = np.zeros((len(act), k + 1, 1), dtype=np.complex128)
            rhs[:, k] = -1
            coef = np.zeros((nb, k), dtype=np.complex128)
            coef[:, -1] = 1.0                      # finished problems: unused
            coef[act] = np.linalg.solve(Bm, rhs)[:, :k, 0]                    # utils.py:135
            cdev = torch.from_numpy(coef).to(dev)
            arr = (ctypes.c_void_p * k)(*[Th[h].data_ptr() for _, h in hist])
            _lib.check(lib.apy_zcombine_seg(k, arr, _vp(cdev), nb, L, _vp(Xn), st), "apy_zcombine_seg")
        else:
            step(X, Xn, R, s)
        energies(Xn, s_new)
        ws = workspace(red, dev)
        _lib.check(lib.apy_zdiffsq_batch(_vp(X), _vp(Xn), nb, L, OV, _vp(dsq), _vp(ws), red, st),
                   "apy_zdiffsq_batch")
        E_new = s_new.cpu().numpy()
        d = dsq.cpu().numpy()
        act = np.flatnonzero(active)
        if len(act) == nb:
            X.copy_(Xn)
            s.copy_(s_new)
        else:
            for q in act:
                X[q].copy_(Xn[q])
                s[q:q + 1].copy_(s_new[q:q + 1])
        for q in act:
            rms1, rms2 = np.sqrt(d[q, 0]), np.sqrt(d[q, 1])
            dE = E[q] - E_new[q]
            E[q] = E_new[q]
            iters[q] = it
            if print_level > 0:
                print(" %02d %3d %20.12f %20.12f %20.12f" % (it, q, (E[q] + shift[q]).real, dE.real, rms2.real))
            if _converged(it, dE, [rms1, rms2] if singles else [rms2], params):
                active[q] = False
    if active.any():
        print("Not converged.")
    return E, X, iters


class SpatialCISet:
    """Spin-adapted CISD (singles=True) or CID of nb independent systems of
    one (o, v) — the displaced geometries and fields of a finite-difference
    set, which fin_diff.py:267-372 solves one after another (ci_wfn.py:420-574
    or :51-167 each) — iterated together: every iteration is one
    apy_cisd_sp_step_multi over all systems (each with its own integral
    blocks, Fock matrix and orbital energies, every product batched), then
    per system the reference's DIIS extrapolation and convergence test
    (_solve_set).  Each system's iterates are those of its own SpatialCI
    solve.

    eri_mos: nb chemists' MO tensors (pq|rs) [m]^4 on one device; focks: nb
    [m, m] MO Fock matrices; eps_list: nb orbital-energy arrays [m]; o: the
    occupied count (the same for all)."""

    def __init__(self, eri_mos, focks, eps_list, o, singles=True):
        self.lib = lib = _lib.load()
        self.nb = nb = len(eri_mos)
        if nb < 1 or not (len(focks) == len(eps_list) == nb):
            raise ValueError("one Fock matrix and one orbital-energy array per system")
        self.device = dev = eri_mos[0].device
        m = eri_mos[0].shape[0]
        ensure_default_plans(dev)
        if any(tuple(e.shape) != (m,) * 4 for e in eri_mos):
            raise ValueError("all systems must have the same number of orbitals")
        self.O, self.V = o, v = int(o), m - int(o)
        self.singles = bool(singles)
        self.OV, self.L = o * v, o * v + o * o * v * v
        st = stream_of(dev)
        sizes = (ctypes.c_int64 * len(_lib.SP_BLOCKS))()
        _lib.check(lib.apy_sp_ints_sizes(o, v, sizes), "apy_sp_ints_sizes")
        # independent systems: every product runs per system, the stacked-w4 block is not used
        names = tuple(b for b in _lib.SP_BLOCKS if b != "w4t") if self.singles else CID_SP_BLOCKS
        self.blocks = {b: torch.empty((nb, int(sizes[k])), dtype=CDT, device=dev)
                       for k, b in enumerate(_lib.SP_BLOCKS) if b in names}
        self.structs = [_lib.SpInts(o, v, *[self.blocks[b][q].data_ptr() if b in self.blocks else None
                                            for b in _lib.SP_BLOCKS]) for q in range(nb)]
        self.fock = torch.stack([f.to(dev, CDT) for f in focks]).contiguous()
        self.eps = torch.stack([torch.as_tensor(np.array(e, dtype=np.float64) if not isinstance(e, torch.Tensor)
                                                else e, dtype=torch.float64).to(dev) for e in eps_list]).contiguous()
        self.hoo = torch.empty((nb, o ** 4), dtype=CDT, device=dev)
        self.g = torch.empty((nb, 2 * o * v), dtype=CDT, device=dev)
        for q in range(nb):
            eri = eri_mos[q].to(CDT).contiguous()
            _lib.check(lib.apy_sp_ints_from_mo(m, _vp(eri), ctypes.byref(self.structs[q]), st),
                       "apy_sp_ints_from_mo")
            _lib.check(lib.apy_cisd_sp_prepare(o, v, _vp(self.fock[q]), ctypes.byref(self.structs[q]),
                                               _vp(self.hoo[q]), _vp(self.g[q]), st), "apy_cisd_sp_prepare")
        self.w21_f = torch.empty_like(self.blocks["w21"])
        _lib.check(lib.apy_cisd_sp_fold_fock(o, v, nb, _vp(self.fock), _vp(self.blocks["w21"]), _vp(self.w21_f), st),
                   "apy_cisd_sp_fold_fock")
        self.ws_bytes = int(lib.apy_cisd_sp_multi_workspace(o, v, int(self.singles), nb))
        self.en_bytes = nb * int(lib.apy_cisd_sp_energy_workspace())

    def _step(self, x, x_new, out, s):
        args = _lib.CisdSpArgs(self.O, self.V, self.fock.data_ptr(), self.eps.data_ptr(),
                               ctypes.pointer(self.structs[0]), self.hoo.data_ptr(), self.g.data_ptr(),
                               int(self.singles), s.data_ptr(), x.data_ptr(), x_new.data_ptr(), out.data_ptr())
        args.w21_f = self.w21_f.data_ptr()
        ws = workspace(self.ws_bytes, self.device)
        _lib.check(self.lib.apy_cisd_sp_step_multi(ctypes.byref(args), self.nb, _vp(ws), self.ws_bytes,
                                                   stream_of(self.device)), "apy_cisd_sp_step_multi")

    def _energies(self, x, out):
        ws = workspace(self.en_bytes, self.device)
        _lib.check(self.lib.apy_cisd_sp_energy_multi(self.O, self.V, int(self.singles), self.nb, _vp(self.g),
                                                     _vp(self.blocks["l"]), _vp(x), _vp(out), _vp(ws),
                                                     self.en_bytes, stream_of(self.device)),
                   "apy_cisd_sp_energy_multi")

    def solve(self, params, print_level=0, max_diis=7):
        """Returns (E [nb] complex host array, t1 [nb,o,v] or None,
        t2 [nb,o,o,v,v] device views, iterations per system)."""
        E, X, iters = _solve_set(self._step, self._energies, self.nb, self.L, self.OV, self.singles,
                                 self.device, params, print_level, max_diis)
        self.X = X
        nb, O, V = self.nb, self.O, self.V
        t1 = X[:, :self.OV].view(nb, O, V) if self.singles else None
        return E, t1, X[:, self.OV:].view(nb, O, O, V, V), iters

    def spin_orbital_amplitudes(self):
        """The solved amplitudes in the spin-orbital layout of solve_CISD_SO /
        solve_CID_SO (ci_wfn.py:171-416; spin orbital 2p + spin, utils.py:283-365):
        (t1 [nb, 2o, 2v], t2 [nb, 2o, 2o, 2v, 2v]) device tensors
        (apy_sp_amplitudes_to_so).  For an RHF reference the spin-orbital
        equations are solved by exactly these amplitudes."""
        nb, o, v = self.nb, self.O, self.V
        t1 = torch.empty((nb, 2 * o, 2 * v), dtype=CDT, device=self.device)
        t2 = torch.empty((nb, 2 * o, 2 * o, 2 * v, 2 * v), dtype=CDT, device=self.device)
        _lib.check(self.lib.apy_sp_amplitudes_to_so(o, v, nb, int(self.singles), _vp(self.X), _vp(t1), _vp(t2),
                                                    stream_of(self.device)), "apy_sp_amplitudes_to_so")
        return t1, t2
