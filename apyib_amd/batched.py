"""Finite-difference instance sets solved as one batch on the GPU.

The reference forms every finite-difference tensor from independent
solves, one per +/- displacement (fin_diff.py:267-372, a Python loop over
energy()/phase_corrected_energy()).  Here the instances of a set are
solved together: `MP2NCBatch` runs the orbital-invariant spin-orbital MP2
iteration (ci_wfn.py:210-212 driven by <ab||ij>; solve_MP2_SO at diagonal
F, mp2_wfn.py:64-87) for all instances, of possibly different (O, V), in
one fused launch per iteration (apy_mp2_nc_batched_step) on amplitudes
stored packed over (i<j, a<b), optionally with chunks of iterations
captured in a HIP graph, the host checking convergence once per chunk.
`plain_solve` runs the same kernels one instance at a time with a
convergence check every iteration: the per-displacement loop.

`fd_mp2_instances` builds a seeded finite-difference set: a base system
plus +/- h along nuclear-like (real symmetric F and real <pq||rs>
perturbation) and field-like (i x antisymmetric F) directions.
"""
import ctypes

import numpy as np
import torch

from . import _lib, synthetic
from .kernels import stream_of, workspace
from .so_ci import MP2_BLOCKS, SOIntegrals

CDT = torch.complex128


class Instance:
    """One MP2 system on the device: F [n,n], eps [n], drv = <ab||ij> at
    [i,j,a,b], w = <ij||ab>; transpose_f / drv_scale select the lambda
    equations (F^T, driver <ij||ab> / 4)."""

    def __init__(self, O, V, F, eps, drv, w, transpose_f=False, drv_scale=1.0):
        self.O, self.V = int(O), int(V)
        self.F = F.contiguous()
        self.eps = eps.contiguous()
        self.drv = drv.contiguous()
        self.w = w.contiguous() if w is not None else None
        self.transpose_f = bool(transpose_f)
        self.drv_scale = float(drv_scale)


class MP2NCBatch:
    """transpose_f / drv_scale, when given, apply to every instance;
    otherwise each instance's own settings are used (amplitude and lambda
    solves may share one batch)."""

    def __init__(self, instances, transpose_f=None, drv_scale=None):
        self.lib = _lib.load()
        self.inst = list(instances)
        self.n = len(self.inst)
        self.dev = self.inst[0].F.device
        sizes = [int(self.lib.apy_t2_packed_length(I.O, I.V)) for I in self.inst]   # x[(i<j)][(a<b)]
        offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
        self._xbuf = torch.zeros(int(offs[-1]), dtype=CDT, device=self.dev)
        self._xnbuf = torch.empty(int(offs[-1]), dtype=CDT, device=self.dev)
        self.x = [self._xbuf[offs[k]:offs[k + 1]] for k in range(self.n)]
        self.x_new = [self._xnbuf[offs[k]:offs[k + 1]] for k in range(self.n)]
        tab = (_lib.Mp2NcInstance * self.n)()
        for k, I in enumerate(self.inst):
            tr = I.transpose_f if transpose_f is None else bool(transpose_f)
            sc = I.drv_scale if drv_scale is None else float(drv_scale)
            tab[k] = _lib.Mp2NcInstance(I.O, I.V, I.F.data_ptr(), I.eps.data_ptr(), I.drv.data_ptr(),
                                        self.x[k].data_ptr(), self.x_new[k].data_ptr(),
                                        I.w.data_ptr() if I.w is not None else None, int(tr), sc)
        self.table = tab
        nbytes = int(self.lib.apy_mp2_nc_batched_workspace(self.n, tab))
        # a private workspace: the tables stay valid while graphs replay
        self.ws = torch.empty(nbytes, dtype=torch.uint8, device=self.dev)
        items = ctypes.c_int64(0)
        _lib.check(self.lib.apy_mp2_nc_batched_setup(self.n, tab, ctypes.c_void_p(self.ws.data_ptr()), nbytes,
                                                     ctypes.byref(items), stream_of(self.dev)),
                   "apy_mp2_nc_batched_setup")
        self.items = items.value
        self.rnorm = torch.zeros(self.n, dtype=torch.float64, device=self.dev)
        self.energy_out = torch.zeros(self.n, dtype=CDT, device=self.dev)
        self.iterations = 0
        self._graph = None

    def step(self, reverse):
        _lib.check(self.lib.apy_mp2_nc_batched_step(self.n, self.items, int(reverse),
                                                    ctypes.c_void_p(self.rnorm.data_ptr()),
                                                    ctypes.c_void_p(self.ws.data_ptr()), stream_of(self.dev)),
                   "apy_mp2_nc_batched_step")

    def _capture(self, chunk):
        """chunk (even) iterations x -> x_new -> x ... -> x as one graph."""
        torch.cuda.synchronize(self.dev)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            for q in range(chunk):
                self.step(q % 2)
        return g

    def reset(self):
        self._xbuf.zero_()
        self.iterations = 0

    def amplitudes(self, k):
        """Instance k's amplitudes as the dense [O, O, V, V] tensor."""
        I = self.inst[k]
        out = torch.empty((I.O, I.O, I.V, I.V), dtype=CDT, device=self.dev)
        _lib.check(self.lib.apy_t2_unpack(I.O, I.V, ctypes.c_void_p(self.x[k].data_ptr()),
                                          ctypes.c_void_p(out.data_ptr()), stream_of(self.dev)), "apy_t2_unpack")
        return out

    def start_from(self, k, dense):
        """Set instance k's amplitudes from a dense [O, O, V, V] tensor (a warm
        start, e.g. the converged amplitudes of the undisplaced system)."""
        I = self.inst[k]
        dense = dense.contiguous()
        _lib.check(self.lib.apy_t2_pack(I.O, I.V, ctypes.c_void_p(dense.data_ptr()),
                                        ctypes.c_void_p(self.x[k].data_ptr()), stream_of(self.dev)), "apy_t2_pack")

    def solve(self, tol=1e-12, maxit=200, chunk=4, graph=False, persistent=False):
        """Iterate all instances until every residual norm is below tol; the
        amplitudes end in x.  Direct launches check the norms after every
        step (one host read each; an odd final step's x_new is copied back
        in one pass over the packed buffer).  graph=True replays chunks of
        `chunk` (even) steps from a captured HIP graph, checked once per
        chunk (worth it when the same batch is solved repeatedly).
        persistent=True runs the whole solve as one cooperative launch that
        tests convergence on the device (apy_mp2_nc_batched_solve; the same
        iterates and stopping step as the direct launches, one host read at
        the end).  Returns the iteration count."""
        it = 0
        if persistent:
            its = torch.zeros(1, dtype=torch.int64, device=self.dev)
            _lib.check(self.lib.apy_mp2_nc_batched_solve(self.n, self.items, float(tol), int(maxit),
                                                         ctypes.c_void_p(self.rnorm.data_ptr()),
                                                         ctypes.c_void_p(its.data_ptr()),
                                                         ctypes.c_void_p(self.ws.data_ptr()), stream_of(self.dev)),
                       "apy_mp2_nc_batched_solve")
            self.iterations = int(its.cpu())
            return self.iterations
        if graph:
            chunk = max(2, chunk + (chunk % 2))
            if self._graph is None or self._graph[0] != chunk:
                self._graph = (chunk, self._capture(chunk))    # recorded, not run: x is untouched
            while it < maxit:
                self._graph[1].replay()
                it += chunk
                if float(self.rnorm.max().cpu()) ** 0.5 < tol:
                    break
        else:
            rev = 0
            while it < maxit:
                self.step(rev)
                rev ^= 1
                it += 1
                if float(self.rnorm.max().cpu()) ** 0.5 < tol:
                    break
            if rev:                                             # the last step wrote x_new
                self._xbuf.copy_(self._xnbuf)
        self.iterations = it
        return it

    def solve_unbatched_steps(self, tol=1e-12, maxit=200):
        """Step-by-step iteration with a host check after every step (the
        plain-loop protocol), used on single-instance batches."""
        it = 0
        rev = 0
        for it in range(1, maxit + 1):
            self.step(rev)
            rev ^= 1
            if it > 1 and float(self.rnorm.max().cpu()) ** 0.5 < tol:
                break
        if rev == 1:                            # last step wrote x_new: copy back
            for x, xn in zip(self.x, self.x_new):
                x.copy_(xn)
        self.iterations = it
        return it

    def energies(self, scale=0.25):
        _lib.check(self.lib.apy_mp2_nc_batched_energy(self.n, self.items, 0, float(scale),
                                                      ctypes.c_void_p(self.energy_out.data_ptr()),
                                                      ctypes.c_void_p(self.ws.data_ptr()), stream_of(self.dev)),
                   "apy_mp2_nc_batched_energy")
        return self.energy_out


class AdjointDensities:
    """(G, t, lambda) of one base system: G = dE/dF [n*n] and the converged
    amplitude and lambda vectors, unpacked to dense [O, O, V, V] on first
    access.  Unpacks like a tuple: G, t, lam = dens."""

    def __init__(self, batch, k_t, k_lam, G):
        # k_lam None: lambda = conj(t) / 4 (Hermitian system)
        self._batch, self._kt, self._kl = batch, k_t, k_lam
        self.G = G
        self._t = self._lam = None

    @property
    def t(self):
        if self._t is None:
            self._t = self._batch.amplitudes(self._kt).view(-1)
        return self._t

    @property
    def lam(self):
        if self._lam is None:
            self._lam = (self._batch.amplitudes(self._kl).view(-1) if self._kl is not None
                         else 0.25 * self.t.conj())
        return self._lam

    def __iter__(self):
        return iter((self.G, self.t, self.lam))

    def __getitem__(self, i):
        return (self.G, self.t, self.lam)[i]


def mp2_adjoint_set(bases, directions, tol=1e-12, maxit=200, graph=False, hermitian=False):
    """Directional derivatives of the MP2 energies of several base systems:
    for base b the directions directions[b] = (dF list [n*n], dW list or
    None, dV list or None; flattened device tensors).  The amplitude and the
    lambda equations (F^T, driver <ij||ab>/4) of every base are one batch
    (one fused launch per iteration for all of them); with hermitian=True
    (every base has F = F^H and <ab||ij> = conj(<ij||ab>), which the caller
    asserts) the lambda equations are the conjugated amplitude equations
    scaled by 1/4, so only the amplitudes are solved and lambda =
    conj(t) / 4.  Then one
    apy_mp2_adjoint_sums call forms every base's gradient densities and the
    directional sums
        dE = sum dE/dF . dF + sum t/4 . dW + sum lambda . dV,
    read back with one host copy.  Returns ([array of dE per base],
    [AdjointDensities (G, t, lambda) per base])."""
    lib = _lib.load()
    dev = bases[0].F.device
    nbase = len(bases)
    lam_inst = [] if hermitian else [Instance(B.O, B.V, B.F, B.eps, B.w, None, transpose_f=True, drv_scale=0.25)
                                     for B in bases]
    batch = MP2NCBatch(list(bases) + lam_inst)
    batch.solve(tol, maxit, chunk=8, graph=graph)
    items = (_lib.Mp2AdjointItem * nbase)()
    keep, counts = [], []
    ncum = [0]
    for B in bases:
        ncum.append(ncum[-1] + (B.O + B.V) ** 2)
    Gbuf = torch.empty(ncum[-1], dtype=CDT, device=dev)
    dens = []
    for b, B in enumerate(bases):
        dF, dW, dV = directions[b]
        K = len(dF)
        arrs = [(ctypes.c_void_p * K)(*[x.data_ptr() for x in xs]) if xs is not None else None
                for xs in (dF, dW, dV)]
        keep.append(arrs)
        G = Gbuf[ncum[b]:ncum[b + 1]]
        as_ptr = lambda a: ctypes.cast(a, ctypes.POINTER(ctypes.c_void_p)) if a is not None else None
        items[b] = _lib.Mp2AdjointItem(B.O, B.V, batch.x[b].data_ptr(),
                                       None if hermitian else batch.x[nbase + b].data_ptr(),
                                       G.data_ptr(), K, *[as_ptr(a) for a in arrs])
        counts.append(K)
        dens.append(AdjointDensities(batch, b, None if hermitian else nbase + b, G))
    nbytes = int(lib.apy_mp2_adjoint_sums_workspace(nbase, items))
    ws = workspace(nbytes, dev)
    out = torch.empty(max(1, sum(counts)), dtype=CDT, device=dev)
    _lib.check(lib.apy_mp2_adjoint_sums(nbase, items, ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(ws.data_ptr()),
                                        nbytes, stream_of(dev)), "apy_mp2_adjoint_sums")
    sums = out.cpu().numpy()
    res, off = [], 0
    for K in counts:
        res.append(sums[off:off + K].copy())
        off += K
    return res, dens


def mp2_adjoint_directional(base, dF, dW, dV, tol=1e-12, maxit=200):
    """Directional derivatives of the MP2 energy of `base` (mp2_adjoint_set
    with one base).  Returns (array of dE, (G, t, lambda))."""
    out, dens = mp2_adjoint_set([base], [(dF, dW, dV)], tol, maxit)
    return out[0], dens[0]


def plain_solve(instances, tol=1e-12, maxit=200):
    """The per-displacement loop: the same kernels, one instance at a time,
    a convergence check after every iteration.  Returns (energies, iters)."""
    E, its = [], []
    for I in instances:
        b = MP2NCBatch([I])
        its.append(b.solve_unbatched_steps(tol, maxit))
        E.append(b.energies()[0])
    return torch.stack(E), its


def base_system(O, V, seed, device, std=0.05):
    """(F, eps, ints) of a seeded synthetic system (MP2 blocks only)."""
    F, eps = synthetic.fock(O, V, seed)
    ints = SOIntegrals.synthetic(O, V, seed, std=std, device=device, blocks=MP2_BLOCKS)
    return (torch.from_numpy(F).to(device), torch.from_numpy(np.asarray(eps, dtype=np.float64)).to(device), ints)


def fd_mp2_instances(O, V, seed, natom, h=1e-4, device="cuda", sizes=None, with_bases=False):
    """A finite-difference set of 2 (3 natom + 3) instances: for each
    nuclear-like direction k (real symmetric dF_k, real <pq||rs> draw with
    tag 100 + k) and each field-like direction (i x antisymmetric dF), the
    base system displaced by +h and -h.  With `sizes` (a list of (O, V) per
    direction), each direction uses its own base system of that size (the
    uneven-size stress case).  Returns (instances, directions) where
    directions[d] = (kind, seed_d, (O, V)) and instances 2d, 2d+1 are +h, -h;
    with_bases=True also returns the undisplaced systems, one Instance per
    distinct size, and base_of[d], the index of direction d's system."""
    ndir = 3 * natom + 3
    dirs, out = [], []
    cache = {}
    for d in range(ndir):
        Od, Vd = sizes[d] if sizes is not None else (O, V)
        if (Od, Vd) not in cache:
            cache[(Od, Vd)] = base_system(Od, Vd, seed + 7919 * Od + 104729 * Vd, device)
        F0, eps, _ = cache[(Od, Vd)]
        kind = 'nuclear' if d < 3 * natom else 'field'
        sd = seed * 1000 + d
        dF = torch.from_numpy(synthetic.fock_perturbation(Od + Vd, sd, kind)).to(device)
        dirs.append((kind, sd, (Od, Vd)))
        for sgn in (+1, -1):
            ints = SOIntegrals.synthetic(Od, Vd, seed + 7919 * Od + 104729 * Vd, device=device, blocks=MP2_BLOCKS)
            if kind == 'nuclear':
                ints.add_synthetic(sd, tag=100 + d, std=1.0, scale=sgn * h, accumulate=True, real_only=True)
            n = Od * Od * Vd * Vd
            out.append(Instance(Od, Vd, F0 + sgn * h * dF, eps, ints.blocks['vvoo_t'].view(n),
                                ints.blocks['oovv'].view(n)))
    if not with_bases:
        return out, dirs
    keys = list(cache)
    bases = []
    for (Od, Vd) in keys:
        F0, eps, ints = cache[(Od, Vd)]
        n = Od * Od * Vd * Vd
        bases.append(Instance(Od, Vd, F0, eps, ints.blocks['vvoo_t'].view(n), ints.blocks['oovv'].view(n)))
    base_of = [keys.index(d[2]) for d in dirs]
    return out, dirs, bases, base_of


def warm_batch_solve(instances, bases, base_of_instance, tol=1e-12, maxit=400, graph=False):
    """The finite-difference set solved batched with warm starts: the
    undisplaced systems first (one batch), then every displaced instance
    started from its system's converged amplitudes (one batch).  Returns
    (energies of the instances, base iterations, instance iterations)."""
    bb = MP2NCBatch(bases)
    itb = bb.solve(tol, maxit, chunk=4, graph=graph)
    b = MP2NCBatch(instances)
    for k, q in enumerate(base_of_instance):
        b.x[k].copy_(bb.x[q])                   # packed amplitudes of the same (O, V) layout
    it = b.solve(tol, maxit, chunk=2, graph=graph)
    return b.energies().clone(), itb, it
