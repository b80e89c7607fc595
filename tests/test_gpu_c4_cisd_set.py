"""The complex128 CISD field-derivative set at its own size: O=24, V=200
spin orbitals sharing one seeded <pq||rs> (6.3 GB packed vvvv), 6 instances
= 3 imaginary-antisymmetric one-electron directions x +/- 1e-4 on F.  The
batched multi-instance Davidson against per-instance plain solves (1e-10 in
E), the shared-subspace field set against the batched solves, the
adjoint dE/dF (one-particle density of the base root, the variational
lambda = conj(c)) against central differences of the set (1e-7 relative to
sum |D| |dF|), and one displaced instance's root checked as an eigenpair of
the oracle's Hermitian CISD matrix (ci_wfn.py:309-334 + f_vo coupling).

Host memory: about 40 GB for the dense oracle blocks of one sigma."""
import numpy as np
import pytest
import torch

from apyib_amd import synthetic
from apyib_amd.davidson import FieldSetDavidson, MultiDavidson
from apyib_amd.so_ci import SOIntegrals
from oracle import davidson as odav
from oracle import so

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(1200)]

O, V, SEED, H = 24, 200, 1, 1e-4
CONV = dict(e_conv=1e-12, r_conv=1e-6)


@pytest.fixture(scope="module")
def c4(dev):
    ints = SOIntegrals.synthetic(O, V, SEED, device=dev)
    F, eps = synthetic.fock(O, V, SEED)
    n = O + V
    F0 = torch.from_numpy(F).to(dev)
    dirs = [torch.from_numpy(synthetic.fock_perturbation(n, SEED * 1000 + 500 + k, 'field')).to(dev)
            for k in range(3)]
    focks = [F0 + s * H * d for d in dirs for s in (+1, -1)]
    md = MultiDavidson(ints, focks, eps, 8)
    Eb = np.asarray(md.solve(**CONV))
    cb = [d.c.clone() for d in md.dav]
    del md
    yield ints, F, eps, F0, dirs, focks, Eb, cb
    del ints
    torch.cuda.empty_cache()


def test_c4_dirs_are_hermitian_imaginary_antisymmetric(c4):
    for d in c4[4]:
        x = d.cpu().numpy()
        assert np.array_equal(x, x.conj().T) and np.abs(x.real).max() == 0.0


def test_c4_batched_equals_plain(c4):
    ints, F, eps, F0, dirs, focks, Eb, _ = c4
    for k in (0, 3, 5):
        md = MultiDavidson(ints, [focks[k]], eps, 8)
        Ep = md.solve(**CONV)[0]
        del md
        assert abs(Eb[k] - Ep) < 1e-10, k


def test_c4_field_set_equals_batched(c4):
    """The shared-subspace field set (root of F0 first, then the six
    instances in one subspace) against the batched solves: 1e-10 in E, and
    fewer full sigmas than six separate solves take (6 x 11 iterations)."""
    ints, F, eps, F0, dirs, focks, Eb, _ = c4
    inst = [(d, s * H) for d in range(3) for s in (+1, -1)]
    fs = FieldSetDavidson(ints, F0, eps, dirs, inst, max_space=16, base_space=8)
    E = fs.solve(**CONV)
    assert fs.converged
    print("field set: base", fs.base_iterations, "shared", fs.iterations, "sigmas", fs.nsigma)
    assert np.abs(E - Eb).max() < 1e-10
    assert fs.nsigma < 6 * 11
    fd_fs = (E[0::2] - E[1::2]) / (2 * H)
    fd_b = (Eb[0::2] - Eb[1::2]) / (2 * H)
    assert np.abs(fd_fs - fd_b).max() < 1e-6 * np.abs(fd_b).max()
    del fs
    torch.cuda.empty_cache()


def test_c4_adjoint_matches_central_differences(c4):
    ints, F, eps, F0, dirs, focks, Eb, _ = c4
    md = MultiDavidson(ints, [F0], eps, 8)
    md.solve(**CONV)
    D = md.dav[0].gradient_fock()
    del md
    g = np.array([complex(torch.sum(D * dF).cpu()) for dF in dirs])
    scale = np.array([float(torch.sum((D * dF).abs()).cpu()) for dF in dirs])
    fd = (Eb[0::2] - Eb[1::2]) / (2 * H)
    assert np.abs(g.imag).max() <= 1e-12 * scale.max()
    assert (np.abs(g.real - fd) / scale).max() < 1e-7


def test_c4_displaced_root_is_an_eigenpair_of_the_oracle(c4):
    ints, F, eps, F0, dirs, focks, Eb, cb = c4
    k = 2
    host = {nm: ints[nm].cpu().numpy() for nm in ints.blocks}
    W = so.blocks_from_packed(host, O, V)
    del host
    Fk = focks[k].cpu().numpy()
    c = cb[k].cpu().numpy()
    L = O * V + O * O * V * V
    cv = (c[L], c[:O * V].reshape(O, V), c[O * V:L].reshape(O, O, V, V))
    sv = odav.sigma(Fk, W, *cv, O)
    del W
    norm = odav.dot(cv, cv).real
    assert abs(norm - 1.0) < 1e-12
    rq = odav.dot(cv, sv).real
    assert abs(rq - Eb[k]) <= 1e-10 * max(1.0, abs(Eb[k]))
    res = tuple(s - Eb[k] * x for s, x in zip(sv, cv))
    assert np.sqrt(abs(odav.dot(res, res))) < 2e-6
