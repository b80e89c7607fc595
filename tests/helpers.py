"""Shared test helpers (inputs only; every expected value comes from the
oracle package or from closed forms)."""
import numpy as np


def rand_amplitudes(O, V, seed, scale=0.1):
    rng = np.random.default_rng(seed)
    t1 = scale * (rng.standard_normal((O, V)) + 1j * rng.standard_normal((O, V)))
    t2 = scale * (rng.standard_normal((O, O, V, V)) + 1j * rng.standard_normal((O, O, V, V)))
    t2 = t2 - t2.transpose(1, 0, 2, 3)
    t2 = t2 - t2.transpose(0, 1, 3, 2)
    return t1, t2


def brillouin(F, O):
    Fb = F.copy()
    Fb[:O, O:] = 0
    Fb[O:, :O] = 0
    return Fb


PARAMS = dict(e_convergence=1e-12, d_convergence=1e-12, DIIS=True, max_iterations=200)
