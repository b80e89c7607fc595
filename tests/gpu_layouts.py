"""Expected device-block layouts (include/apyib_amd.h, apy_so_ints) built on
the host from the oracle's antisymmetrized blocks."""
import numpy as np

from oracle import so


def expected_blocks(A, O):
    """A: full antisymmetrized <pq||rs>.  Returns {name: flat complex array}."""
    n = A.shape[0]
    V = n - O
    W = so.blocks_from_eri(A / 2, O)
    io, jo = np.triu_indices(O, 1)
    ia, ib = np.triu_indices(V, 1)
    out = {}
    out['vvvv_p'] = W['vvvv'][ia, ib][:, ia, ib]
    out['ring'] = W['ovvo'].transpose(3, 1, 0, 2).reshape(O * V, O * V)
    out['oooo_p'] = W['oooo'].transpose(2, 3, 0, 1)[io, jo][:, io, jo]
    out['ovoo_p'] = W['ovoo'][:, :, io, jo]
    out['vvvo_p'] = W['vvvo'][ia, ib].transpose(0, 2, 1)
    out['vovv'] = W['vovv']
    out['ooov'] = W['ooov']
    out['vvoo_t'] = W['vvoo'].transpose(2, 3, 0, 1)
    out['oovv'] = W['oovv']
    return {k: np.ascontiguousarray(v).ravel() for k, v in out.items()}
