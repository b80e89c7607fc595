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
    out['vovv_p'] = W['vovv'][:, :, ia, ib]
    out['ooov'] = W['ooov']
    out['vvoo_t'] = W['vvoo'].transpose(2, 3, 0, 1)
    out['oovv'] = W['oovv']
    return {k: np.ascontiguousarray(v).ravel() for k, v in out.items()}


def block_element_indices(name, flat, O, V):
    """Full-tensor indices (p, q, r, s) of the entries at flat positions
    `flat` of the device block `name` (layouts of apy_so_ints): the stored
    value is <pq||rs> at those indices."""
    flat = np.asarray(flat, dtype=np.int64)
    io, jo = np.triu_indices(O, 1)
    ia, ib = np.triu_indices(V, 1)
    Po, Pv = len(io), len(ia)
    if name == 'vvvv_p':                       # [(a<b)][(c<d)] <ab||cd>
        x, y = np.divmod(flat, Pv)
        return O + ia[x], O + ib[x], O + ia[y], O + ib[y]
    if name == 'ring':                         # [(j,b)][(k,c)] <kb||cj>
        jb, kc = np.divmod(flat, O * V)
        j, b = np.divmod(jb, V)
        k, c = np.divmod(kc, V)
        return k, O + b, O + c, j
    if name == 'oooo_p':                       # [(i<j)][(k<l)] <kl||ij>
        x, y = np.divmod(flat, Po)
        return io[y], jo[y], io[x], jo[x]
    if name == 'ovoo_p':                       # [k][b][(i<j)] <kb||ij>
        kb, x = np.divmod(flat, Po)
        k, b = np.divmod(kb, V)
        return k, O + b, io[x], jo[x]
    if name == 'vvvo_p':                       # [(a<b)][j][c] <ab||cj>
        x, jc = np.divmod(flat, O * V)
        j, c = np.divmod(jc, V)
        return O + ia[x], O + ib[x], O + c, j
    if name == 'vovv_p':                       # [a][j][(c<b)] <aj||cb>
        a, j, x = np.unravel_index(flat, (V, O, Pv))
        return O + a, j, O + ia[x], O + ib[x]
    if name == 'ooov':                         # [k][j][i][b] <kj||ib>
        k, j, i, b = np.unravel_index(flat, (O, O, O, V))
        return k, j, i, O + b
    if name == 'vvoo_t':                       # [i][j][a][b] <ab||ij>
        i, j, a, b = np.unravel_index(flat, (O, O, V, V))
        return O + a, O + b, i, j
    if name == 'oovv':                         # [i][j][a][b] <ij||ab>
        i, j, a, b = np.unravel_index(flat, (O, O, V, V))
        return i, j, O + a, O + b
    raise KeyError(name)
