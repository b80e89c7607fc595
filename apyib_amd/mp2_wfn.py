"""MP2 wavefunction with apyib's interface (mp2_wfn.py:16-87) on the GPU."""
import numpy as np

from . import transforms as tr
from .so_ci import MP2_BLOCKS, SOIntegrals
from .utils import compute_ERI_MO_device, device, get_slices


class mp2_wfn(object):
    def __init__(self, parameters, wfn):
        self.parameters = parameters
        self.H = wfn.H
        self.wfn = wfn
        self.C = wfn.C
        self.C_list, self.I_list = get_slices(self.parameters, self.wfn)
        o = self.C_list[1]
        v = self.C_list[2]
        self.eps_o = wfn.eps[o]
        self.eps_v = wfn.eps[v]
        self.D_ijab = (self.eps_o.reshape(-1, 1, 1, 1) + self.eps_o.reshape(-1, 1, 1) - self.eps_v.reshape(-1, 1)
                       - self.eps_v)

    def solve_MP2(self):
        """mp2_wfn.py:42-59: t = (ai|bj)/D, E = sum (2(ia|jb) - (ib|ja)) t."""
        eri = compute_ERI_MO_device(self.parameters, self.wfn, self.C_list)
        eps = np.concatenate([self.eps_o, self.eps_v])
        E, t2 = tr.mp2_spatial(eri, eps, len(self.eps_o))
        return np.complex128(complex(E.cpu()[0])), t2.cpu().numpy()

    def solve_MP2_SO(self):
        """mp2_wfn.py:64-87: t = <ab||ij>/D, E = 1/4 sum <ij||ab> t."""
        import torch
        dev = device()
        eri = compute_ERI_MO_device(self.parameters, self.wfn, self.C_list)
        O = 2 * len(self.eps_o)
        ints = SOIntegrals.from_mo(eri, O, blocks=MP2_BLOCKS)
        eps = torch.from_numpy(np.concatenate([np.repeat(self.eps_o, 2), np.repeat(self.eps_v, 2)])).to(dev)
        E, (t2,) = tr.mp2_so_batched([(O, ints.V, eps, ints['vvoo_t'], ints['oovv'])], dev)
        return np.complex128(complex(E.cpu()[0])), t2.cpu().numpy()
