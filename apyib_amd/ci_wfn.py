"""CID / CISD wavefunction with apyib's interface (ci_wfn.py:19-574) on the
GPU.

The spin-orbital solvers run the reference's Jacobi + DIIS iteration with
the sigma of apy_cisd_so_step; the spin-adapted (spatial) solvers run the
same iteration over spatial amplitudes with the residual of
apy_cisd_sp_step (ci_wfn.py:420-574 and :51-167), so both follow the
reference iterate for iterate.
"""
import numpy as np

from . import transforms as tr
from .so_ci import CID_BLOCKS, SOCI, SOIntegrals
from .spatial_ci import SPIntegrals, SpatialCI
from .utils import compute_ERI_MO_device, compute_F_MO_device, device, get_slices


class ci_wfn(object):
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
        self.D_ia = self.eps_o.reshape(-1, 1) - self.eps_v
        self.D_ijab = (self.eps_o.reshape(-1, 1, 1, 1) + self.eps_o.reshape(-1, 1, 1) - self.eps_v.reshape(-1, 1)
                       - self.eps_v)
        self._F_MO, self.E_fc = compute_F_MO_device(self.parameters, self.wfn, self.C_list)
        self._ERI_MO = compute_ERI_MO_device(self.parameters, self.wfn, self.C_list)

    @property
    def F_MO(self):
        return self._F_MO.cpu().numpy()

    @property
    def ERI_MO(self):
        return self._ERI_MO.cpu().numpy()

    def _so_solver(self, singles):
        O = 2 * len(self.eps_o)
        if singles:
            ints = SOIntegrals.from_mo(self._ERI_MO, O)
        else:
            ints = SOIntegrals.from_mo(self._ERI_MO, O, blocks=CID_BLOCKS)
        F_SO = tr.so_fock(self._F_MO)
        eps = np.concatenate([np.repeat(self.eps_o, 2), np.repeat(self.eps_v, 2)])
        return SOCI(ints, F_SO, eps, singles=singles)

    def _e_offset(self):
        return complex(self.wfn.E_SCF) + self.H.E_nuc

    def solve_CISD_SO(self, print_level=0):
        """ci_wfn.py:263-416.  Returns (E_CISD, t1, t2)."""
        s = self._so_solver(True)
        E, t1, t2, _ = s.solve(self.parameters, print_level, e_offset=self._e_offset())
        return np.complex128(E), t1.cpu().numpy(), t2.cpu().numpy()

    def solve_CID_SO(self, print_level=0):
        """ci_wfn.py:171-259.  Returns (E_CID, t2)."""
        s = self._so_solver(False)
        E, _, t2, _ = s.solve(self.parameters, print_level, e_offset=self._e_offset())
        return np.complex128(E), t2.cpu().numpy()

    def _spatial_solver(self, singles):
        ints = SPIntegrals.from_mo(self._ERI_MO, len(self.eps_o), singles=singles)
        eps = np.concatenate([self.eps_o, self.eps_v])
        return SpatialCI(ints, self._F_MO, eps, singles=singles)

    def solve_CISD(self, print_level=0):
        """ci_wfn.py:420-574.  Returns (E_CISD, t1, t2)."""
        s = self._spatial_solver(True)
        E, t1, t2, _ = s.solve(self.parameters, print_level, e_offset=self._e_offset())
        t1s = t1.cpu().numpy()
        t2s = t2.cpu().numpy()
        print("t-Amplitude Data:")
        print("Maximum t1: ", np.max(t1s))
        print("Maximum t2: ", np.max(t2s))
        return np.complex128(E), t1s, t2s

    def solve_CID(self, print_level=0):
        """ci_wfn.py:51-167.  Returns (E_CID, t2)."""
        s = self._spatial_solver(False)
        E, _, t2, _ = s.solve(self.parameters, print_level, e_offset=self._e_offset())
        t2s = t2.cpu().numpy()
        print("t-Amplitude Data:")
        print("Maximum t2: ", np.max(t2s))
        return np.complex128(E), t2s
