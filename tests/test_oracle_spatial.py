"""Spatial (spin-adapted) oracle against the spin-orbital oracle and the
reference's golden energies.  CPU only."""
import numpy as np
import pytest

from oracle import spatial, so, transforms
from tests.test_oracle_golden import GOLD_CID_H2O, GOLD_MP2_H2O, h2o, h22  # noqa: F401
from tests.helpers import PARAMS


def _inputs(sys_):
    mol, bs, S, T, V, G, En, E, C, eps = sys_
    nbf = C.shape[0]
    C_list = [slice(0, 0), slice(0, 0), slice(0, 0), slice(0, nbf)]
    eri_mo = transforms.compute_ERI_MO(G, C, C_list)
    F_mo = transforms.compute_F_MO(T + V, G, C, [slice(0, 0), slice(0, mol.Z.sum().astype(int) // 2),
                                                  slice(0, 0), slice(0, nbf)], False)[0]
    return E + En, eri_mo, F_mo, eps


def test_spatial_mp2_golden(h2o):
    Eref, eri_mo, F, eps = _inputs(h2o)
    E, _ = spatial.solve_mp2(eri_mo.swapaxes(1, 2), eps, 5)
    assert abs(Eref + E - GOLD_MP2_H2O) < 1e-11


def test_spatial_cid_golden(h2o):
    Eref, eri_mo, F, eps = _inputs(h2o)
    E, t2, _ = spatial.solve_cid(F, eri_mo.swapaxes(1, 2), eps, 5, PARAMS)
    assert abs(Eref + E - GOLD_CID_H2O) < 1e-11


@pytest.mark.parametrize("which", ["h2o", "h22"])
def test_spatial_cisd_equals_alpha_beta_block_of_spin_orbital(which, request):
    sys_ = request.getfixturevalue(which)
    Eref, eri_mo, F, eps = _inputs(sys_)
    O = int(sys_[0].Z.sum()) // 2
    E, t1, t2, _ = spatial.solve_cisd(F, eri_mo.swapaxes(1, 2), eps, O, PARAMS)
    eri_so = transforms.compute_ERI_SO(eri_mo).swapaxes(1, 2)
    F_so = transforms.compute_F_SO(F)
    e = np.repeat(eps, 2)
    Es, s1, s2, _ = so.solve_cisd_so(F_so, eri_so, e[:2 * O], e[2 * O:], 2 * O, PARAMS)
    assert abs(E - Es) < 1e-11
    assert np.abs(t1 - s1[0::2, 0::2]).max() < 1e-9
    assert np.abs(t2 - s2[0::2, 1::2, 0::2, 1::2]).max() < 1e-9
