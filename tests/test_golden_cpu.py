"""The oracle reproduces the committed golden fixtures (tests/golden/)."""
import os

import numpy as np
import pytest

from helpers import O

GOLD = os.path.join(os.path.dirname(__file__), "golden", "qem_golden.npz")


@pytest.fixture(scope="module")
def g():
    return np.load(GOLD, allow_pickle=False)


def test_two_atom(g):
    ri, rj = g["two_atom_ri"], g["two_atom_rj"]
    for row in g["two_atom_table"]:
        g1, g2, alpha, single, sem = row[:5]
        p = O.params((10.0, 10.0, 2.0), 1.0, alpha, g1, g2, 1.0, 1e-6, 30, int(sem))
        f, fzj = O.F_short_ij(p, 1.0, 1.0, ri, rj, single=bool(single))
        fs = O.F_short_i(p, 1.0, ri, single=bool(single))
        assert np.allclose([*f, fzj, fs[2]], row[5:], rtol=1e-13, atol=0)


def test_four_atom(g):
    xyz, q = g["four_atom_xyz"], g["four_atom_q"]
    zl = O.sortperm_z(xyz)
    for (g1, g2, kx, ky, k), ref in zip(g["four_atom_params"], g["four_atom_parts"]):
        k3 = np.array([kx, ky, k])
        p = O.params((10.0, 10.0, 2.0), 1.0, 1.0, g1, g2, 1.0, 1e-6, 30)
        got = [O.Fl_non_sort(k3, xyz, q, g1, g2, 2.0), O.Fl_sort(k3, xyz, q, zl),
               O.Flz_self(k3, xyz, q, g1, g2, 2.0), O.Flz_sort_k0(q, zl),
               O.F_l_total(k3, xyz, q, zl, p)]
        for a, b in zip(got, ref):
            assert np.allclose(a, b, rtol=1e-13, atol=1e-300)


def test_slab_n100(g):
    from qem4molly_amd.workloads import CONFIGS
    cfg = CONFIGS["slab_n100"]
    xyz, q = g["slab_xyz"], g["slab_q"]
    pairs = O.neighbor_pairs(xyz, cfg.Lx, cfg.Ly, cfg.r_cutoff)
    assert np.array_equal(pairs, g["slab_pairs"])
    assert np.array_equal(O.sortperm_z(xyz), g["slab_zlist"])
    po = O.params(cfg.L, cfg.r_cutoff, cfg.alpha, cfg.gamma_1, cfg.gamma_2, cfg.eps_0,
                  cfg.accuracy, cfg.N_t, 0)
    assert np.allclose(O.short_forces(po, xyz, q, pairs), g["slab_short_as_written"],
                       rtol=1e-13, atol=1e-15)
    K, P, sK, _ = O.kset(cfg.Lx, cfg.Ly, cfg.alpha, cfg.accuracy)
    assert np.array_equal(K[g["slab_kidx"]], g["slab_kvecs"])
    assert sK == float(g["slab_sum_K"])
    assert np.allclose(O.long_forces(po, xyz, q, g["slab_kvecs"], weight=sK / cfg.P),
                       g["slab_long"], rtol=1e-12, atol=1e-15)
