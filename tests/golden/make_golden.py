"""Writes tests/golden/qem_golden.npz: seeded inputs and FP64 oracle outputs.

The reference ships no fixtures (its tests compare live against other
implementations), so these are the oracle's outputs on the geometries its
tests use, plus the 100-atom slab:
  two_atom   test/Force_short_test.jl:30-34 geometry: F_short_ij and F_short_i
             for contrasts (0.9, 0.8) (the commented values at lines 13-14),
             (0.3, 0.1), (-0.7, 0.95); alpha 1 and 0.1; single on/off; three
             semantics settings
  four_atom  test/Force_long_test.jl:13-14 geometry: Fl_non_sort, Fl_sort,
             Flz_self, Flz_sort_k0, F_l_total for three seeded (k, gamma)
  slab_n100  neighbour pairs, short forces (as written and corrected),
             a batch of 100 k vectors, long forces, total
Run from the repository root:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402
from qem4molly_amd.workloads import CONFIGS, make_system  # noqa: E402

SEMS = (0, O.SEM_SWAP_LIMITS, O.SEM_CORRECTED)


def main():
    out = {}
    # two-atom short range
    ri = np.array([5.0, 5.0, 0.2]); rj = np.array([5.1, 5.2, 0.15])
    rows = []
    for g1, g2 in [(0.9, 0.8), (0.3, 0.1), (-0.7, 0.95)]:
        for alpha in (1.0, 0.1):
            for single in (0, 1):
                for sem in SEMS:
                    p = O.params((10.0, 10.0, 2.0), 1.0, alpha, g1, g2, 1.0, 1e-6, 30, sem)
                    f, fzj = O.F_short_ij(p, 1.0, 1.0, ri, rj, single=bool(single))
                    fs = O.F_short_i(p, 1.0, ri, single=bool(single))
                    rows.append([g1, g2, alpha, single, sem, *f, fzj, fs[2]])
    out["two_atom_ri"] = ri
    out["two_atom_rj"] = rj
    out["two_atom_table"] = np.array(rows)  # g1 g2 alpha single sem fx fy fz fz_j fz_self

    # four-atom long range
    xyz = np.array([[5.0, 5.0, 0.2], [5.1, 5.2, 0.15], [4.9, 4.9, 0.3], [5.5, 5.3, 0.4]])
    q = np.array([1.0, -1.0, 1.0, -1.0])
    rng = np.random.default_rng(20240601)
    zl = O.sortperm_z(xyz)
    p4 = []
    parts = []
    for t in range(3):
        g1, g2 = rng.random(2)
        kx, ky = rng.random(2) + 0.1
        k3 = np.array([kx, ky, np.hypot(kx, ky)])
        p = O.params((10.0, 10.0, 2.0), 1.0, 1.0, g1, g2, 1.0, 1e-6, 30)
        p4.append([g1, g2, *k3])
        parts.append(np.stack([O.Fl_non_sort(k3, xyz, q, g1, g2, 2.0),
                               O.Fl_sort(k3, xyz, q, zl),
                               O.Flz_self(k3, xyz, q, g1, g2, 2.0),
                               O.Flz_sort_k0(q, zl),
                               O.F_l_total(k3, xyz, q, zl, p)]))
    out["four_atom_xyz"] = xyz
    out["four_atom_q"] = q
    out["four_atom_params"] = np.array(p4)          # g1 g2 kx ky k
    out["four_atom_parts"] = np.stack(parts)        # [case][non_sort, sort, self, k0, total][atom][3]

    # 100-atom slab
    cfg = CONFIGS["slab_n100"]
    xyz, q = make_system(cfg, seed=0)
    pairs = O.neighbor_pairs(xyz, cfg.Lx, cfg.Ly, cfg.r_cutoff)
    K, P, sK, _ = O.kset(cfg.Lx, cfg.Ly, cfg.alpha, cfg.accuracy)
    prob = P.astype(np.float64) / P.astype(np.float64).sum()
    kidx = np.random.default_rng(7).choice(len(K), cfg.P, p=prob)
    po = O.params(cfg.L, cfg.r_cutoff, cfg.alpha, cfg.gamma_1, cfg.gamma_2, cfg.eps_0,
                  cfg.accuracy, cfg.N_t, 0)
    pc = O.params(cfg.L, cfg.r_cutoff, cfg.alpha, cfg.gamma_1, cfg.gamma_2, cfg.eps_0,
                  cfg.accuracy, cfg.N_t, O.SEM_CORRECTED)
    out["slab_xyz"] = xyz
    out["slab_q"] = q
    out["slab_pairs"] = pairs
    out["slab_kidx"] = kidx.astype(np.int64)
    out["slab_kvecs"] = K[kidx]
    out["slab_sum_K"] = np.array(sK)
    out["slab_short_as_written"] = O.short_forces(po, xyz, q, pairs)
    out["slab_short_corrected"] = O.short_forces(pc, xyz, q, pairs)
    out["slab_long"] = O.long_forces(po, xyz, q, K[kidx], weight=sK / cfg.P)
    out["slab_zlist"] = O.sortperm_z(xyz)
    np.savez_compressed(os.path.join(HERE, "qem_golden.npz"), **out)
    print("wrote", os.path.join(HERE, "qem_golden.npz"))


if __name__ == "__main__":
    main()
