"""Synthetic slab workloads (seeded) for tests and benchmarks.

Names follow the measurement plan of the force path:
  slab_n100     N=100,     20 x 20 x 5,   q = (-1)^i,            gamma 0.9/0.9,
                alpha 1.0, r_c 5,  N_t 30, P 100   (CPU-runnable reference case)
  slab_n10k     N=10,000,  100 x 100 x 10, alternating +-1,      gamma_down -0.5,
                gamma_up 0.9, alpha 0.5, r_c 8, N_t 40, P 200
  slab_n100k    N=100,000, 316 x 316 x 20, 33,333 x (+2), 66,666 x (-1), 1 x 0
                (exactly neutral), gamma 0.95/0.95, alpha 0.3, r_c 10, N_t 50, P 400
x, y ~ U[0, L), z ~ U[0.5, Lz - 0.5), accuracy 1e-6, drawn from
numpy.random.default_rng(seed) in the order x, y, z (then the charge
permutation for slab_n100k).  gamma_1 is the lower interface (z = 0),
gamma_2 the upper one (z = Lz).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass(frozen=True)
class SlabConfig:
    name: str
    n: int
    Lx: float
    Ly: float
    Lz: float
    gamma_1: float   # gamma_down, z = 0
    gamma_2: float   # gamma_up,   z = Lz
    alpha: float
    r_cutoff: float
    N_t: int
    P: int
    accuracy: float = 1e-6
    eps_0: float = 1.0

    @property
    def L(self):
        return (self.Lx, self.Ly, self.Lz)


CONFIGS = {
    "slab_n100": SlabConfig("slab_n100", 100, 20.0, 20.0, 5.0, 0.9, 0.9, 1.0, 5.0, 30, 100),
    "slab_n10k": SlabConfig("slab_n10k", 10_000, 100.0, 100.0, 10.0, -0.5, 0.9, 0.5, 8.0, 40, 200),
    "slab_n100k": SlabConfig("slab_n100k", 100_000, 316.0, 316.0, 20.0, 0.95, 0.95, 0.3, 10.0,
                             50, 400),
}


def make_system(cfg: SlabConfig | str, seed: int = 0):
    """Returns (xyz [N,3] float64, q [N] float64)."""
    if isinstance(cfg, str):
        cfg = CONFIGS[cfg]
    rng = np.random.default_rng(seed)
    n = cfg.n
    x = rng.uniform(0.0, cfg.Lx, n)
    y = rng.uniform(0.0, cfg.Ly, n)
    z = rng.uniform(0.5, cfg.Lz - 0.5, n)
    xyz = np.ascontiguousarray(np.column_stack([x, y, z]))
    if cfg.name == "slab_n100k":
        n_pos = n // 3                       # 33,333 x (+2)
        n_neg = 2 * n_pos                    # 66,666 x (-1)
        q = np.concatenate([np.full(n_pos, 2.0), np.full(n_neg, -1.0),
                            np.zeros(n - n_pos - n_neg)])
        q = q[rng.permutation(n)]
    elif cfg.name == "slab_n100":
        q = np.array([(-1.0) ** i for i in range(1, n + 1)])
    else:
        q = np.where(np.arange(n) % 2 == 0, 1.0, -1.0)
    return xyz, np.ascontiguousarray(q)


def scaled_config(cfg: SlabConfig, n: int) -> SlabConfig:
    """Same density, interfaces and parameters with n atoms (box rescaled in x, y)."""
    s = np.sqrt(n / cfg.n)
    return SlabConfig(f"{cfg.name}_n{n}", n, cfg.Lx * s, cfg.Ly * s, cfg.Lz, cfg.gamma_1,
                      cfg.gamma_2, cfg.alpha, cfg.r_cutoff, cfg.N_t, cfg.P, cfg.accuracy,
                      cfg.eps_0)


def rel_rms(f, ref) -> float:
    """||f - ref||_2 / ||ref||_2 over all components."""
    f = np.asarray(f, dtype=np.float64).ravel()
    ref = np.asarray(ref, dtype=np.float64).ravel()
    den = np.linalg.norm(ref)
    return float(np.linalg.norm(f - ref) / den) if den > 0 else float(np.linalg.norm(f - ref))
