"""Shared helpers for the test suite (oracle parameters, contexts, metrics)."""
from __future__ import annotations

import numpy as np

import oracle as O
from qem4molly_amd import _lib as L
from qem4molly_amd.workloads import CONFIGS, SlabConfig, make_system, rel_rms, scaled_config

__all__ = ["O", "L", "CONFIGS", "SlabConfig", "make_system", "rel_rms", "scaled_config",
           "oracle_params", "lib_params", "ctx_for", "four_atom", "two_atom", "max_rel"]


def oracle_params(cfg: SlabConfig, semantics=0, N_t=None) -> O.Params:
    return O.params(cfg.L, cfg.r_cutoff, cfg.alpha, cfg.gamma_1, cfg.gamma_2, cfg.eps_0,
                    cfg.accuracy, N_t or cfg.N_t, semantics)


def lib_params(cfg: SlabConfig, precision="fp64", semantics=0, N_t=None, iter_period=100,
               rbe_p=None) -> L.Params:
    p = L.default_params()
    p.Lx, p.Ly, p.Lz = cfg.L
    p.r_cutoff = cfg.r_cutoff
    p.alpha = cfg.alpha
    p.gamma_1 = cfg.gamma_1
    p.gamma_2 = cfg.gamma_2
    p.eps_0 = cfg.eps_0
    p.accuracy = cfg.accuracy
    p.N_t = N_t or cfg.N_t
    p.iter_period = iter_period
    p.rbe_p = cfg.P if rbe_p is None else rbe_p
    p.precision = L.PREC_MIXED if precision == "mixed" else L.PREC_FP64
    p.semantics = semantics
    p.device = 0
    return p


def ctx_for(cfg: SlabConfig, **kw) -> L.Context:
    return L.Context(lib_params(cfg, **kw))


def max_rel(a, b):
    a = np.asarray(a, dtype=np.float64); b = np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


# geometries held by the reference's own tests
def two_atom():
    """test/Force_short_test.jl:30-34: q_i = q_j = 1, r_i, r_j, L = (10, 10, 2)."""
    return (np.array([5.0, 5.0, 0.2]), np.array([5.1, 5.2, 0.15]), (10.0, 10.0, 2.0))


def four_atom():
    """test/Force_long_test.jl:13-14,36-37: the 4-atom slab, L = (10, 10, 2),
    charges (-1)^(i+1), z_list = [2, 1, 3, 4] (1-based)."""
    xyz = np.array([[5.0, 5.0, 0.2], [5.1, 5.2, 0.15], [4.9, 4.9, 0.3], [5.5, 5.3, 0.4]])
    q = np.array([1.0, -1.0, 1.0, -1.0])
    return xyz, q, (10.0, 10.0, 2.0), np.array([1, 0, 2, 3])
