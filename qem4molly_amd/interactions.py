"""Host-side mirror of QEM4Molly's interaction API, running on libqem_hip.so.

Same type names, constructor keywords, fields and `forces(inter, sys)`
entry point as the reference's Molly plugin (QEM4Molly, Julia):

    QEM_short(L, r_cutoff, alpha; iter_period, gamma_1, gamma_2, eps_0,
              accuracy, N_t)                      src/QEM_short.jl:4-25
    QEM_long(L; k_cutoff, z_list, gamma_1, gamma_2, eps_0, accuracy, alpha,
             rbe_mode, rbe_p)                     src/QEM_long.jl:3-30
    QEM_inter(n_atoms, L; s, accuracy, rbe_mode, rbe_p, eps_0, gamma_1,
              gamma_2, N_t, iter_period)          src/Quasi_Ewald_Method.jl:3-24
    forces(inter, sys, neighbors=None)            Molly.forces methods,
                                                  src/QEM_short.jl:27-59,
                                                  src/QEM_long.jl:32-44
    F_short_ij, F_short_i                         src/Force_short.jl:25-107
    K_set_generator                               src/K_set_generator.jl:5-45
    Fl_non_sort, Fl_sort, Flz_sort_k0, Flz_self,
    F_l_total, F_long                             src/Force_long.jl:5-242
    Substrate_LJ(; cutoff, sigma, eps, sub_pos, direction)
                                                  src/Substrate_LJ.jl:7-46

Every numeric result comes from the GPU library; a missing library raises.
Indices here are 0-based (Python), as are `z_list` values.

Additions beyond the reference keywords: `precision` ("fp64" | "mixed"),
`semantics` (QEM_SEM_* bits, 0 = as written), `device`, and on QEM_long a
`seed` for the batch sampler (the reference draws from Julia's global RNG,
src/Force_long.jl:235).  `rbe_mode=False` evaluates the full k sum over the
K set with weights exp(-k^2/(4 alpha)) (the reference's non-batch branch,
src/Force_long.jl:220-233, does not run as written).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Any, Sequence

import numpy as np

from . import _lib as L

_PREC = {"fp64": L.PREC_FP64, "mixed": L.PREC_MIXED,
         L.PREC_FP64: L.PREC_FP64, L.PREC_MIXED: L.PREC_MIXED}


@dataclass
class Atom:
    """The part of Molly.Atom the QEM interactions read (its charge)."""
    charge: float = 0.0
    mass: float = 1.0


@dataclass
class System:
    """Minimal stand-in for Molly.System: coordinates (N x 3) and atoms."""
    atoms: Sequence[Any]
    coords: Any

    def __post_init__(self):
        self.coords = np.ascontiguousarray(self.coords, dtype=np.float64).reshape(-1, 3)


def _charges(sys) -> np.ndarray:
    q = getattr(sys, "charges", None)
    if q is not None:
        return np.ascontiguousarray(q, dtype=np.float64)
    return np.array([a.charge for a in sys.atoms], dtype=np.float64)


def _params(L_, r_cutoff, alpha, gamma_1, gamma_2, eps_0, accuracy, N_t, iter_period,
            rbe_p, precision, semantics, device) -> L.Params:
    p = L.default_params()
    p.Lx, p.Ly, p.Lz = (float(v) for v in L_)
    p.r_cutoff = float(r_cutoff)
    p.alpha = float(alpha)
    p.gamma_1 = float(gamma_1)
    p.gamma_2 = float(gamma_2)
    p.eps_0 = float(eps_0)
    p.accuracy = float(accuracy)
    p.N_t = int(N_t)
    p.iter_period = int(iter_period)
    p.rbe_p = int(rbe_p)
    p.precision = _PREC[precision]
    p.semantics = int(semantics)
    p.device = int(device)
    return p


class QEM_short:
    """Short-range QEM interaction (src/QEM_short.jl:4-25)."""

    def __init__(self, L_, r_cutoff, alpha, *, iter_period=100, gamma_1=0.0, gamma_2=0.0,
                 eps_0=1.0, accuracy=1e-6, N_t=30, precision="fp64", semantics=0, device=0):
        self.r_cutoff = float(r_cutoff)
        self.n_steps = 0
        self.iter_period = int(iter_period)
        self.gamma_1 = float(gamma_1)
        self.gamma_2 = float(gamma_2)
        self.eps_0 = float(eps_0)
        self.L = tuple(float(v) for v in L_)
        self.accuracy = float(accuracy)
        self.N_t = int(N_t)
        self.alpha = float(alpha)
        self.precision = precision
        self.semantics = int(semantics)
        self.ctx = L.Context(_params(self.L, r_cutoff, alpha, gamma_1, gamma_2, eps_0, accuracy,
                                     N_t, iter_period, 0, precision, semantics, device))

    @property
    def neighbor_list(self):
        """Pair count of the device-resident list (the list itself stays on
        the GPU; `neighborlist()` returns a host copy)."""
        return self.ctx.short_state()[0]

    def neighborlist(self, coords) -> np.ndarray:
        return self.ctx.neighbor_list(coords)


class QEM_long:
    """Long-range (random batch) QEM interaction (src/QEM_long.jl:3-30)."""

    def __init__(self, L_, *, k_cutoff=1.0, z_list=(), gamma_1=0.0, gamma_2=0.0, eps_0=1.0,
                 accuracy=1e-6, alpha=1.0, rbe_mode=True, rbe_p=30, seed=0,
                 precision="fp64", device=0):
        self.k_cutoff = float(k_cutoff)
        self.z_list = np.asarray(z_list, dtype=np.int64)
        self.gamma_1 = float(gamma_1)
        self.gamma_2 = float(gamma_2)
        self.eps_0 = float(eps_0)
        self.L = tuple(float(v) for v in L_)
        self.accuracy = float(accuracy)
        self.alpha = float(alpha)
        self.rbe_mode = bool(rbe_mode)
        self.rbe_p = int(rbe_p)
        self.precision = precision
        self.seed = int(seed)
        self._draws = 0
        # r_cutoff is irrelevant for the long range; any value inside the box
        rc = 0.25 * min(self.L[0], self.L[1])
        self.ctx = L.Context(_params(self.L, rc, alpha, gamma_1, gamma_2, eps_0, accuracy, 1,
                                     1, max(self.rbe_p, 1), precision, 0, device))
        K, P, sK = self.ctx.kset()
        self.K_set, self.Prob, self.sum_K = K, P, sK

    def sample(self) -> np.ndarray:
        """Next batch of rbe_p k indices (seeded, one stream per interaction)."""
        idx = self.ctx.sample_k(self.seed + 0x9E3779B97F4A7C15 * self._draws, self.rbe_p)
        self._draws += 1
        return idx


_wall_ctx = None


def _utility_context() -> L.Context:
    """A small context for interactions that need only the device (walls)."""
    global _wall_ctx
    if _wall_ctx is None:
        p = L.default_params()
        p.rbe_p = 0
        _wall_ctx = L.Context(p)
    return _wall_ctx


class Substrate_LJ:
    """12-6 Lennard-Jones wall on the z coordinate (src/Substrate_LJ.jl:7-17).

    `semantics` (WALL_SEM_* bits of _lib) is an addition: 0 reproduces the
    reference as written (signed-dz cutoff, force -dU/dz / |dz|)."""

    def __init__(self, *, cutoff=1.0, sigma=0.5, eps=1.0, sub_pos=0.0, direction=+1,
                 semantics=0, ctx=None):
        self.cutoff = float(cutoff)
        self.sigma = float(sigma)
        self.eps = float(eps)
        self.sub_pos = float(sub_pos)
        self.direction = int(direction)
        self.semantics = int(semantics)
        self.ctx = ctx

    def wall(self) -> L.Wall:
        return L.wall(self.cutoff, self.sigma, self.eps, self.sub_pos, self.direction,
                      self.semantics)


def QEM_inter(n_atoms, L_, *, s=1.0, accuracy=1e-6, rbe_mode=True, rbe_p=30, eps_0=1.0,
              gamma_1=0.0, gamma_2=0.0, N_t=30, iter_period=100, precision="fp64", seed=0,
              device=0):
    """Factory (src/Quasi_Ewald_Method.jl:3-24): alpha, r_cutoff, k_cutoff from N and s."""
    alpha, r_cutoff, k_cutoff = L.inter_params(n_atoms, L_[0], L_[1], s, rbe_mode)
    short = QEM_short(L_, r_cutoff, alpha, iter_period=iter_period, gamma_1=gamma_1,
                      gamma_2=gamma_2, eps_0=eps_0, accuracy=accuracy, N_t=N_t,
                      precision=precision, device=device)
    long = QEM_long(L_, k_cutoff=k_cutoff, gamma_1=gamma_1, gamma_2=gamma_2, eps_0=eps_0,
                    accuracy=accuracy, alpha=alpha, rbe_mode=rbe_mode,
                    rbe_p=rbe_p if rbe_mode else 0, seed=seed, precision=precision,
                    device=device)
    return short, long


def forces(inter, sys, neighbors=None, *, k_idx=None) -> np.ndarray:
    """Molly.forces(inter, sys, neighbors): N x 3 forces of one interaction.

    QEM_short: step counter, neighbour list every iter_period calls
    (src/QEM_short.jl:27-59).  QEM_long: z order, then F_long with a batch
    from the interaction's sampler, or the given `k_idx`
    (src/QEM_long.jl:32-44).  Substrate_LJ: the wall's z forces
    (src/Substrate_LJ.jl:24-46)."""
    coords = np.ascontiguousarray(sys.coords, dtype=np.float64).reshape(-1, 3)
    q = _charges(sys)
    if isinstance(inter, QEM_short):
        f = inter.ctx.short_forces(coords, q)
        inter.n_steps += 1
        return f
    if isinstance(inter, QEM_long):
        inter.z_list = inter.ctx.sortperm_z(coords)
        if not inter.rbe_mode:
            return inter.ctx.long_terms(coords, q, inter.K_set, weight=-1.0, parts=L.LONG_ALL)
        if k_idx is None:
            k_idx = inter.sample()
        return inter.ctx.long_forces(coords, q, k_idx)
    if isinstance(inter, Substrate_LJ):
        # Molly.forces(::Substrate_LJ) (src/Substrate_LJ.jl:24-46); an atom on
        # the wrong side of the wall raises (the reference's @assert)
        ctx = inter.ctx if inter.ctx is not None else _utility_context()
        return ctx.substrate_lj(coords, inter.wall())
    raise TypeError(f"forces: unsupported interaction {type(inter).__name__}")


# ---- reference helper functions ------------------------------------------------

def F_short_ij(q_i, q_j, coord_i, coord_j, inter: QEM_short, *, single=False):
    """Force on i from j (src/Force_short.jl:25-78).  Returns a 3-vector."""
    fi, _ = inter.ctx.short_pair([q_i], [q_j], [coord_i], [coord_j], single=single)
    return fi[0]


def F_short_i(q_i, coord_i, inter: QEM_short, *, single=False):
    """Self-image force on i (src/Force_short.jl:82-107).  Returns a 3-vector."""
    return inter.ctx.short_self([q_i], [coord_i], single=single)[0]


def K_set_generator(L_x, L_y, alpha, accuracy):
    """(K, Prob, sum_K) of src/K_set_generator.jl:5-45 (K as an |K| x 3 array)."""
    rc = 0.25 * min(L_x, L_y)
    ctx = L.Context(_params((L_x, L_y, 1.0), rc, alpha, 0.0, 0.0, 1.0, accuracy, 1, 1, 1,
                            "fp64", 0, 0))
    try:
        return ctx.kset()
    finally:
        ctx.close()


@dataclass
class greens_element:
    """The fields of greens_element_init (src/greens_function.jl:51-65) that
    the long-range functions read."""
    gamma_1: float
    gamma_2: float
    L_z: float
    alpha: float = 1.0
    _ctx: Any = field(default=None, repr=False)

    def ctx(self):
        if self._ctx is None:
            self._ctx = L.Context(_params((1.0, 1.0, self.L_z), 0.25, self.alpha, self.gamma_1,
                                          self.gamma_2, 1.0, 1e-6, 1, 1, 1, "fp64", 0, 0))
        return self._ctx


def greens_element_init(gamma_1, gamma_2, L_z, alpha):
    return greens_element(float(gamma_1), float(gamma_2), float(L_z), float(alpha))


def _long_part(k_set, atoms, coords, element: greens_element, parts):
    # unit box area and eps_0, unit weight: the un-normalised per-k sums
    q = np.array([a.charge for a in atoms], dtype=np.float64) if not isinstance(
        atoms, np.ndarray) else np.asarray(atoms, dtype=np.float64)
    k = np.asarray(k_set, dtype=np.float64).reshape(1, 3)
    return element.ctx().long_terms(coords, q, k, weight=1.0, parts=parts)


def Fl_non_sort(k_set, atoms, coords, element):
    """src/Force_long.jl:5-70 (N x 3)."""
    return _long_part(k_set, atoms, coords, element, L.LONG_NONSORT)


def Fl_sort(k_set, atoms, z_list, coords, element):
    """src/Force_long.jl:72-127 (N x 3).  The z order is recomputed on the
    device (stable sortperm), so `z_list` must equal sortperm(z)."""
    return _long_part(k_set, atoms, coords, element, L.LONG_SORT)


def Flz_self(k_set, atoms, coords, element):
    """src/Force_long.jl:164-184 (N x 3, z only)."""
    return _long_part(k_set, atoms, coords, element, L.LONG_SELF)


def Flz_sort_k0(atoms, z_list, coords, element):
    """src/Force_long.jl:130-161: q_i (Q1 - Q2) (N x 3, z only)."""
    q = np.array([a.charge for a in atoms], dtype=np.float64)
    # the device applies 1/(2 Lx Ly eps_0) = 1/2 with a unit box
    f = element.ctx().long_terms(coords, q, np.zeros((0, 3)), weight=1.0, parts=L.LONG_K0)
    return 2.0 * f


def F_l_total(k_set, sys, inter: QEM_long, element=None):
    """src/Force_long.jl:186-199: one k, divided by Lx Ly eps_0."""
    coords = np.ascontiguousarray(sys.coords, dtype=np.float64).reshape(-1, 3)
    k = np.asarray(k_set, dtype=np.float64).reshape(1, 3)
    return inter.ctx.long_terms(coords, _charges(sys), k, weight=1.0,
                                parts=L.LONG_NONSORT | L.LONG_SORT | L.LONG_SELF)


def F_long(sys, inter: QEM_long, k_idx=None):
    """src/Force_long.jl:201-242 (random-batch branch)."""
    coords = np.ascontiguousarray(sys.coords, dtype=np.float64).reshape(-1, 3)
    if not inter.rbe_mode:
        return inter.ctx.long_terms(coords, _charges(sys), inter.K_set, weight=-1.0)
    if k_idx is None:
        k_idx = inter.sample()
    return inter.ctx.long_forces(coords, _charges(sys), k_idx)
