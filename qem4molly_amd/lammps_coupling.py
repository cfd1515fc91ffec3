"""LAMMPS coupling: the legacy python_lib driver's calling convention on the
GPU library.

Reference (read, not run; numba and LAMMPS are absent):
  input_read_lammps.py:14-81   PARCAR / INCAR loading and the derived fields
  init_lammps.py:43-57         POSCAR from the LAMMPS atoms (rows by atom id)
  update_lammps.py:55-107      update_Force_njit / update_velocity_njit: QEM
                               forces added to LAMMPS's per-atom f, or the
                               velocity kick v += F dt/m + q E dt/m
(all under src/python_lib/).  The driver works with any object that has the
LAMMPS python module's interface (`lammps.lammps`): get_natoms() and
extract_atom("x" | "f" | "v" | "id" | "type"), whose per-atom arrays index as
a[i][j] in LAMMPS's local order, atom ids 1-based.

The forces come from the library through the torch binding: positions and
charges go into device tensors that stay allocated across steps, the QEM
step (short range with its neighbour-list period, long range for a fresh
batch of RBM_p k vectors, or the full k sum in "exact" mode) runs on torch's
current stream, and the forces come back once per step.  The legacy
numerics (their own quadrature limits, Bessel approximations and K list)
are not reproduced: the physics is the library's, with the corrected
semantics by default (the legacy code integrates the gaussian part to its
own limit and takes the j-side z force from d/dz_j, as QEM_SEM_CORRECTED
does) and the pole-splitting rule with the INCAR's 'eps' for
gamma_1 gamma_2 > 1, as the legacy Force_short_norm path.
"""
from __future__ import annotations

import json
import math
from collections.abc import Mapping
from typing import Any

import numpy as np

from . import _lib as L


def load_parcar(path="./INPUT/PARCAR") -> dict:
    """PARCAR (input_read_lammps.py:14-19): JSON, type -> [mass, charge, count]."""
    with open(path) as fh:
        return json.load(fh)


def load_incar(parcar: dict, path="./INPUT/INCAR", incar: dict | None = None) -> dict:
    """INCAR with the derived fields of input_read_lammps.py:22-81 (the Gauss
    orders, K set and cell layout are the library's own and are not added)."""
    if incar is None:
        with open(path) as fh:
            incar = json.load(fh)
    incar = dict(incar)
    n = sum(int(parcar[str(t)][2]) for t in range(1, len(parcar) + 1))
    incar["NUM_particle"] = n
    g1, g2 = incar["g_1"], incar["g_2"]
    incar["eps_1"] = (1 - g1) / (1 + g1) if g1 != -1 else math.inf
    incar["eps_2"] = 1
    incar["eps_3"] = (1 - g2) / (1 + g2) if g2 != -1 else math.inf
    if incar["mode"] not in ("rbm", "exact"):
        raise ValueError(f"INCAR mode must be 'rbm' or 'exact', not {incar['mode']!r}")
    alpha = n / (incar["L_x"] * incar["L_y"])
    incar["alpha"] = alpha
    s = math.sqrt(-math.log(incar["accuracy"]))
    incar["r_c"] = s / math.sqrt(alpha)
    incar["k_f1"] = 2 * math.sqrt(alpha) * s
    incar["k_f2"] = s ** 2 / (2 * incar["L_z"])
    return incar


def _extract(lmp, name):
    """The per-atom array: numpy's view when the object offers one
    (lammps.numpy.extract_atom, or an array already), else the ctypes one."""
    num = getattr(lmp, "numpy", None)
    if num is not None and hasattr(num, "extract_atom"):
        return num.extract_atom(name)
    return lmp.extract_atom(name)


def _per_atom(lmp, name, n, width):
    a = _extract(lmp, name)
    if isinstance(a, np.ndarray):
        return np.array(a[:n, :width], dtype=np.float64)
    return np.array([[a[i][j] for j in range(width)] for i in range(n)], dtype=np.float64)


def _ids_types(lmp, n):
    ids = _extract(lmp, "id")
    types = _extract(lmp, "type")
    if isinstance(ids, np.ndarray) and isinstance(types, np.ndarray):
        return ids[:n].astype(np.int64), types[:n].astype(np.int64)
    return (np.array([int(ids[i]) for i in range(n)], dtype=np.int64),
            np.array([int(types[i]) for i in range(n)], dtype=np.int64))


def init_poscar(lmp, parcar: dict) -> np.ndarray:
    """POSCAR rows (q, x, y, z) by atom id (init_lammps.py:43-57)."""
    n = int(lmp.get_natoms())
    x = _per_atom(lmp, "x", n, 3)
    ids, types = _ids_types(lmp, n)
    poscar = np.zeros((n, 4))
    charge = np.zeros(int(types.max()) + 1 if n else 1)
    for t in np.unique(types):
        charge[t] = float(parcar[str(int(t))][1])
    poscar[ids - 1, 0] = charge[types]
    poscar[ids - 1, 1:] = x
    return poscar


def add_forces(lmp, f_by_id: np.ndarray) -> dict:
    """f[i] += F(id[i]) for every local atom (update_lammps.py:68-80); returns
    the reference's Force_dict (str(id) -> [Fx, Fy, Fz])."""
    n = int(lmp.get_natoms())
    ids, _ = _ids_types(lmp, n)
    f = _extract(lmp, "f")
    if isinstance(f, np.ndarray):
        f[:n, :3] += f_by_id[ids - 1]
    else:
        for i in range(n):
            row = f_by_id[ids[i] - 1]
            for j in range(3):
                f[i][j] += float(row[j])
    return _force_dict(f_by_id)


class ForceDict(Mapping):
    """The reference's Force_dict (str(id) -> [Fx, Fy, Fz]) as a read-only
    view of the forces by id, built without a per-atom loop."""

    def __init__(self, f_by_id: np.ndarray):
        self._f = f_by_id

    def __getitem__(self, key):
        i = int(key) - 1
        if not 0 <= i < self._f.shape[0]:
            raise KeyError(key)
        return self._f[i].tolist()

    def __iter__(self):
        return (str(i + 1) for i in range(self._f.shape[0]))

    def __len__(self):
        return self._f.shape[0]


def _force_dict(f_by_id):
    return ForceDict(f_by_id)


def kick_velocities(lmp, f_by_id: np.ndarray, parcar: dict, E, dt) -> dict:
    """v[i] += F(id[i]) dt/m + q E dt/m with m, q of the atom's type
    (update_lammps.py:94-107)."""
    n = int(lmp.get_natoms())
    ids, types = _ids_types(lmp, n)
    v = _extract(lmp, "v")
    if isinstance(v, np.ndarray):
        mt = np.zeros(int(types.max()) + 1 if n else 1)
        qt = np.zeros_like(mt)
        for t in np.unique(types):
            mt[t], qt[t] = float(parcar[str(int(t))][0]), float(parcar[str(int(t))][1])
        mass = mt[types][:, None]
        q = qt[types][:, None]
        Ev = np.asarray(E, dtype=np.float64)[None, :3]
        v[:n, :3] += f_by_id[ids - 1] * dt / mass + q * Ev * dt / mass
    else:
        for i in range(n):
            mass, q = float(parcar[str(types[i])][0]), float(parcar[str(types[i])][1])
            row = f_by_id[ids[i] - 1]
            for j in range(3):
                v[i][j] += float(row[j]) * dt / mass + q * float(E[j]) * dt / mass
    return _force_dict(f_by_id)


class QEMLammps:
    """The QEM force provider of a LAMMPS run (update_lammps.py).

    incar: load_incar's dict (needs L_x, L_y, L_z, g_1, g_2, accuracy, N_t,
    mode, RBM_p for mode "rbm", eps for gamma_1 gamma_2 > 1, optional eps_0);
    parcar: type -> [mass, charge, count]."""

    def __init__(self, incar: dict, parcar: dict, precision="mixed", semantics=None,
                 iter_period=100, seed=0, device=0):
        import torch  # noqa: F401  (the torch binding: one HIP runtime, torch's)
        from . import torch_qem
        self.torch = torch
        self.tq = torch_qem
        self.incar, self.parcar = incar, parcar
        p = L.default_params()
        p.Lx, p.Ly, p.Lz = incar["L_x"], incar["L_y"], incar["L_z"]
        p.r_cutoff, p.alpha = incar["r_c"], incar["alpha"]
        p.gamma_1, p.gamma_2 = incar["g_1"], incar["g_2"]
        p.eps_0 = incar.get("eps_0", 1.0)
        p.accuracy = incar["accuracy"]
        p.N_t = int(incar["N_t"])
        p.iter_period = int(iter_period)
        self.mode = incar["mode"]
        p.rbe_p = int(incar["RBM_p"]) if self.mode == "rbm" else 0
        p.precision = L.PREC_MIXED if precision == "mixed" else L.PREC_FP64
        sem = L.SEM_CORRECTED if semantics is None else int(semantics)
        if p.gamma_1 * p.gamma_2 > 1.0:
            sem |= L.SEM_POLE_SPLIT
            p.pole_eps = float(incar.get("eps", 0.0))
        p.semantics = sem
        p.device = device
        if self.mode == "exact":
            # the K set is built only for a batch size > 0; the full sum uses it
            p.rbe_p = 1
        self.ctx = L.Context(p)
        self.seed = int(seed)
        self.step = 0
        self.dev = torch.device("cuda", device)
        self._x = self._q = self._f = None

    def close(self):
        self.ctx.close()

    def forces_by_id(self, poscar: np.ndarray) -> np.ndarray:
        """QEM forces (N x 3, rows by atom id) for POSCAR rows (q, x, y, z)."""
        torch = self.torch
        n = poscar.shape[0]
        if self._x is None or self._x.shape[0] != n:
            self._x = torch.empty((n, 3), dtype=torch.float64, device=self.dev)
            self._q = torch.empty(n, dtype=torch.float64, device=self.dev)
            self._f = torch.empty((n, 3), dtype=torch.float64, device=self.dev)
        self._x.copy_(torch.from_numpy(np.ascontiguousarray(poscar[:, 1:4])))
        self._q.copy_(torch.from_numpy(np.ascontiguousarray(poscar[:, 0])))
        if self.mode == "rbm":
            kidx = self.ctx.sample_k(self.seed + self.step, self.ctx.params.rbe_p)
            self.tq.qem_forces(self.ctx, self._x, self._q, kidx, out=self._f)
        else:
            self.tq.qem_forces(self.ctx, self._x, self._q, full_sum=True, out=self._f)
        self.step += 1
        return self._f.cpu().numpy()

    def update_force(self, lmp) -> dict:
        """update_Force_njit (update_lammps.py:55-80)."""
        f = self.forces_by_id(init_poscar(lmp, self.parcar))
        return add_forces(lmp, f)

    def update_velocity(self, lmp) -> dict:
        """update_velocity_njit (update_lammps.py:82-107); INCAR E and dt."""
        f = self.forces_by_id(init_poscar(lmp, self.parcar))
        return kick_velocities(lmp, f, self.parcar, self.incar["E"], self.incar["dt"])

    def __enter__(self):
        return self

    def __exit__(self, *exc: Any):
        self.close()
