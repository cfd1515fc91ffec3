"""A stand-in for the LAMMPS python module's `lammps` object (the module is
not installed): get_natoms() and extract_atom() over numpy per-atom arrays in
a local order that differs from the atom ids."""
import numpy as np


class FakeLammps:
    """lists=True hands out per-atom arrays the way the plain module does
    (indexable as a[i][j], not numpy arrays; the rows write through)."""

    def __init__(self, xyz_by_id, types_by_id, seed=0, lists=False):
        self.lists = lists
        n = xyz_by_id.shape[0]
        order = np.random.default_rng(seed).permutation(n)   # local index -> id - 1
        self.id = (order + 1).astype(np.int64)
        self.type = np.asarray(types_by_id)[order].astype(np.int64)
        self.x = np.ascontiguousarray(xyz_by_id[order], dtype=np.float64)
        self.f = np.random.default_rng(seed + 1).normal(size=(n, 3))
        self.v = np.random.default_rng(seed + 2).normal(size=(n, 3))

    def get_natoms(self):
        return int(self.id.shape[0])

    def extract_atom(self, name):
        a = getattr(self, name)
        if not self.lists:
            return a
        return [row for row in a] if a.ndim == 2 else [int(v) for v in a]
