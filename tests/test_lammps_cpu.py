"""The LAMMPS coupling's host side (src/python_lib/input_read_lammps.py:14-81,
init_lammps.py:43-57, update_lammps.py:55-107) on the CPU: INCAR derivation,
POSCAR rows by atom id, and the force and velocity updates in LAMMPS's local
order."""
import json
import math

import numpy as np
import pytest

from lammps_fake import FakeLammps
from qem4molly_amd import lammps_coupling as LC

PARCAR = {"1": [1.0, 1.0, 30], "2": [2.5, -2.0, 15]}
INCAR = {"L_x": 40.0, "L_y": 30.0, "L_z": 6.0, "g_1": 0.6, "g_2": -0.4, "accuracy": 1e-5,
         "mode": "rbm", "RBM_p": 20, "N_t": 30, "dt": 0.005, "E": [0.1, -0.2, 0.3]}


def test_load_parcar_and_incar(tmp_path):
    (tmp_path / "PARCAR").write_text(json.dumps(PARCAR))
    (tmp_path / "INCAR").write_text(json.dumps(INCAR))
    par = LC.load_parcar(tmp_path / "PARCAR")
    assert par == PARCAR
    inc = LC.load_incar(par, tmp_path / "INCAR")
    n = 45
    alpha = n / (40.0 * 30.0)
    s = math.sqrt(-math.log(1e-5))
    assert inc["NUM_particle"] == n and inc["alpha"] == alpha
    assert inc["r_c"] == s / math.sqrt(alpha)
    assert inc["k_f1"] == 2 * math.sqrt(alpha) * s and inc["k_f2"] == s ** 2 / 12.0
    assert inc["eps_1"] == pytest.approx(0.4 / 1.6) and inc["eps_3"] == pytest.approx(1.4 / 0.6)
    assert inc["eps_2"] == 1
    with pytest.raises(ValueError):
        LC.load_incar(par, incar=dict(INCAR, mode="other"))


def _system(seed=0):
    rng = np.random.default_rng(seed)
    xyz = np.column_stack([rng.uniform(0, 40, 45), rng.uniform(0, 30, 45), rng.uniform(0.5, 5.5, 45)])
    types = np.array([1] * 30 + [2] * 15)
    return xyz, types


@pytest.mark.parametrize("lists", [False, True])
def test_poscar_rows_follow_atom_ids(lists):
    xyz, types = _system()
    lmp = FakeLammps(xyz, types, lists=lists)
    pos = LC.init_poscar(lmp, PARCAR)
    np.testing.assert_array_equal(pos[:, 1:], xyz)
    np.testing.assert_array_equal(pos[:, 0], np.where(types == 1, 1.0, -2.0))


@pytest.mark.parametrize("lists", [False, True])
def test_force_and_velocity_updates_in_local_order(lists):
    xyz, types = _system(1)
    lmp = FakeLammps(xyz, types, seed=3, lists=lists)
    f0, v0 = lmp.f.copy(), lmp.v.copy()
    F = np.random.default_rng(9).normal(size=(45, 3))      # rows by id
    d = LC.add_forces(lmp, F)
    idx = lmp.id - 1
    np.testing.assert_array_equal(lmp.f, f0 + F[idx])
    assert d["7"] == list(F[6])
    LC.kick_velocities(lmp, F, PARCAR, INCAR["E"], INCAR["dt"])
    m = np.where(lmp.type == 1, 1.0, 2.5)[:, None]
    q = np.where(lmp.type == 1, 1.0, -2.0)[:, None]
    E = np.array(INCAR["E"])[None, :]
    np.testing.assert_allclose(lmp.v, v0 + F[idx] * 0.005 / m + q * E * 0.005 / m, rtol=1e-15)
