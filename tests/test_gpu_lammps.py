"""GPU: the LAMMPS coupling (src/python_lib/update_lammps.py:55-107) over the
torch binding, with a stand-in LAMMPS object whose local order differs from
the atom ids.  The forces it adds equal the library's forces for the
id-ordered POSCAR (same k batch), in "rbm" and "exact" mode, and with
gamma_1 gamma_2 > 1 (pole splitting with the INCAR's eps).

Runs in a child process that imports torch before the library, as the torch
binding requires; the rest of the suite never imports torch.
"""
import json
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CHILD = r"""
import json, sys
import numpy as np
import torch                                   # first: one HIP runtime
sys.path[:0] = [ROOT, ROOT + "/tests", ROOT + "/oracle"]
from qem4molly_amd import _lib as L
from qem4molly_amd import lammps_coupling as LC
from lammps_fake import FakeLammps
from helpers import rel_rms
out = {}
PARCAR = {"1": [1.0, 1.0, 300], "2": [2.0, -1.0, 300]}
base = {"L_x": 60.0, "L_y": 60.0, "L_z": 8.0, "accuracy": 1e-6, "RBM_p": 50, "N_t": 30,
        "dt": 0.002, "E": [0.0, 0.0, 0.5]}
rng = np.random.default_rng(1)
xyz = np.column_stack([rng.uniform(0, 60, 600), rng.uniform(0, 60, 600), rng.uniform(0.5, 7.5, 600)])
types = np.array([1, 2] * 300)
q = np.where(types == 1, 1.0, -1.0)
for name, extra, prec in [("rbm", {"mode": "rbm", "g_1": 0.7, "g_2": -0.5}, "mixed"),
                          ("rbm_fp64", {"mode": "rbm", "g_1": 0.7, "g_2": -0.5}, "fp64"),
                          ("exact", {"mode": "exact", "g_1": 0.7, "g_2": -0.5}, "fp64"),
                          ("pole", {"mode": "rbm", "g_1": 2.0, "g_2": 2.0, "eps": 0.005,
                                    "N_t": 120}, "fp64")]:
    inc = LC.load_incar(PARCAR, incar=dict(base, **extra))
    lmp = FakeLammps(xyz, types, seed=4)
    f0 = lmp.f.copy()
    with LC.QEMLammps(inc, PARCAR, precision=prec, seed=11) as drv:
        drv.update_force(lmp)
        p = drv.ctx.params
        with L.Context(p) as c:
            if inc["mode"] == "rbm":
                ref = c.forces(xyz, q, c.sample_k(11, inc["RBM_p"]))
            else:
                K, _, _ = c.kset()
                ref = c.short_forces(xyz, q) + c.long_terms(xyz, q, K, weight=-1.0)
        got = lmp.f - f0
        out[name] = rel_rms(got, ref[lmp.id - 1])
        out[name + "_finite"] = bool(np.all(np.isfinite(got)))
        v0 = lmp.v.copy()
        drv.step = 0                  # the same batch again
        drv.update_velocity(lmp)
        m = np.where(lmp.type == 1, 1.0, 2.0)[:, None]
        qq = np.where(lmp.type == 1, 1.0, -1.0)[:, None]
        want = v0 + ref[lmp.id - 1] * 0.002 / m + qq * np.array([[0.0, 0.0, 0.5]]) * 0.002 / m
        out[name + "_v"] = float(np.max(np.abs(lmp.v - want)))
print("RESULT " + json.dumps(out), flush=True)
"""


def test_lammps_coupling_forces():
    r = subprocess.run([sys.executable, "-u", "-c", CHILD.replace("ROOT", repr(ROOT))],
                       capture_output=True, text=True, timeout=600)
    print(r.stdout[-2000:], r.stderr[-4000:])
    assert r.returncode == 0
    line = [l for l in r.stdout.splitlines() if l.startswith("RESULT ")][-1]
    res = json.loads(line[len("RESULT "):])
    for name in ("rbm", "rbm_fp64", "exact", "pole"):
        assert res[name + "_finite"]
        # same kernels, same batch: only the accumulation into LAMMPS's f
        # (f0 + F - f0) rounds
        assert res[name] < 1e-12, (name, res[name])
        assert res[name + "_v"] < 1e-9, (name, res[name + "_v"])
