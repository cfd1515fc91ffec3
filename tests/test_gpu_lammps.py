"""GPU: the LAMMPS coupling (src/python_lib/update_lammps.py:55-107) over the
torch binding, with a stand-in LAMMPS object whose local order differs from
the atom ids, checked against the FP64 oracle.

The reference parameters are rebuilt here from the INCAR fields themselves
(input_read_lammps.py:22-81: alpha = N / (L_x L_y), r_c = sqrt(-ln acc) /
sqrt(alpha)), independently of the driver's own mapping, and the oracle
(short range over the O(N^2) pair list + long range) gives the reference
forces by atom id: the same sampled batch for "rbm" (the driver's seed and
step), the whole K set weighted by exp(-k^2/(4 alpha)) for "exact", and the
principal-value rule with the INCAR's eps for gamma_1 gamma_2 > 1.  The
forces the driver adds to LAMMPS's f must match within 1e-12 (FP64) and 1e-6
(mixed); the velocity kick v += F dt/m + q E dt/m is checked against a NumPy
restatement of update_lammps.py:94-107 with those oracle forces.

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
import json, math, sys
import numpy as np
import torch                                   # first: one HIP runtime
sys.path[:0] = [ROOT, ROOT + "/tests", ROOT + "/oracle"]
import oracle as O
from qem4molly_amd import lammps_coupling as LC
from lammps_fake import FakeLammps
from helpers import rel_rms
out = {}
PARCAR = {"1": [1.0, 1.0, 300], "2": [2.0, -1.0, 300]}
base = {"L_x": 60.0, "L_y": 60.0, "L_z": 8.0, "accuracy": 1e-6, "RBM_p": 50, "N_t": 30,
        "dt": 0.002, "E": [0.0, 0.0, 0.5]}
rng = np.random.default_rng(1)
n = 600
xyz = np.column_stack([rng.uniform(0, 60, n), rng.uniform(0, 60, n), rng.uniform(0.5, 7.5, n)])
types = np.array([1, 2] * 300)
q = np.where(types == 1, 1.0, -1.0)
mass = np.where(types == 1, 1.0, 2.0)
# the reference's derived fields (input_read_lammps.py:22-81), written out here
alpha = n / (60.0 * 60.0)
r_c = math.sqrt(-math.log(1e-6)) / math.sqrt(alpha)
pairs = O.neighbor_pairs(xyz, 60.0, 60.0, r_c)
K, Pk, sK, _ = O.kset(60.0, 60.0, alpha, 1e-6)
for name, extra, prec in [("rbm", {"mode": "rbm", "g_1": 0.7, "g_2": -0.5}, "mixed"),
                          ("rbm_fp64", {"mode": "rbm", "g_1": 0.7, "g_2": -0.5}, "fp64"),
                          ("exact", {"mode": "exact", "g_1": 0.7, "g_2": -0.5}, "fp64"),
                          ("pole", {"mode": "rbm", "g_1": 2.0, "g_2": 2.0, "eps": 0.005,
                                    "N_t": 120}, "fp64")]:
    inc = LC.load_incar(PARCAR, incar=dict(base, **extra))
    g1, g2, N_t = extra["g_1"], extra["g_2"], extra.get("N_t", 30)
    sem = O.SEM_CORRECTED | (O.SEM_POLE_SPLIT if g1 * g2 > 1 else 0)
    p = O.params((60.0, 60.0, 8.0), r_c, alpha, g1, g2, 1.0, 1e-6, N_t, sem,
                 pole_eps=extra.get("eps"))
    lmp = FakeLammps(xyz, types, seed=4)
    f0 = lmp.f.copy()
    with LC.QEMLammps(inc, PARCAR, precision=prec, seed=11) as drv:
        if inc["mode"] == "rbm":
            kv, w, full = K[drv.ctx.sample_k(11, 50)], sK / 50, False
        else:
            kv, w, full = K, 1.0, True
        ref = (O.short_forces(p, xyz, q, pairs, nthreads=16)
               + O.long_forces(p, xyz, q, kv, weight=w, full_sum=full, nthreads=16))
        drv.update_force(lmp)
        got = lmp.f - f0
        out[name] = rel_rms(got, ref[lmp.id - 1])
        out[name + "_finite"] = bool(np.all(np.isfinite(got)))
        v0 = lmp.v.copy()
        drv.step = 0                  # the same batch again
        drv.update_velocity(lmp)
        # update_lammps.py:94-107: v[i] += F(id) dt / m + q E dt / m
        want = v0.copy()
        for i in range(n):
            a = lmp.id[i] - 1
            for j in range(3):
                want[i, j] += ref[a, j] * 0.002 / mass[a] + q[a] * base["E"][j] * 0.002 / mass[a]
        out[name + "_v"] = float(np.max(np.abs(lmp.v - want)) / np.max(np.abs(want - v0)))
print("RESULT " + json.dumps(out), flush=True)
"""


def test_lammps_coupling_forces_vs_oracle():
    r = subprocess.run([sys.executable, "-u", "-c", CHILD.replace("ROOT", repr(ROOT))],
                       capture_output=True, text=True, timeout=900)
    print(r.stdout[-2000:], r.stderr[-4000:])
    assert r.returncode == 0
    line = [l for l in r.stdout.splitlines() if l.startswith("RESULT ")][-1]
    res = json.loads(line[len("RESULT "):])
    for name in ("rbm", "rbm_fp64", "exact", "pole"):
        tol = 1e-6 if name == "rbm" else 1e-12
        assert res[name + "_finite"]
        assert res[name] < tol, (name, res[name])
        # largest velocity error relative to the largest kick (a max over
        # components, so ten times the rel-RMS bar)
        assert res[name + "_v"] < 10 * tol, (name, res[name + "_v"])
