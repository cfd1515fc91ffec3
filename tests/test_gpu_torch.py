"""GPU: the torch.autograd binding over the device-pointer entry points.

Runs in a child process that imports torch before the library, so torch and
libqem_hip.so share torch's HIP runtime (the order the binding requires);
the rest of the suite never imports torch.
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
from qem4molly_amd.torch_qem import qem_energy_forces
from helpers import CONFIGS, L, ctx_for, make_system, rel_rms, scaled_config
out = {}
cfg = scaled_config(CONFIGS["slab_n10k"], 800)
xyz_h, q_h = make_system(cfg, seed=2)
dev = torch.device("cuda:0")
with ctx_for(cfg, semantics=L.SEM_CORRECTED, iter_period=100000) as c:
    kidx = c.sample_k(3, cfg.P)
    F_ref = c.forces(xyz_h, q_h, kidx)
    E_ref, phi_ref, dg_ref = c.energy(xyz_h, q_h, kidx)
    v = np.random.default_rng(0).normal(size=xyz_h.shape)
    hv_ref = c.hvp(xyz_h, q_h, kidx, v)
    xyz = torch.tensor(xyz_h, device=dev, requires_grad=True)
    q = torch.tensor(q_h, device=dev, requires_grad=True)
    gamma = torch.tensor([cfg.gamma_1, cfg.gamma_2], dtype=torch.float64, device=dev,
                         requires_grad=True)
    E, F = qem_energy_forces(c, xyz, q, kidx, gamma=gamma)
    out["E"] = abs(float(E) - E_ref) / abs(E_ref)
    out["F"] = rel_rms(F.detach().cpu().numpy(), F_ref)
    E.backward()
    out["gx"] = rel_rms(xyz.grad.cpu().numpy(), -F_ref)
    out["gq"] = rel_rms(q.grad.cpu().numpy(), phi_ref)
    out["gg"] = float(np.max(np.abs(gamma.grad.cpu().numpy() - dg_ref) / np.abs(dg_ref)))
    xyz2 = torch.tensor(xyz_h, device=dev, requires_grad=True)
    E2, F2 = qem_energy_forces(c, xyz2, torch.tensor(q_h, device=dev), kidx)
    (F2 * torch.tensor(v, device=dev)).sum().backward()
    out["hv"] = rel_rms(xyz2.grad.cpu().numpy(), -hv_ref)
    # a loss on F differentiated with respect to gamma (or q) raises instead
    # of returning the energy term alone
    for wrt in ("gamma", "q"):
        g3 = torch.tensor([cfg.gamma_1, cfg.gamma_2], dtype=torch.float64, device=dev,
                          requires_grad=(wrt == "gamma"))
        q3 = torch.tensor(q_h, device=dev, requires_grad=(wrt == "q"))
        E3, F3 = qem_energy_forces(c, torch.tensor(xyz_h, device=dev), q3, kidx, gamma=g3)
        try:
            (E3 + F3.sum()).backward()
            out["raise_" + wrt] = False
        except NotImplementedError:
            out["raise_" + wrt] = True
    # torch-allocated device memory through the plain device entry point
    x = torch.tensor(xyz_h, device=dev); qq = torch.tensor(q_h, device=dev)
    f = torch.empty_like(x)
    c.forces_dev(cfg.n, x.data_ptr(), qq.data_ptr(), kidx, f.data_ptr(),
                 torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    out["dev"] = rel_rms(f.cpu().numpy(), c.forces(xyz_h, q_h, kidx))
# mixed precision: the forward keeps the tangent sums and the backward's HVP
# combines them; a second forward on the context before the first backward
# replaces the sums, and that backward then recomputes the quadrature
with ctx_for(cfg, precision="mixed", semantics=L.SEM_CORRECTED, iter_period=100000) as c:
    kidx = c.sample_k(3, cfg.P)
    v = np.random.default_rng(0).normal(size=xyz_h.shape)
    c.forces_energy(xyz_h, q_h, kidx, save=True)
    hs_ref = c.hvp(xyz_h, q_h, kidx, v, saved=True)
    hd_ref = c.hvp(xyz_h, q_h, kidx, v)
    vt = torch.tensor(v, device=dev)
    xa = torch.tensor(xyz_h, device=dev, requires_grad=True)
    Ea, Fa = qem_energy_forces(c, xa, torch.tensor(q_h, device=dev), kidx)
    (Fa * vt).sum().backward()
    out["mixed_saved_exact"] = bool(np.array_equal(xa.grad.cpu().numpy(), -hs_ref))
    xb = torch.tensor(xyz_h, device=dev, requires_grad=True)
    Eb, Fb = qem_energy_forces(c, xb, torch.tensor(q_h, device=dev), kidx, keep_sums=False)
    (Fb * vt).sum().backward()
    out["mixed_direct_exact"] = bool(np.array_equal(xb.grad.cpu().numpy(), -hd_ref))
    xc = torch.tensor(xyz_h, device=dev, requires_grad=True)
    xd = torch.tensor(xyz_h + 0.01, device=dev, requires_grad=True)
    Ec, Fc = qem_energy_forces(c, xc, torch.tensor(q_h, device=dev), kidx)
    Ed, Fd = qem_energy_forces(c, xd, torch.tensor(q_h, device=dev), kidx)
    (Fc * vt).sum().backward()
    out["mixed_replaced_exact"] = bool(np.array_equal(xc.grad.cpu().numpy(), -hd_ref))
    out["mixed_saved_vs_direct"] = rel_rms(hs_ref, hd_ref)
print("RESULT " + json.dumps(out), flush=True)
"""


def test_torch_autograd_binding():
    r = subprocess.run([sys.executable, "-u", "-c", CHILD.replace("ROOT", repr(ROOT))],
                       capture_output=True, text=True, timeout=600)
    print(r.stdout[-2000:], r.stderr[-4000:])
    assert r.returncode == 0
    line = [l for l in r.stdout.splitlines() if l.startswith("RESULT ")][-1]
    res = json.loads(line[len("RESULT "):])
    assert res["E"] < 1e-13 and res["F"] < 1e-13
    assert res["gx"] < 1e-13 and res["gq"] < 1e-13 and res["gg"] < 1e-12
    assert res["hv"] < 1e-13
    assert res["dev"] < 1e-13
    assert res["raise_gamma"] and res["raise_q"]
    assert res["mixed_saved_exact"] and res["mixed_direct_exact"]
    assert res["mixed_replaced_exact"]
    assert res["mixed_saved_vs_direct"] < 1e-6
