"""GPU: the adjoint at scale -- slab_n100k (N = 100,000), one fixed seeded
batch of P = 400 k vectors, mixed precision, corrected semantics (the
conservative force), positions perturbed over consecutive steps as in the
reference's multi-step run (test/QEM_test.jl:53-58): each step N(0, 0.01)
per component, x and y wrapped, z reflected into [0.5, Lz - 0.5].

Every step runs the forward (forces, energy, dE/dq, dE/dgamma) and the
backward (force Hessian-vector product along a fresh gaussian vector).  At
the first and the last step the mixed results are checked against an FP64
context on the same inputs (forces, HVP, E and dE/dq within 1e-6; dE/dgamma,
a sum of FP32 node sums over every pair, within 1e-5, the bound of the
small-system mixed-precision energy test);
every step's outputs are finite, the last step's are bitwise repeatable,
the one-call forces + energy (qem_forces_energy_dev) gives the same forces
bit for bit and the energy within the same bars, and the context's device
memory stays under 1.6 GB (measured 1.49 GB: the endpoint records take
32 bytes per pair, 0.5 GB at 15.7 million pairs).  At the end the same
call keeping the tangent node sums (qem_forces_energy_save_dev) repeats
its outputs bit for bit, the HVP from those sums (qem_hvp_saved_dev) is
within 1e-6 of FP64, and the sums add at most 1.25 x 224 bytes per pair.
"""
import numpy as np
import pytest

from helpers import CONFIGS, L, ctx_for, make_system, rel_rms

pytestmark = pytest.mark.gpu

STEPS = 10
SIGMA = 0.01
MAX_DEVICE_BYTES = 1.6e9


def _perturb(xyz, rng, cfg):
    xyz = xyz + rng.normal(scale=SIGMA, size=xyz.shape)
    xyz[:, 0] %= cfg.Lx
    xyz[:, 1] %= cfg.Ly
    lo, hi = 0.5, cfg.Lz - 0.5
    z = xyz[:, 2]
    z = np.where(z < lo, 2 * lo - z, z)
    xyz[:, 2] = np.where(z > hi, 2 * hi - z, z)
    return xyz


def _step(c, n, xd, qd, vd, fd, hd, pd, kidx):
    c.forces_dev(n, xd.ptr, qd.ptr, kidx, fd.ptr)
    E, dg = c.energy_dev(n, xd.ptr, qd.ptr, kidx, pd.ptr)   # synchronises
    c.hvp_dev(n, xd.ptr, qd.ptr, kidx, vd.ptr, hd.ptr)
    L.synchronize()
    return fd.download(), hd.download(), E, pd.download(), np.array(dg)


def test_adjoint_slab_n100k_perturbed_steps():
    cfg = CONFIGS["slab_n100k"]
    xyz, q = make_system(cfg)
    rng = np.random.default_rng(5)
    xd, qd = L.DeviceArray(xyz), L.DeviceArray(q)
    fd, hd, vd = (L.DeviceArray(np.zeros_like(xyz)) for _ in range(3))
    pd = L.DeviceArray(np.zeros(cfg.n))
    c = ctx_for(cfg, precision="mixed", semantics=L.SEM_CORRECTED)
    c64 = ctx_for(cfg, precision="fp64", semantics=L.SEM_CORRECTED)
    with c, c64:
        kidx = c.sample_k(0, cfg.P)
        for s in range(STEPS):
            xyz = _perturb(xyz, rng, cfg)
            v = rng.normal(size=xyz.shape)
            xd.upload(xyz)
            vd.upload(v)
            F, hv, E, phi, dg = _step(c, cfg.n, xd, qd, vd, fd, hd, pd, kidx)
            print(f"step {s}: E {E:.9e} dE/dgamma {dg}", flush=True)
            for a in (F, hv, phi, dg):
                assert np.all(np.isfinite(a))
            assert np.isfinite(E)
            if s in (0, STEPS - 1):
                F64, hv64, E64, phi64, dg64 = _step(c64, cfg.n, xd, qd, vd, fd, hd, pd, kidx)
                eF, eH, eP = rel_rms(F, F64), rel_rms(hv, hv64), rel_rms(phi, phi64)
                print(f"step {s} vs fp64: F {eF:.2e} hv {eH:.2e} E {abs(E - E64) / abs(E64):.2e}"
                      f" phi {eP:.2e} dg {np.abs(dg - dg64) / np.abs(dg64)}", flush=True)
                assert eF < 1e-6 and eH < 1e-6 and eP < 1e-6
                assert abs(E - E64) <= 1e-6 * abs(E64)
                assert np.allclose(dg, dg64, rtol=1e-5, atol=1e-9 * abs(E64))
        # the last step again: bitwise the same outputs
        F2, hv2, E2, phi2, dg2 = _step(c, cfg.n, xd, qd, vd, fd, hd, pd, kidx)
        assert np.array_equal(F2, F) and np.array_equal(hv2, hv) and E2 == E
        assert np.array_equal(phi2, phi) and np.array_equal(dg2, dg)
        # forces and energy in one call (one pair quadrature for both): the
        # same forces bit for bit, the energy within the mixed bars of FP64
        E3, dg3 = c.forces_energy_dev(cfg.n, xd.ptr, qd.ptr, kidx, fd.ptr, pd.ptr)
        L.synchronize()
        F3, phi3, dg3 = fd.download(), pd.download(), np.array(dg3)
        print(f"fused vs fp64: E {abs(E3 - E64) / abs(E64):.2e} phi {rel_rms(phi3, phi64):.2e}"
              f" dg {np.abs(dg3 - dg64) / np.abs(dg64)}", flush=True)
        assert np.array_equal(F3, F)
        assert abs(E3 - E64) <= 1e-6 * abs(E64) and rel_rms(phi3, phi64) < 1e-6
        assert np.allclose(dg3, dg64, rtol=1e-5, atol=1e-9 * abs(E64))
        mem = c.device_bytes()
        print(f"peak device bytes (mixed context): {mem}", flush=True)
        assert mem < MAX_DEVICE_BYTES
        # the same call keeping the tangent node sums: its outputs bit for bit;
        # the HVP from the sums within the mixed bar of FP64 (FP32 sums in
        # another order than the direct HVP's, so not bitwise the same)
        E4, dg4 = c.forces_energy_dev(cfg.n, xd.ptr, qd.ptr, kidx, fd.ptr, pd.ptr, save=True)
        L.synchronize()
        assert np.array_equal(fd.download(), F3) and E4 == E3
        assert np.array_equal(pd.download(), phi3) and np.array_equal(np.array(dg4), dg3)
        c.hvp_dev(cfg.n, xd.ptr, qd.ptr, kidx, vd.ptr, hd.ptr, saved=True)
        L.synchronize()
        hs = hd.download()
        print(f"saved-sums HVP vs fp64: {rel_rms(hs, hv64):.2e}, vs direct {rel_rms(hs, hv):.2e}",
              flush=True)
        assert rel_rms(hs, hv64) < 1e-6 and not np.array_equal(hs, hv)
        # the sums: 224 bytes per pair, allocated with 25% headroom
        npairs = c.short_state()[0]
        mem_saved = c.device_bytes()
        print(f"device bytes with the saved sums: {mem_saved} ({npairs} pairs)", flush=True)
        assert mem_saved <= mem + 1.25 * 224 * npairs + 4096
