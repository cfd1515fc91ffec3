#!/usr/bin/env python3
"""QEM force-path throughput on one MI355X (libqem_hip.so, gfx950).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload slab_n100k]

A step is one full electrostatic force evaluation of the quasi-Ewald method
(short range: neighbour list at its iter_period=100 frequency + pair
quadrature + self term + gather; long range: z sort + random batch of P
k-vectors + k = 0 term) on device-resident synthetic inputs, with a fresh
k batch drawn per step.  Prints ONE JSON line (other output goes to stderr).

Metric: atom-force evals/s = N / t_step, mixed precision (FP32 quadrature,
Bessel, exp, sincos; FP64 accumulation), next to the FP64 plain GPU path on
the same inputs and k batches, with the mixed result's rel-RMS error against
the FP64 path.  The workload is slab_n100k (N = 100,000), the size at which
the mixed-vs-FP64 target is stated; slab_n10k is reported as a secondary
figure.  roofline: the pair-quadrature kernel (the dominant kernel) against
the FP32 VALU peak, its duration measured with HIP events on the stream it
runs on.  cpu_baseline: the C FP64 restatement of the reference algorithm
(oracle/) on the host cores, timed on a bounded sample and extrapolated.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# algorithmic flop count per unit (add/mul 1, FMA 2, exp/sin/cos/rcp/sqrt 1
# plus argument arithmetic, joint J0+J1 evaluation 40)
FLOP_NODE_POINT = 62      # one point-charge quadrature node (rho and z integrands)
FLOP_NODE_GAUSS = 64      # one gaussian-screened quadrature node
FLOP_NODE_SELF_POINT = 22
FLOP_NODE_SELF_GAUSS = 24
FLOP_ATOM_K = 90          # long range per (atom, k): structure factor 20 + force 70
PEAK_FP32_TFLOPS = 157.3  # MI355X published vector FP32
PEAK_FP64_TFLOPS = 78.6   # MI355X published vector FP64
PEAK_HBM_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def pair_flops(npairs: int, N_t: int) -> float:
    return float(npairs) * 4 * N_t * (FLOP_NODE_POINT + FLOP_NODE_GAUSS)


def run_precision(cfg, xyz_d, q_d, f_d, precision, steps, warmup, seed0=1000):
    from qem4molly_amd import _lib as L
    p = L.default_params()
    p.Lx, p.Ly, p.Lz = cfg.L
    p.r_cutoff, p.alpha = cfg.r_cutoff, cfg.alpha
    p.gamma_1, p.gamma_2 = cfg.gamma_1, cfg.gamma_2
    p.eps_0, p.accuracy = cfg.eps_0, cfg.accuracy
    p.N_t, p.iter_period, p.rbe_p = cfg.N_t, 100, cfg.P
    p.precision = L.PREC_MIXED if precision == "mixed" else L.PREC_FP64
    ctx = L.Context(p)
    step = 0

    def one():
        nonlocal step
        kidx = ctx.sample_k(seed0 + step, cfg.P)
        ctx.forces_dev(cfg.n, xyz_d.ptr, q_d.ptr, kidx, f_d.ptr)
        step += 1

    for _ in range(warmup):
        one()
    L.synchronize()
    b0 = ctx.short_state()[1]
    t0 = time.perf_counter()
    for _ in range(steps):
        one()
    L.synchronize()
    dt = time.perf_counter() - t0
    npairs, builds, nsteps = ctx.short_state()
    builds_timed = builds - b0
    f = f_d.download()
    kidx_last = ctx.sample_k(seed0 + step - 1, cfg.P)
    # The neighbour list is rebuilt at steps 1, 101, 201, ... (iter_period =
    # 100, src/QEM_short.jl:32-36).  Two more windows after the timed steps:
    # 20 steps holding no rebuild, then 100 consecutive steps holding exactly
    # one (its cost amortised at its frequency).
    while ctx.short_state()[2] % 100 != 1:
        one()
    L.synchronize()
    b1 = ctx.short_state()[1]
    t0 = time.perf_counter()
    for _ in range(20):
        one()
    L.synchronize()
    dt_none = (time.perf_counter() - t0) / 20
    while ctx.short_state()[2] % 100 != 0:
        one()
    L.synchronize()
    b2 = ctx.short_state()[1]
    t0 = time.perf_counter()
    for _ in range(100):
        one()
    L.synchronize()
    dt_amort = (time.perf_counter() - t0) / 100
    windows_ok = ctx.short_state()[1] - b2 == 1 and b2 == b1
    # Per-part times (HIP events, after the timed steps, which run without
    # them).  The step runs the long range on a second stream beside the
    # short range, so each part is timed alone, on the same inputs: the short
    # range over 100 steps (one neighbour-list rebuild among them, as in the
    # step), then the long range over 20 k batches.
    g = L.DeviceArray(np.zeros(3 * cfg.n))
    ctx.timing_enable(True)
    for _ in range(100):
        ctx.short_forces_dev(cfg.n, xyz_d.ptr, q_d.ptr, g.ptr)
    L.synchronize()
    short = ctx.timing_read(reset=True)
    for t in range(20):
        ctx.long_forces_dev(cfg.n, xyz_d.ptr, q_d.ptr, ctx.sample_k(seed0 + t, cfg.P), g.ptr)
    L.synchronize()
    long_ = ctx.timing_read(reset=True)
    # part -> (ms per step, ms per launch of its events)
    phases = {k: (short[k][0] / 100, short[k][0] / max(short[k][1], 1))
              for k in ("neighbor", "pair", "gather")}
    phases["long"] = (long_["long"][0] / 20, long_["long"][0] / max(long_["long"][1], 1))
    del g
    ctx.close()
    return {"ms_per_step": dt / steps * 1e3, "phases": phases, "npairs": npairs,
            "builds": builds, "builds_timed": builds_timed, "f": f, "kidx_last": kidx_last,
            "ms_per_step_rebuild_amortized": dt_amort * 1e3,
            "ms_per_step_no_rebuild": dt_none * 1e3, "rebuild_windows_ok": windows_ok}


def parts_per_step(run):
    """ms per step of each part, timed alone."""
    return {k: v[0] for k, v in run["phases"].items()}


def cpu_baseline(cfg, xyz, q, pairs_dev_ctx_pairs, kvecs, weight, threads, target_s=15.0):
    """Times the oracle on a bounded sample (about target_s seconds) and
    extrapolates to one step."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    lib_path = O.use_avx512_build()
    p = O.params(cfg.L, cfg.r_cutoff, cfg.alpha, cfg.gamma_1, cfg.gamma_2, cfg.eps_0,
                 cfg.accuracy, cfg.N_t, 0)
    pairs = pairs_dev_ctx_pairs
    rng = np.random.default_rng(0)
    # calibrate on a small sample (large enough to amortise the thread start),
    # then size the timed sample to ~0.6 target_s
    m0 = min(len(pairs), 2000 * threads)
    sel = rng.choice(len(pairs), m0, replace=False)
    t = time.perf_counter()
    O.short_pair_forces(p, xyz, q, pairs[sel], nthreads=threads)
    per_pair = (time.perf_counter() - t) / m0
    m = int(min(len(pairs), max(m0, 0.6 * target_s / max(per_pair, 1e-9))))
    sel = rng.choice(len(pairs), m, replace=False)
    t = time.perf_counter()
    O.short_pair_forces(p, xyz, q, pairs[sel], nthreads=threads)
    t_pairs = (time.perf_counter() - t) * len(pairs) / m
    # self term on a subset of atoms
    ns = min(cfg.n, 20000)
    idx = rng.choice(cfg.n, ns, replace=False)
    t = time.perf_counter()
    O.short_forces(p, xyz[idx], q[idx], np.zeros((0, 2), np.int32), nthreads=threads)
    t_self = (time.perf_counter() - t) * cfg.n / ns
    # long range: all atoms, a subset of the batch
    ps = max(1, min(len(kvecs), int(0.3 * target_s / max(cfg.n * 4e-7 / threads, 1e-9))))
    t = time.perf_counter()
    O.long_forces(p, xyz, q, kvecs[:ps], weight=weight, nthreads=threads)
    t_long = (time.perf_counter() - t) * len(kvecs) / ps
    t_step = t_pairs + t_self + t_long
    return {
        "value": cfg.n / t_step, "unit": "atom-force evals/s", "cores": threads,
        "kind": "port",
        "sample": (f"{cfg.name}: {m} of {len(pairs)} pairs, {ns} of {cfg.n} self terms, "
                   f"{ps} of {len(kvecs)} k vectors over all atoms; each part scaled "
                   f"linearly to one full step (C FP64 restatement, OpenMP)"),
        "seconds_per_step_est": t_step,
        "parts_s": {"pairs": t_pairs, "self": t_self, "long": t_long},
        "build": os.path.basename(lib_path) + (" (gcc -O3 -march=x86-64-v4)" if "v4" in lib_path
                                               else " (gcc -O3 -march=x86-64-v3)"),
    }


def oracle_sampled_errors(cfg, xyz, q, runs, nrows=200, threads=16):
    """rel-RMS force error of each run's last step against the FP64 oracle,
    over a seeded sample of nrows atoms: the oracle's short range for those
    rows (qo_short_forces_rows, each row's pairs by an O(N) scan) plus its
    long range for the run's last k batch (all atoms; the sampled rows
    kept).  runs: {precision: run_precision result}."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from qem4molly_amd.workloads import rel_rms
    O.use_avx512_build()
    t0 = time.perf_counter()
    p = O.params(cfg.L, cfg.r_cutoff, cfg.alpha, cfg.gamma_1, cfg.gamma_2, cfg.eps_0,
                 cfg.accuracy, cfg.N_t, 0)
    K, _, sK, _ = O.kset(cfg.Lx, cfg.Ly, cfg.alpha, cfg.accuracy)
    rows = np.sort(np.random.default_rng(17).choice(cfg.n, min(nrows, cfg.n), replace=False))
    short = O.short_forces_rows(p, xyz, q, rows, nthreads=threads)
    out, longs = {}, {}
    for prec, run in runs.items():
        key = run["kidx_last"].tobytes()
        if key not in longs:
            longs[key] = O.long_forces(p, xyz, q, K[run["kidx_last"]], weight=sK / cfg.P,
                                       nthreads=threads)[rows]
        out[prec] = rel_rms(run["f"][rows], short + longs[key])
    out["rows"] = int(rows.shape[0])
    out["oracle_seconds"] = time.perf_counter() - t0
    return out


def host_cpu():
    """The host's CPU model and logical CPU count."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"model": model, "logical_cpus": os.cpu_count()}


def _params(cfg, precision, semantics=0, iter_period=100):
    from qem4molly_amd import _lib as L
    p = L.default_params()
    p.Lx, p.Ly, p.Lz = cfg.L
    p.r_cutoff, p.alpha = cfg.r_cutoff, cfg.alpha
    p.gamma_1, p.gamma_2 = cfg.gamma_1, cfg.gamma_2
    p.eps_0, p.accuracy = cfg.eps_0, cfg.accuracy
    p.N_t, p.iter_period, p.rbe_p = cfg.N_t, iter_period, cfg.P
    p.precision = L.PREC_MIXED if precision == "mixed" else L.PREC_FP64
    p.semantics = semantics
    return p


def adjoint_cost(name, precision, reps=10):
    """Forward (forces) vs backward (force HVP, quadrature recomputed) on the
    same inputs and k batch; energy + dE/dq + dE/dgamma timed alongside."""
    from qem4molly_amd import _lib as L
    from qem4molly_amd.workloads import CONFIGS, make_system
    cfg = CONFIGS[name]
    xyz, q = make_system(cfg)
    v = np.random.default_rng(1).normal(size=xyz.shape)
    xd, qd, vd = L.DeviceArray(xyz), L.DeviceArray(q), L.DeviceArray(v)
    fd, hd, pd = L.DeviceArray(np.zeros_like(xyz)), L.DeviceArray(np.zeros_like(xyz)), \
        L.DeviceArray(np.zeros(cfg.n))
    ctx = L.Context(_params(cfg, precision, semantics=7, iter_period=1000000))
    kidx = ctx.sample_k(1, cfg.P)

    def timed(fn, blocks=3):
        # the fastest of three blocks of reps calls (each block's mean), so a
        # stray host delay in one block does not move the ratios
        fn()
        L.synchronize()
        best = None
        for _ in range(blocks):
            t = time.perf_counter()
            for _ in range(reps):
                fn()
            L.synchronize()
            ms = (time.perf_counter() - t) / reps * 1e3
            best = ms if best is None else min(best, ms)
        return best

    t_f = timed(lambda: ctx.forces_dev(cfg.n, xd.ptr, qd.ptr, kidx, fd.ptr))
    t_h = timed(lambda: ctx.hvp_dev(cfg.n, xd.ptr, qd.ptr, kidx, vd.ptr, hd.ptr))
    t_e = timed(lambda: ctx.energy_dev(cfg.n, xd.ptr, qd.ptr, kidx, pd.ptr))
    t_fe = timed(lambda: ctx.forces_energy_dev(cfg.n, xd.ptr, qd.ptr, kidx, fd.ptr, pd.ptr))
    mem = ctx.device_bytes()
    # the same step with the forward keeping the pair quadrature's tangent
    # node sums (qem_forces_energy_save_dev) and the HVP combining them
    # (qem_hvp_saved_dev): each call alone, then the pair as one step
    t_fes = timed(lambda: ctx.forces_energy_dev(cfg.n, xd.ptr, qd.ptr, kidx, fd.ptr, pd.ptr,
                                                save=True))
    t_hs = timed(lambda: ctx.hvp_dev(cfg.n, xd.ptr, qd.ptr, kidx, vd.ptr, hd.ptr, saved=True))

    def step_saved():
        ctx.forces_energy_dev(cfg.n, xd.ptr, qd.ptr, kidx, fd.ptr, pd.ptr, save=True)
        ctx.hvp_dev(cfg.n, xd.ptr, qd.ptr, kidx, vd.ptr, hd.ptr, saved=True)
    t_ss = timed(step_saved)
    mem_saved = ctx.device_bytes()
    ctx.close()
    # the differentiable op's forward is the one-call forces + energy (F, E,
    # dE/dq, dE/dgamma); its backward is the HVP
    return {"workload": name, "precision": precision, "forward_ms": t_f, "hvp_ms": t_h,
            "energy_dE_dq_dE_dgamma_ms": t_e, "forces_energy_ms": t_fe,
            "forces_energy_over_forces": t_fe / t_f,
            "backward_over_forward": t_h / t_fe,
            "forward_plus_backward_over_forward": (t_fe + t_h) / t_fe,
            "adjoint_step_over_force_step": (t_fe + t_h) / t_f,
            "forces_energy_save_ms": t_fes, "hvp_saved_ms": t_hs, "saved_step_ms": t_ss,
            "saved_step_over_force_step": t_ss / t_f,
            "device_bytes": mem, "device_bytes_saved": mem_saved}


def general_step(cfg, xyz_d, q_d, f_d, steps, warmup, seed0=1000):
    """The device-resident general-interaction step of test/QEM_test.jl:20-28
    (two Substrate_LJ walls, at z = 0 facing up and z = Lz facing down, plus
    QEM short + long) through qem_general_forces_dev, mixed precision; the
    wall pass and its domain-flag readback are inside the timed region."""
    from qem4molly_amd import _lib as L
    ctx = L.Context(_params(cfg, "mixed"))
    walls = [L.wall(sub_pos=0.0, direction=1), L.wall(sub_pos=cfg.Lz, direction=-1)]
    step = 0

    def one():
        nonlocal step
        kidx = ctx.sample_k(seed0 + step, cfg.P)
        ctx.general_forces_dev(cfg.n, xyz_d.ptr, q_d.ptr, kidx, walls, f_d.ptr)
        step += 1

    for _ in range(warmup):
        one()
    L.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        one()
    L.synchronize()
    dt = (time.perf_counter() - t0) / steps
    ctx.close()
    return {"workload": cfg.name, "walls": 2, "ms_per_step": dt * 1e3,
            "value": cfg.n / dt, "unit": "atom-force evals/s"}


def rbe_variance(cfg, xyz_d, q_d, reps=5):
    """Cost of the random-batch variance diagnostic (qem_long_variance_dev,
    mixed precision) next to the long-range force of the same batch, and the
    batch's relative RMS force noise sqrt(sum_i var_i) / |F_long|."""
    from qem4molly_amd import _lib as L
    ctx = L.Context(_params(cfg, "mixed"))
    kidx = ctx.sample_k(77, cfg.P)
    vd = L.DeviceArray(np.zeros(cfg.n))
    fd = L.DeviceArray(np.zeros((cfg.n, 3)))

    def timed(fn):
        fn()
        L.synchronize()
        t = time.perf_counter()
        for _ in range(reps):
            fn()
        L.synchronize()
        return (time.perf_counter() - t) / reps * 1e3

    t_v = timed(lambda: ctx.long_variance_dev(cfg.n, xyz_d.ptr, q_d.ptr, kidx, vd.ptr))
    t_f = timed(lambda: ctx.long_forces_dev(cfg.n, xyz_d.ptr, q_d.ptr, kidx, fd.ptr))
    var, f = vd.download(), fd.download()
    ctx.close()
    return {"workload": cfg.name, "P": cfg.P, "variance_ms": t_v, "long_forces_ms": t_f,
            "relative_rms_noise": float(np.sqrt(var.sum() / (f ** 2).sum()))}


def full_sum(cfg, xyz_d, q_d, reps=3):
    """The deterministic full k sum (rbe_mode = false): QEM short plus the
    long range over the whole K set with weights exp(-k^2/(4 alpha)),
    device-resident, mixed precision."""
    from qem4molly_amd import _lib as L
    ctx = L.Context(_params(cfg, "mixed"))
    K, _, _ = ctx.kset()
    fd = L.DeviceArray(np.zeros((cfg.n, 3)))

    def one():
        ctx.short_forces_dev(cfg.n, xyz_d.ptr, q_d.ptr, fd.ptr)
        ctx.long_terms_dev(cfg.n, xyz_d.ptr, q_d.ptr, K, -1.0, L.LONG_ALL, fd.ptr,
                           accumulate=True)

    one()
    L.synchronize()
    t = time.perf_counter()
    for _ in range(reps):
        one()
    L.synchronize()
    dt = (time.perf_counter() - t) / reps
    mem = ctx.device_bytes()
    ctx.close()
    return {"workload": cfg.name, "n_k": int(K.shape[0]), "ms_per_step": dt * 1e3,
            "value": cfg.n / dt, "unit": "atom-force evals/s", "device_bytes": mem}


def pole_step(steps=20, warmup=2):
    """gamma_1 = gamma_2 = -2 on the 100k slab (gamma_1 gamma_2 = 4 > 1): the
    pole-splitting rule (QEM_SEM_POLE_SPLIT, eps = k0/20), N_t = 240 so that
    every segment of the rule has nodes (k0 = ln 4 / 40), mixed precision,
    device-resident QEM step."""
    from dataclasses import replace
    from qem4molly_amd import _lib as L
    from qem4molly_amd.workloads import CONFIGS, make_system
    cfg = replace(CONFIGS["slab_n100k"], name="slab_n100k_pole", gamma_1=-2.0, gamma_2=-2.0,
                  N_t=240)
    xyz, q = make_system(CONFIGS["slab_n100k"], seed=0)
    xd, qd, fd = L.DeviceArray(xyz), L.DeviceArray(q), L.DeviceArray(np.zeros_like(xyz))
    ctx = L.Context(_params(cfg, "mixed", semantics=L.SEM_CORRECTED | L.SEM_POLE_SPLIT))
    for s in range(warmup):
        ctx.forces_dev(cfg.n, xd.ptr, qd.ptr, ctx.sample_k(500 + s, cfg.P), fd.ptr)
    L.synchronize()
    t = time.perf_counter()
    for s in range(steps):
        ctx.forces_dev(cfg.n, xd.ptr, qd.ptr, ctx.sample_k(600 + s, cfg.P), fd.ptr)
    L.synchronize()
    dt = (time.perf_counter() - t) / steps
    npairs = ctx.short_state()[0]
    f = fd.download()
    # the pair kernel's time from HIP events, the short range timed alone
    g = L.DeviceArray(np.zeros_like(xyz))
    ctx.timing_enable(True)
    for s in range(5):
        ctx.short_forces_dev(cfg.n, xd.ptr, qd.ptr, g.ptr)
    L.synchronize()
    ph = ctx.timing_read(reset=True)
    ctx.close()
    return {"workload": cfg.name, "gamma": [-2.0, -2.0], "N_t": 240, "ms_per_step": dt * 1e3,
            "value": cfg.n / dt, "unit": "atom-force evals/s", "n_pairs": npairs,
            "pair_ms": ph["pair"][0] / 5, "finite": bool(np.all(np.isfinite(f)))}


LAMMPS_CHILD = r"""
import json, sys, time
import numpy as np
import torch                                   # first: one HIP runtime
sys.path.insert(0, ROOT)
from qem4molly_amd import lammps_coupling as LC
from qem4molly_amd.workloads import CONFIGS, make_system

class Atoms:                                   # numpy per-atom arrays, local order
    def __init__(self, xyz, types, rng):
        order = rng.permutation(xyz.shape[0])
        self.id = order + 1
        self.type = types[order]
        self.x = xyz[order].copy()
        self.f = np.zeros_like(self.x)
        self.v = np.zeros_like(self.x)
    def get_natoms(self):
        return self.id.shape[0]
    def extract_atom(self, name):
        return getattr(self, name)

cfg = CONFIGS["slab_n10k"]
xyz, q = make_system(cfg, seed=0)
types = np.where(q > 0, 1, 2)
parcar = {"1": [1.0, 1.0, int((q > 0).sum())], "2": [1.0, -1.0, int((q < 0).sum())]}
incar = LC.load_incar(parcar, incar={"L_x": cfg.Lx, "L_y": cfg.Ly, "L_z": cfg.Lz,
                                     "g_1": cfg.gamma_1, "g_2": cfg.gamma_2,
                                     "accuracy": cfg.accuracy, "mode": "rbm",
                                     "RBM_p": cfg.P, "N_t": cfg.N_t})
lmp = Atoms(xyz, types, np.random.default_rng(1))
with LC.QEMLammps(incar, parcar, precision="mixed") as drv:
    for _ in range(3):
        drv.update_force(lmp)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(STEPS):
        drv.update_force(lmp)
    dt = (time.perf_counter() - t) / STEPS
print("RESULT " + json.dumps({"workload": "slab_n10k", "alpha": incar["alpha"],
                              "r_c": incar["r_c"], "ms_per_step": dt * 1e3,
                              "value": cfg.n / dt, "unit": "atom-force evals/s"}), flush=True)
"""


def lammps_step(steps=20):
    """One LAMMPS-coupled force update per step (the legacy driver's
    update_Force_njit through the torch binding), slab_n10k charges and
    box with the driver's own alpha = N/(Lx Ly) and r_c: POSCAR from the
    per-atom arrays by id, host-to-device copy, QEM step, forces back and
    added in local order.  Runs in a child process that imports torch first."""
    import subprocess
    root = os.path.dirname(os.path.abspath(__file__))
    code = LAMMPS_CHILD.replace("ROOT", repr(root)).replace("STEPS", str(int(steps)))
    r = subprocess.run([sys.executable, "-u", "-c", code], capture_output=True, text=True,
                       timeout=600)
    lines = [l for l in r.stdout.splitlines() if l.startswith("RESULT ")]
    if r.returncode != 0 or not lines:
        return {"error": (r.stderr or r.stdout)[-500:]}
    return json.loads(lines[-1][len("RESULT "):])


def adjoint_md(name="slab_n100k", steps=100, sigma=0.01):
    """100 steps of positions perturbed by N(0, sigma) per component
    (x, y wrapped, z reflected into [0.5, Lz - 0.5]), fixed seeded k batch,
    mixed precision; each step: forward (forces, energy, dE/dq, dE/dgamma in
    one call) and backward (force HVP along a fresh gaussian cotangent)."""
    from qem4molly_amd import _lib as L
    from qem4molly_amd.workloads import CONFIGS, make_system
    cfg = CONFIGS[name]
    xyz, q = make_system(cfg)
    rng = np.random.default_rng(5)
    xd, qd = L.DeviceArray(xyz), L.DeviceArray(q)
    fd, hd, vd = (L.DeviceArray(np.zeros_like(xyz)) for _ in range(3))
    pd = L.DeviceArray(np.zeros(cfg.n))
    ctx = L.Context(_params(cfg, "mixed"))
    ctx.timing_enable(True)
    kidx = ctx.sample_k(0, cfg.P)
    lo, hi = 0.5, cfg.Lz - 0.5
    t_dev = 0.0
    t0 = time.perf_counter()
    for s in range(steps):
        xyz = xyz + rng.normal(scale=sigma, size=xyz.shape)
        xyz[:, 0] %= cfg.Lx
        xyz[:, 1] %= cfg.Ly
        z = xyz[:, 2]
        z = np.where(z < lo, 2 * lo - z, z)
        xyz[:, 2] = np.where(z > hi, 2 * hi - z, z)
        # the previous step's HVP runs on the context's own stream, which a
        # plain (null-stream) copy does not wait for
        L.synchronize()
        xd.upload(xyz)
        vd.upload(rng.normal(size=xyz.shape))
        ctx.forces_energy_dev(cfg.n, xd.ptr, qd.ptr, kidx, fd.ptr, pd.ptr)
        ctx.hvp_dev(cfg.n, xd.ptr, qd.ptr, kidx, vd.ptr, hd.ptr)
    L.synchronize()
    wall = time.perf_counter() - t0
    ph = ctx.timing_read()
    f = fd.download()
    mem = ctx.device_bytes()
    npairs, builds, _ = ctx.short_state()
    ctx.close()
    return {"workload": name, "steps": steps, "sigma": sigma, "precision": "mixed",
            "ms_per_step_wall": wall / steps * 1e3,
            "gpu_ms_per_step": {k: v[0] / steps for k, v in ph.items()},
            "neighbor_builds": builds, "n_pairs_last": npairs,
            "finite": bool(np.isfinite(f).all()), "peak_device_bytes": mem}


def traffic_from_profiles(workload):
    path = os.path.join(ROOT, "profiles", f"pair_kernel_traffic_{workload}.json")
    if not os.path.exists(path):
        return None
    try:
        with open(path) as fh:
            d = json.load(fh)
        return float(d["hbm_bytes_per_dispatch"])
    except Exception:
        return None


def valu_issue_from_profiles(workload, kernel):
    """Measured VALU issue fraction of a kernel (rocprofv3 SQ_INSTS_VALU over
    GRBM_GUI_ACTIVE, committed under profiles/), or None."""
    path = os.path.join(ROOT, "profiles", f"valu_issue_{workload}.json")
    if not os.path.exists(path):
        return None
    try:
        with open(path) as fh:
            d = json.load(fh)
        return float(d["kernels"][kernel]["issue_frac"])
    except Exception:
        return None


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="slab_n100k",
                    choices=["slab_n100", "slab_n10k", "slab_n100k"])
    ap.add_argument("--fp64-steps", type=int, default=None,
                    help="steps for the FP64 plain comparison (default: --steps)")
    ap.add_argument("--no-secondary", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-oracle-check", action="store_true",
                    help="skip the oracle check of the last step on 200 sampled atoms")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the full-sum, pole-splitting and LAMMPS-coupling lines")
    ap.add_argument("--cpu-threads", type=int, default=None)
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="after the run, copy this many MiB three times with the library's "
                         "plain copy kernel (a known-size reference for HBM byte counters)")
    ap.add_argument("--adjoint-only", default=None, metavar="WORKLOAD:PRECISION",
                    help="time only the forward, the HVP and the energy of one workload "
                         "(e.g. slab_n100k:mixed) and print that one JSON line")
    a = ap.parse_args(argv)

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if a.gpus != 1 or world != 1:
        if int(os.environ.get("RANK", "0")) == 0:
            print("bench.py: this build runs the QEM force path on one GPU only; "
                  f"--gpus {a.gpus} (WORLD_SIZE={world}) is not supported. Across GPUs the "
                  "short range would shard by xy domain with an r_cutoff halo and the long "
                  "range by k batch (positions all-gathered, forces reduce-scattered).",
                  file=sys.stderr)
        return 2

    if a.adjoint_only:
        name, _, prec = a.adjoint_only.partition(":")
        if name not in ("slab_n100", "slab_n10k", "slab_n100k") or prec not in ("mixed", "fp64"):
            ap.error("--adjoint-only takes WORKLOAD:PRECISION, e.g. slab_n100k:mixed")
        print(json.dumps(adjoint_cost(name, prec)), flush=True)
        return 0

    from qem4molly_amd import _lib as L
    from qem4molly_amd.workloads import CONFIGS, make_system, rel_rms

    def measure(name, steps, warmup, fp64_steps):
        cfg = CONFIGS[name]
        xyz, q = make_system(cfg, seed=0)
        xd, qd, fd = L.DeviceArray(xyz), L.DeviceArray(q), L.DeviceArray(np.zeros_like(xyz))
        log(f"[bench] {name}: N={cfg.n} mixed x {steps} steps")
        mix = run_precision(cfg, xd, qd, fd, "mixed", steps, warmup)
        log(f"[bench] {name}: mixed {mix['ms_per_step']:.3f} ms/step; fp64 x {fp64_steps} steps")
        f64 = run_precision(cfg, xd, qd, fd, "fp64", fp64_steps, warmup)
        log(f"[bench] {name}: fp64 {f64['ms_per_step']:.3f} ms/step")
        # both runs end on the same k batch (same seeds, same step count)
        same_batch = (steps + warmup) == (fp64_steps + warmup)
        err = rel_rms(mix["f"], f64["f"]) if same_batch else None
        if not a.no_oracle_check:
            log(f"[bench] {name}: oracle check on sampled rows")
            err_o = oracle_sampled_errors(cfg, xyz, q, {"mixed": mix, "fp64": f64})
            mix["oracle_err"], f64["oracle_err"] = err_o["mixed"], err_o["fp64"]
            mix["oracle_check"] = err_o
        return cfg, xyz, q, mix, f64, err

    fp64_steps = a.fp64_steps or a.steps
    cfg, xyz, q, mix, f64, err = measure(a.workload, a.steps, a.warmup, fp64_steps)
    value = cfg.n / (mix["ms_per_step"] * 1e-3)
    value64 = cfg.n / (f64["ms_per_step"] * 1e-3)

    pair_avg_ms = mix["phases"]["pair"][1]
    fl = pair_flops(mix["npairs"], cfg.N_t)
    achieved = fl / (pair_avg_ms * 1e-3) / 1e12
    p64_avg = f64["phases"]["pair"][1]
    achieved64 = fl / (p64_avg * 1e-3) / 1e12

    out = {
        "metric": "atom-force evals/s, mixed vs FP64 plain on 1 GPU at 1e-6 rel RMS force err",
        "value": value,
        "unit": "atom-force evals/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": mix["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32/f64",
        "data": "synthetic (seeded uniform slab, see qem4molly_amd/workloads.py)",
        "config": {
            "workload": cfg.name, "N": cfg.n, "box": list(cfg.L), "gamma_down": cfg.gamma_1,
            "gamma_up": cfg.gamma_2, "alpha": cfg.alpha, "r_cutoff": cfg.r_cutoff,
            "N_t": cfg.N_t, "P": cfg.P, "accuracy": cfg.accuracy, "iter_period": 100,
            "n_pairs": mix["npairs"], "precision": "mixed", "parallelism": "single GPU",
        },
        "fp64_plain": {"value": value64, "ms_per_step": f64["ms_per_step"],
                       "ms_per_step_rebuild_amortized": f64["ms_per_step_rebuild_amortized"],
                       "ms_per_step_no_rebuild": f64["ms_per_step_no_rebuild"]},
        "mixed_over_fp64": value / value64,
        "rel_rms_err_vs_fp64": err,
        "rel_rms_err_vs_oracle_sampled": mix.get("oracle_check"),
        # the timed steps hold mix["builds_timed"] neighbour-list rebuilds;
        # these two windows hold exactly one in 100 steps, and none
        "neighbor_rebuilds_in_timed_steps": mix["builds_timed"],
        "ms_per_step_rebuild_amortized": mix["ms_per_step_rebuild_amortized"],
        "ms_per_step_no_rebuild": mix["ms_per_step_no_rebuild"],
        "rebuild_windows_ok": mix["rebuild_windows_ok"] and f64["rebuild_windows_ok"],
        # each part timed alone (the step overlaps the long range with the
        # short range, so these add up to more than ms_per_step)
        "phases_ms_per_step": parts_per_step(mix),
        "phases_ms_per_step_fp64": parts_per_step(f64),
        "overlap": "mixed precision: long range on a second stream alongside the short range",
        "roofline": {
            "bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS,
            "traffic": traffic_from_profiles(cfg.name),
            "valu_issue_frac_profiled": valu_issue_from_profiles(cfg.name, "k_pair_list<mixed>"),
            "valu_issue_source": f"profiles/valu_issue_{cfg.name}.json (rocprofv3 --pmc, "
                                 "not measured in this run)",
            "kernel": "k_pair_list<mixed>", "avg_ms": pair_avg_ms,
            "flop_per_launch": fl,
        },
        "roofline_fp64": {
            "bound": "valu", "achieved": achieved64, "peak": PEAK_FP64_TFLOPS,
            "unit": "TFLOP/s", "frac": achieved64 / PEAK_FP64_TFLOPS,
            "valu_issue_frac_profiled": valu_issue_from_profiles(cfg.name, "k_pair_list<fp64>"),
            "kernel": "k_pair_list<fp64>", "avg_ms": p64_avg,
        },
    }

    xd_g, qd_g, fd_g = (L.DeviceArray(xyz), L.DeviceArray(q),
                        L.DeviceArray(np.zeros_like(xyz)))
    log(f"[bench] {cfg.name}: general step (2 walls + QEM) x {a.steps}")
    out["general_step"] = general_step(cfg, xd_g, qd_g, fd_g, a.steps, a.warmup)
    log(f"[bench] {cfg.name}: random-batch variance diagnostic")
    out["rbe_variance"] = rbe_variance(cfg, xd_g, qd_g)
    if not a.no_extras:
        log(f"[bench] {cfg.name}: full k sum")
        out["full_sum"] = full_sum(cfg, xd_g, qd_g)
    del xd_g, qd_g, fd_g
    if not a.no_extras:
        log("[bench] slab_n100k_pole: pole-splitting step")
        out["pole_step"] = pole_step()
        log("[bench] slab_n10k: LAMMPS-coupled force update")
        out["lammps_step"] = lammps_step()

    if not a.no_secondary and a.workload != "slab_n10k":
        c2, _, _, m2, d2, e2 = measure("slab_n10k", 50, 3, 50)
        out["secondary"] = {
            "workload": "slab_n10k", "value": c2.n / (m2["ms_per_step"] * 1e-3),
            "ms_per_step": m2["ms_per_step"],
            "fp64_plain_value": c2.n / (d2["ms_per_step"] * 1e-3),
            "mixed_over_fp64": d2["ms_per_step"] / m2["ms_per_step"],
            "rel_rms_err_vs_fp64": e2,
            "rel_rms_err_vs_oracle_sampled": m2.get("oracle_check"),
            "ms_per_step_rebuild_amortized": m2["ms_per_step_rebuild_amortized"],
            "ms_per_step_no_rebuild": m2["ms_per_step_no_rebuild"],
        }

    if not a.no_adjoint:
        log("[bench] adjoint cost (slab_n10k fp64, mixed; slab_n100k mixed) and 100-step run")
        out["adjoint"] = {
            "cost": [adjoint_cost("slab_n10k", "fp64"), adjoint_cost("slab_n10k", "mixed"),
                     adjoint_cost("slab_n100k", "mixed", reps=5)],
            "md_slab_n100k": adjoint_md(),
        }

    if not a.no_cpu_baseline:
        threads = a.cpu_threads or int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or \
            min(16, os.cpu_count() or 1)
        pctx_p = L.default_params()
        pctx_p.Lx, pctx_p.Ly, pctx_p.Lz = cfg.L
        pctx_p.r_cutoff, pctx_p.alpha = cfg.r_cutoff, cfg.alpha
        pctx_p.gamma_1, pctx_p.gamma_2 = cfg.gamma_1, cfg.gamma_2
        pctx_p.N_t, pctx_p.rbe_p = cfg.N_t, cfg.P
        with L.Context(pctx_p) as c:
            pairs = c.neighbor_list(xyz)
            K, Pk, sK = c.kset()
            kidx = c.sample_k(1000, cfg.P)
        log(f"[bench] cpu baseline on {threads} threads")
        out["cpu_baseline"] = cpu_baseline(cfg, xyz, q, pairs, K[kidx], sK / cfg.P, threads)
        out["cpu_baseline"]["host"] = host_cpu()
        # the same restatement on one core, and on the smaller workloads
        more = [dict(cpu_baseline(cfg, xyz, q, pairs, K[kidx], sK / cfg.P, 1, target_s=10.0),
                     workload=cfg.name)]
        for name in ("slab_n100", "slab_n10k"):
            c2 = CONFIGS[name]
            x2, q2 = make_system(c2, seed=0)
            with L.Context(_params(c2, "fp64")) as c:
                pr2 = c.neighbor_list(x2)
                K2, _, sK2 = c.kset()
                ki2 = c.sample_k(1000, c2.P)
            for th in (1, threads):
                log(f"[bench] cpu baseline {name} on {th} threads")
                more.append(dict(cpu_baseline(c2, x2, q2, pr2, K2[ki2], sK2 / c2.P, th,
                                              target_s=3.0), workload=name))
        out["cpu_baseline_more"] = more

    if a.copy_calibration_mib > 0:
        nb = a.copy_calibration_mib << 20
        src = L.DeviceArray(np.zeros(nb, np.uint8))
        dst = L.DeviceArray(np.zeros(nb, np.uint8))
        for _ in range(3):
            L.dev_copy(dst, src, nb)
        L.synchronize()
        out["copy_calibration_bytes"] = nb
        log(f"[bench] copy calibration: 3 x {a.copy_calibration_mib} MiB")

    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
