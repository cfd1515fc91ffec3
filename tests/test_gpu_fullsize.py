"""GPU parity at full size: slab_n10k (N = 10,000) and slab_n100k
(N = 100,000), as the benchmark runs them, against the FP64 oracle.

The oracle's short range is evaluated for a sample of 200 atoms
(qo_short_forces_rows: each row's pairs by an O(N) scan, summed as
Molly.forces(::QEM_short) sums them, src/QEM_short.jl:41-56); its long range
(F_long, src/Force_long.jl:201-242) for every atom, with the same k batch.
Bars: FP64 within 1e-12 rel-RMS (the short and long parts separately and
their sum), mixed precision within 1e-6, over the sampled rows; the FP64
long range over all atoms.
"""
import numpy as np
import pytest

from helpers import CONFIGS, O, ctx_for, make_system, oracle_params, rel_rms

pytestmark = pytest.mark.gpu

NROWS = 200


@pytest.mark.parametrize("name", ["slab_n10k", "slab_n100k"])
def test_full_size_forces_vs_oracle_sampled_rows(name):
    cfg = CONFIGS[name]
    xyz, q = make_system(cfg, seed=0)
    rows = np.sort(np.random.default_rng(17).choice(cfg.n, NROWS, replace=False))
    got = {}
    with ctx_for(cfg, precision="fp64") as c:
        K, _, sK = c.kset()
        kidx = c.sample_k(0, cfg.P)
        got["fp64"] = c.forces(xyz, q, kidx)
        short64 = c.short_forces(xyz, q)
        long64 = c.long_forces(xyz, q, kidx)
    with ctx_for(cfg, precision="mixed") as c:
        got["mixed"] = c.forces(xyz, q, kidx)
    p = oracle_params(cfg)
    ref_short = O.short_forces_rows(p, xyz, q, rows, nthreads=16)
    ref_long = O.long_forces(p, xyz, q, K[kidx], weight=sK / cfg.P, nthreads=16)
    ref = ref_short + ref_long[rows]
    e_short = rel_rms(short64[rows], ref_short)
    e_long = rel_rms(long64, ref_long)
    e64 = rel_rms(got["fp64"][rows], ref)
    emix = rel_rms(got["mixed"][rows], ref)
    print(f"{name}: short {e_short:.2e} long(all atoms) {e_long:.2e} fp64 {e64:.2e} "
          f"mixed {emix:.2e}")
    assert e_short < 1e-12
    assert e_long < 1e-12
    assert e64 < 1e-12
    assert emix < 1e-6
