"""Random-batch variance diagnostic on the GPU: the per-atom batch estimate
against the oracle (qo_long_variance) on the same batch, and, at a larger
size, its meaning: averaged over batches it matches the measured spread of
the batch force about the full k sum (src/Force_long.jl:201-242)."""
import numpy as np
import pytest

from helpers import CONFIGS, L, O, ctx_for, make_system, oracle_params, rel_rms, scaled_config

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,n,P", [("slab_n100", 100, 40), ("slab_n10k", 600, 96),
                                      ("slab_n10k", 600, 200)])
def test_variance_fp64_matches_oracle(name, n, P):
    cfg = scaled_config(CONFIGS[name], n)
    xyz, q = make_system(cfg, seed=11)
    with ctx_for(cfg) as c:
        K, _, sK = c.kset()
        kidx = c.sample_k(5, P)
        f, var = c.long_variance(xyz, q, kidx)
        f_ref = c.long_forces(xyz, q, kidx)
    ref = O.long_variance(oracle_params(cfg), xyz, q, K[kidx], sK / P)
    assert np.all(ref > 0)
    assert rel_rms(var, ref) < 1e-11
    assert np.max(np.abs(var - ref) / ref) < 1e-9
    # the force output is the long-range force of the same batch
    assert np.array_equal(f, f_ref)


def test_variance_mixed_and_device_entry():
    cfg = scaled_config(CONFIGS["slab_n10k"], 800)
    xyz, q = make_system(cfg, seed=12)
    P = 128
    with ctx_for(cfg, precision="mixed") as c:
        K, _, sK = c.kset()
        kidx = c.sample_k(6, P)
        f, var = c.long_variance(xyz, q, kidx)
        d_x, d_q = L.DeviceArray(xyz), L.DeviceArray(q)
        d_v, d_f = L.DeviceArray(np.zeros(cfg.n)), L.DeviceArray(np.zeros_like(xyz))
        c.long_variance_dev(cfg.n, d_x.ptr, d_q.ptr, kidx, d_v.ptr, d_f.ptr)
        c.long_variance_dev(cfg.n, d_x.ptr, d_q.ptr, kidx, d_v.ptr)     # force not wanted
        L.synchronize()
        assert np.array_equal(d_v.download(), var)
        assert np.array_equal(d_f.download(), f)
        with pytest.raises(L.QEMError) as e:
            c.long_variance(xyz, q, kidx[:1])
        assert e.value.status == L.QEM_ERR_INVALID_ARG
        with pytest.raises(L.QEMError) as e:       # the pass's batch bound
            c.long_variance(xyz, q, c.sample_k(7, 65537))
        assert e.value.status == L.QEM_ERR_UNSUPPORTED
    ref = O.long_variance(oracle_params(cfg), xyz, q, K[kidx], sK / P)
    assert rel_rms(var, ref) < 1e-5


def test_variance_calibrated_against_full_sum():
    # 40 batches of 64 k vectors on 2,000 atoms; the full sum over the K set
    # is the batch force's expectation
    cfg = scaled_config(CONFIGS["slab_n10k"], 2000)
    xyz, q = make_system(cfg, seed=13)
    P, M = 64, 40
    with ctx_for(cfg) as c:
        K, _, _ = c.kset()
        full = c.long_terms(xyz, q, K, weight=-1.0)
        spread, est, F = [], [], []
        for m in range(M):
            f, var = c.long_variance(xyz, q, c.sample_k(100 + m, P))
            F.append(f)
            spread.append(((f - full) ** 2).sum())
            est.append(var.sum())
    F = np.array(F)
    # the batch mean approaches the full sum (relative to the batch noise)
    noise = np.sqrt(np.mean(spread))
    assert np.sqrt(((F.mean(axis=0) - full) ** 2).sum()) < 3 * noise / np.sqrt(M)
    assert np.mean(est) == pytest.approx(np.mean(spread), rel=0.15)
