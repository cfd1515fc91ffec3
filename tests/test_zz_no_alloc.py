"""GPU: qem_long_forces_noalloc_dev, the long-range force call that
allocates nothing (its work buffers in the library's fixed per-GPU array),
against qem_long_forces_dev through the C ABI: the same forces bit for bit,
on two calls in a row, the second with a smaller input (the array keeps the
first call's contents), in FP64 and mixed precision; the context's device
memory does not grow; sizes beyond the limits are refused before any launch.
"""
import numpy as np
import pytest

from helpers import CONFIGS, L, ctx_for, make_system

pytestmark = pytest.mark.gpu


def _both(c, xyz, q, kidx):
    n = xyz.shape[0]
    xd, qd = L.DeviceArray(xyz), L.DeviceArray(q)
    fa, fb = L.DeviceArray(np.zeros_like(xyz)), L.DeviceArray(np.zeros_like(xyz))
    c.long_forces_dev(n, xd.ptr, qd.ptr, kidx, fa.ptr)
    c.long_forces_noalloc_dev(n, xd.ptr, qd.ptr, kidx, fb.ptr)
    L.synchronize()
    return fa.download(), fb.download()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_noalloc_long_forces_equal_the_allocating_call(precision):
    cfg = CONFIGS["slab_n100k"]
    xyz, q = make_system(cfg, seed=2)
    with ctx_for(cfg, precision=precision) as c:
        bytes0 = c.device_bytes()
        # first call: N = 100,000, P = 400 (the 100k slab's batch)
        xd, qd = L.DeviceArray(xyz), L.DeviceArray(q)
        f0 = L.DeviceArray(np.zeros_like(xyz))
        kidx = c.sample_k(4, cfg.P)
        c.long_forces_noalloc_dev(cfg.n, xd.ptr, qd.ptr, kidx, f0.ptr)
        L.synchronize()
        assert c.device_bytes() == bytes0          # nothing allocated in the context
        got1 = f0.download()
        ref1, again1 = _both(c, xyz, q, kidx)
        assert np.array_equal(got1, ref1) and np.array_equal(again1, ref1)
        # second call, smaller: the first 30,000 atoms, P = 123
        m = 30000
        k2 = c.sample_k(5, 123)
        ref2, got2 = _both(c, xyz[:m], q[:m], k2)
        assert np.array_equal(got2, ref2)
        assert np.all(np.isfinite(got2))
        # limits: refused on the host, nothing launched
        n_big = L.NOALLOC_MAX_N + 1
        with pytest.raises(L.QEMError) as e:
            c.long_forces_noalloc_dev(n_big, xd.ptr, qd.ptr, kidx, f0.ptr)
        assert e.value.status == L.QEM_ERR_CAPACITY
        with pytest.raises(L.QEMError) as e:
            c.long_forces_noalloc_dev(cfg.n, xd.ptr, qd.ptr, c.sample_k(6, L.NOALLOC_MAX_P + 1),
                                      f0.ptr)
        assert e.value.status == L.QEM_ERR_CAPACITY


def test_noalloc_without_k_set_is_refused():
    cfg = CONFIGS["slab_n100"]
    xyz, q = make_system(cfg)
    with ctx_for(cfg, rbe_p=0) as c:       # no K set built at creation
        xd, qd = L.DeviceArray(xyz), L.DeviceArray(q)
        fd = L.DeviceArray(np.zeros_like(xyz))
        with pytest.raises(L.QEMError) as e:
            c.long_forces_noalloc_dev(cfg.n, xd.ptr, qd.ptr, np.zeros(0, np.int64), fd.ptr)
        assert e.value.status == L.QEM_ERR_UNSUPPORTED
        c.kset()                            # builds it; then P = 0 (the k = 0 term only) runs
        c.long_forces_noalloc_dev(cfg.n, xd.ptr, qd.ptr, np.zeros(0, np.int64), fd.ptr)
        ref = L.DeviceArray(np.zeros_like(xyz))
        c.long_forces_dev(cfg.n, xd.ptr, qd.ptr, np.zeros(0, np.int64), ref.ptr)
        L.synchronize()
        assert np.array_equal(fd.download(), ref.download())
