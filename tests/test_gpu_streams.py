"""GPU: the mixed-precision force step runs the long range on a second
stream beside the short range and joins back for one output pass.  Its
forces equal, bit for bit, the short-range forces plus the long-range forces
of the same k batch computed by the one-stream entry points (the output
pass forms short + long in that order), over several steps and across a
neighbour-list rebuild, and on a caller's stream as on the context's."""
import numpy as np
import pytest

from helpers import CONFIGS, L, ctx_for, make_system, scaled_config

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,n", [("slab_n10k", 3000), ("slab_n100k", 20000)])
def test_two_stream_step_equals_its_parts(name, n):
    cfg = scaled_config(CONFIGS[name], n)
    xyz, q = make_system(cfg, seed=21)
    with ctx_for(cfg, precision="mixed", iter_period=2) as c1, \
            ctx_for(cfg, precision="mixed", iter_period=2) as c2:
        d_x, d_q = L.DeviceArray(xyz), L.DeviceArray(q)
        d_f = L.DeviceArray(np.zeros_like(xyz))
        d_s, d_l = L.DeviceArray(np.zeros_like(xyz)), L.DeviceArray(np.zeros_like(xyz))
        for step in range(4):          # rebuilds at steps 0 and 2 (iter_period 2)
            kidx = c1.sample_k(30 + step, cfg.P)
            c1.forces_dev(cfg.n, d_x.ptr, d_q.ptr, kidx, d_f.ptr)
            c2.short_forces_dev(cfg.n, d_x.ptr, d_q.ptr, d_s.ptr)
            c2.long_forces_dev(cfg.n, d_x.ptr, d_q.ptr, kidx, d_l.ptr)
            L.synchronize()
            f, fs, fl = d_f.download(), d_s.download(), d_l.download()
            assert np.array_equal(f, fs + fl), step
        assert c1.short_state()[1] == 2


def test_two_stream_step_with_walls_on_the_null_stream():
    cfg = scaled_config(CONFIGS["slab_n10k"], 2500)
    xyz, q = make_system(cfg, seed=22)
    walls = [L.wall(sub_pos=0.0, direction=1), L.wall(sub_pos=cfg.Lz, direction=-1)]
    with ctx_for(cfg, precision="mixed") as c1, ctx_for(cfg, precision="mixed") as c2:
        kidx = c1.sample_k(5, cfg.P)
        d_x, d_q = L.DeviceArray(xyz), L.DeviceArray(q)
        d_f1, d_f2 = L.DeviceArray(np.zeros_like(xyz)), L.DeviceArray(np.zeros_like(xyz))
        c1.general_forces_dev(cfg.n, d_x.ptr, d_q.ptr, kidx, walls, d_f1.ptr,
                              stream=L.STREAM_NULL)
        c2.general_forces_dev(cfg.n, d_x.ptr, d_q.ptr, kidx, walls, d_f2.ptr)
        L.synchronize()
        assert np.array_equal(d_f1.download(), d_f2.download())
