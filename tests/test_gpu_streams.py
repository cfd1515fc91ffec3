"""GPU: the mixed-precision force step runs the long range on a second
stream beside the short range and joins back for one output pass.  Its
forces equal, bit for bit, the short-range forces plus the long-range forces
of the same k batch computed by the one-stream entry points (the output
pass forms short + long in that order), over several steps and across a
neighbour-list rebuild, and on a caller's stream as on the context's.  The
energy and the HVP overlap the same way and are checked for repeatability
and against the one-stream FP64 path."""
import numpy as np
import pytest

from helpers import CONFIGS, L, ctx_for, make_system, rel_rms, scaled_config

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


@pytest.mark.parametrize("name,n", [("slab_n10k", 3000), ("slab_n100k", 20000)])
def test_two_stream_energy_and_hvp(name, n):
    """The mixed energy and HVP run their long range on the second stream too.
    Against the FP64 device path (one stream, pinned to the oracle by the
    adjoint tests) they agree to the mixed bound, and they are bitwise
    repeatable, over three k batches, on a caller's null stream as on the
    context's."""
    cfg = scaled_config(CONFIGS[name], n)
    xyz, q = make_system(cfg, seed=23)
    v = np.random.default_rng(24).normal(size=xyz.shape)
    with ctx_for(cfg, precision="mixed") as cm, ctx_for(cfg, precision="mixed") as cn, \
            ctx_for(cfg, precision="fp64") as cd:
        d_x, d_q, d_v = L.DeviceArray(xyz), L.DeviceArray(q), L.DeviceArray(v)
        bufs = {k: L.DeviceArray(np.zeros(cfg.n)) for k in ("pm", "pn", "pd")}
        hvs = {k: L.DeviceArray(np.zeros_like(xyz)) for k in ("hm", "hn", "hd")}
        for step in range(3):
            kidx = cm.sample_k(40 + step, cfg.P)
            Em, dgm = cm.energy_dev(cfg.n, d_x.ptr, d_q.ptr, kidx, bufs["pm"].ptr)
            En, dgn = cn.energy_dev(cfg.n, d_x.ptr, d_q.ptr, kidx, bufs["pn"].ptr,
                                    stream=L.STREAM_NULL)
            Ed, dgd = cd.energy_dev(cfg.n, d_x.ptr, d_q.ptr, kidx, bufs["pd"].ptr)
            cm.hvp_dev(cfg.n, d_x.ptr, d_q.ptr, kidx, d_v.ptr, hvs["hm"].ptr)
            cn.hvp_dev(cfg.n, d_x.ptr, d_q.ptr, kidx, d_v.ptr, hvs["hn"].ptr,
                       stream=L.STREAM_NULL)
            cd.hvp_dev(cfg.n, d_x.ptr, d_q.ptr, kidx, d_v.ptr, hvs["hd"].ptr)
            L.synchronize()
            pm, pn, pd = (bufs[k].download() for k in ("pm", "pn", "pd"))
            hm, hn, hd = (hvs[k].download() for k in ("hm", "hn", "hd"))
            assert Em == En and np.array_equal(dgm, dgn), step
            assert np.array_equal(pm, pn) and np.array_equal(hm, hn), step
            assert Em == pytest.approx(Ed, rel=1e-6), step
            assert rel_rms(pm, pd) < 1e-6 and rel_rms(hm, hd) < 1e-6, step
            assert np.allclose(dgm, dgd, rtol=1e-5), step
