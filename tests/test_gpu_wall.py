"""Substrate_LJ wall on the GPU against the oracle (src/Substrate_LJ.jl:7-46),
and the device-resident general-interaction step (walls + QEM)."""
import numpy as np
import pytest

from helpers import CONFIGS, L, O, ctx_for, make_system, max_rel, scaled_config
from qem4molly_amd.interactions import Atom, Substrate_LJ, System, forces

pytestmark = pytest.mark.gpu


def _atoms(n, Lz, seed):
    rng = np.random.default_rng(seed)
    return np.column_stack([rng.uniform(0, 20, n), rng.uniform(0, 20, n),
                            rng.uniform(0.3, Lz - 0.3, n)])


@pytest.mark.parametrize("sem", [0, 1, 2, 3])
@pytest.mark.parametrize("direction,sub_pos", [(+1, 0.0), (-1, 10.0)])
def test_wall_matches_oracle(sem, direction, sub_pos):
    xyz = _atoms(3000, 10.0, 1)
    w = L.wall(cutoff=1.0, sigma=0.5, eps=1.2, sub_pos=sub_pos, direction=direction,
               semantics=sem)
    with ctx_for(CONFIGS["slab_n100"], rbe_p=0) as c:
        got = c.substrate_lj(xyz, w)
        f0 = np.random.default_rng(2).normal(size=xyz.shape)
        acc = c.substrate_lj(xyz, w, f=f0)
    ref, st = O.substrate_lj(xyz, 1.0, 0.5, 1.2, sub_pos, direction, sem)
    assert st == 0
    assert np.count_nonzero(ref[:, 2]) > 0
    assert max_rel(got, ref) < 1e-14
    assert np.allclose(acc, f0 + ref, rtol=1e-14, atol=1e-14 * np.abs(ref).max())


def test_wall_side_assertion():
    xyz = _atoms(100, 10.0, 3)
    xyz[17, 2] = -0.2
    w = L.wall(sub_pos=0.0, direction=1)
    with ctx_for(CONFIGS["slab_n100"], rbe_p=0) as c:
        with pytest.raises(L.QEMError) as e:
            c.substrate_lj(xyz, w)
        assert e.value.status == L.QEM_ERR_DOMAIN
        d_x, d_f = L.DeviceArray(xyz), L.DeviceArray(np.zeros_like(xyz))
        with pytest.raises(L.QEMError) as e:
            c.substrate_lj_dev(xyz.shape[0], w, d_x.ptr, d_f.ptr)
        assert e.value.status == L.QEM_ERR_DOMAIN
        xyz[17, 2] = 0.2
        d_x.upload(xyz)
        c.substrate_lj_dev(xyz.shape[0], w, d_x.ptr, d_f.ptr)   # no error once fixed
        ref, _ = O.substrate_lj(xyz, 1.0, 0.5, 1.0, 0.0, 1)
        assert max_rel(d_f.download(), ref) < 1e-14


def test_python_mirror_wall_forces():
    xyz = _atoms(500, 10.0, 4)
    sys_ = System(atoms=[Atom(charge=1.0)] * 500, coords=xyz)
    lo = Substrate_LJ(cutoff=1.0, sigma=0.5, eps=1.0, sub_pos=0.0, direction=+1)
    hi = Substrate_LJ(cutoff=1.0, sigma=0.5, eps=1.0, sub_pos=10.0, direction=-1)
    got = forces(lo, sys_) + forces(hi, sys_)
    ref = O.substrate_lj(xyz, 1.0, 0.5, 1.0, 0.0, 1)[0] + O.substrate_lj(xyz, 1.0, 0.5, 1.0, 10.0, -1)[0]
    assert max_rel(got, ref) < 1e-14


def test_general_forces_dev_is_walls_plus_qem():
    # the device-resident step of test/QEM_test.jl:28 (two walls + QEM)
    cfg = scaled_config(CONFIGS["slab_n10k"], 2000)
    xyz, q = make_system(cfg, seed=5)
    walls = [L.wall(sub_pos=0.0, direction=1), L.wall(sub_pos=cfg.Lz, direction=-1)]
    with ctx_for(cfg, precision="mixed") as c1, ctx_for(cfg, precision="mixed") as c2:
        kidx = c1.sample_k(9, cfg.P)
        d_x, d_q = L.DeviceArray(xyz), L.DeviceArray(q)
        d_f1, d_f2 = L.DeviceArray(np.zeros_like(xyz)), L.DeviceArray(np.zeros_like(xyz))
        for _ in range(3):          # several steps: the step counters stay in sync
            c1.general_forces_dev(cfg.n, d_x.ptr, d_q.ptr, kidx, walls, d_f1.ptr)
            c2.forces_dev(cfg.n, d_x.ptr, d_q.ptr, kidx, d_f2.ptr)
            for w in walls:
                c2.substrate_lj_dev(cfg.n, w, d_x.ptr, d_f2.ptr, accumulate=True)
            L.synchronize()
            a, b = d_f1.download(), d_f2.download()
            # same terms; the general step sums them in the reference's order
            # (walls, short, long), the sequential calls in another
            assert np.allclose(a, b, rtol=1e-14, atol=1e-15 * np.abs(b).max())
    wall_only = sum(O.substrate_lj(xyz, 1.0, 0.5, 1.0, w.sub_pos, w.direction)[0] for w in walls)
    assert np.count_nonzero(wall_only[:, 2]) > 0


def test_general_forces_dev_many_walls_in_order():
    # more walls than one pass takes (8): the second pass adds to the first
    cfg = scaled_config(CONFIGS["slab_n10k"], 1500)
    xyz, q = make_system(cfg, seed=6)
    walls = [L.wall(cutoff=1.0 + 0.05 * i, sigma=0.4 + 0.01 * i, eps=1.0 + 0.1 * i,
                    sub_pos=-0.03 * i, direction=1, semantics=i % 4) for i in range(6)] + \
            [L.wall(cutoff=0.9 + 0.05 * i, sigma=0.45, eps=0.7 + 0.2 * i,
                    sub_pos=cfg.Lz + 0.02 * i, direction=-1, semantics=(i + 1) % 4)
             for i in range(5)]
    with ctx_for(cfg, precision="mixed") as c1, ctx_for(cfg, precision="mixed") as c2:
        kidx = c1.sample_k(3, cfg.P)
        d_x, d_q = L.DeviceArray(xyz), L.DeviceArray(q)
        d_f1, d_f2 = L.DeviceArray(np.zeros_like(xyz)), L.DeviceArray(np.zeros_like(xyz))
        c1.general_forces_dev(cfg.n, d_x.ptr, d_q.ptr, kidx, walls, d_f1.ptr)
        c2.forces_dev(cfg.n, d_x.ptr, d_q.ptr, kidx, d_f2.ptr)
        L.synchronize()             # the context stream does not order the copy
        qem_only = d_f2.download()
        for w in walls:
            c2.substrate_lj_dev(cfg.n, w, d_x.ptr, d_f2.ptr, accumulate=True)
        L.synchronize()
        a, b = d_f1.download(), d_f2.download()
        assert np.allclose(a, b, rtol=1e-14, atol=1e-15 * np.abs(b).max())
        ref = sum(O.substrate_lj(xyz, w.cutoff, w.sigma, w.eps, w.sub_pos, w.direction,
                                 w.semantics)[0] for w in walls)
        assert np.count_nonzero(ref[:, 2]) > 0
        assert np.allclose(a, qem_only + ref, rtol=1e-13, atol=1e-14 * np.abs(a).max())
        # atoms behind the ninth wall (second pass): the whole call reports it,
        # and the next call with the original walls does not see the stale flag
        assert np.any(xyz[:, 2] < 1.0)
        bad = walls[:8] + [L.wall(sub_pos=1.0, direction=1)] + walls[8:]
        with pytest.raises(L.QEMError) as e:
            c1.general_forces_dev(cfg.n, d_x.ptr, d_q.ptr, kidx, bad, d_f1.ptr)
        assert e.value.status == L.QEM_ERR_DOMAIN
        c1.general_forces_dev(cfg.n, d_x.ptr, d_q.ptr, kidx, walls, d_f1.ptr)
