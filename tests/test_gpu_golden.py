"""GPU checks against the committed fixtures and size-independent properties
at the benchmark sizes."""
import os

import numpy as np
import pytest

from helpers import CONFIGS, L, O, ctx_for, make_system, max_rel, oracle_params, rel_rms, scaled_config

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden", "qem_golden.npz")


@pytest.fixture(scope="module")
def g():
    return np.load(GOLD, allow_pickle=False)


def test_golden_two_atom(g):
    from qem4molly_amd.workloads import SlabConfig
    ri, rj = g["two_atom_ri"], g["two_atom_rj"]
    for prec, tol in (("fp64", 1e-13), ("mixed", 2e-6)):
        for row in g["two_atom_table"]:
            g1, g2, alpha, single, sem = row[:5]
            cfg = SlabConfig("two", 2, 10.0, 10.0, 2.0, g1, g2, alpha, 1.0, 30, 1)
            with ctx_for(cfg, semantics=int(sem), rbe_p=0, precision=prec) as c:
                f, fzj = c.short_pair([1.0], [1.0], [ri], [rj], single=bool(single))
                fs = c.short_self([1.0], [ri], single=bool(single))
            got = np.array([*f[0], fzj[0], fs[0, 2]])
            assert max_rel(got, row[5:]) < tol, (prec, row[:5])


def test_golden_four_atom(g):
    from qem4molly_amd import interactions as I
    xyz, q = g["four_atom_xyz"], g["four_atom_q"]
    atoms = [I.Atom(charge=v) for v in q]
    zl = np.array([1, 0, 2, 3])
    for (g1, g2, kx, ky, k), ref in zip(g["four_atom_params"], g["four_atom_parts"]):
        k3 = np.array([kx, ky, k])
        el = I.greens_element_init(g1, g2, 2.0, 1.0)
        got = [I.Fl_non_sort(k3, atoms, xyz, el), I.Fl_sort(k3, atoms, zl, xyz, el),
               I.Flz_self(k3, atoms, xyz, el), I.Flz_sort_k0(atoms, zl, xyz, el)]
        for a, b in zip(got, ref[:4]):
            assert max_rel(a, b) < 1e-13
        lng = I.QEM_long((10.0, 10.0, 2.0), gamma_1=g1, gamma_2=g2, rbe_p=1)
        tot = I.F_l_total(k3, I.System(atoms, xyz), lng)
        assert max_rel(tot, ref[4]) < 1e-13


def test_golden_slab_n100(g):
    cfg = CONFIGS["slab_n100"]
    xyz, q = g["slab_xyz"], g["slab_q"]
    with ctx_for(cfg) as c:
        pairs = c.neighbor_list(xyz)
        assert set(map(tuple, pairs.tolist())) == set(map(tuple, g["slab_pairs"].tolist()))
        assert np.array_equal(c.sortperm_z(xyz), g["slab_zlist"])
        assert rel_rms(c.short_forces(xyz, q), g["slab_short_as_written"]) < 1e-12
        assert rel_rms(c.long_forces(xyz, q, g["slab_kidx"]), g["slab_long"]) < 1e-12
        K, P, sK = c.kset()
        assert sK == float(g["slab_sum_K"])
    with ctx_for(cfg, semantics=L.SEM_CORRECTED) as c:
        assert rel_rms(c.short_forces(xyz, q), g["slab_short_corrected"]) < 1e-12


@pytest.mark.parametrize("name", ["slab_n100", "slab_n10k", "slab_n100k"])
def test_kset_bit_exact(name):
    cfg = CONFIGS[name]
    with ctx_for(cfg) as c:
        K, P, sK = c.kset()
    Ko, Po, sKo, _ = O.kset(cfg.Lx, cfg.Ly, cfg.alpha, cfg.accuracy)
    assert np.array_equal(K, Ko) and np.array_equal(P, Po) and sK == sKo


def test_sampler_frequencies():
    # alias sampler: counts follow P_k / sum(P) (chi-square on the heaviest k)
    cfg = CONFIGS["slab_n100"]
    with ctx_for(cfg) as c:
        K, P, sK = c.kset()
        draws = np.concatenate([c.sample_k(s, 20000) for s in range(10)])
        again = c.sample_k(3, 20000)
    assert np.array_equal(again, draws[3 * 20000:4 * 20000])      # seeded
    p = P.astype(np.float64) / P.astype(np.float64).sum()
    top = np.argsort(-p)[:200]
    cnt = np.bincount(draws, minlength=len(p))
    expected = p[top] * draws.size
    chi2 = np.sum((cnt[top] - expected) ** 2 / expected)
    assert chi2 < 200 + 6 * np.sqrt(2 * 200)
    assert cnt.sum() == draws.size and cnt[p == 0].sum() == 0


def test_batch_mean_is_full_sum():
    # random-batch estimator is unbiased: mean over many batches -> the
    # exp(-k^2/4a)-weighted sum over the whole K set (same parts, same k = 0)
    cfg = CONFIGS["slab_n100"]
    xyz, q = make_system(cfg)
    with ctx_for(cfg) as c:
        K, P, sK = c.kset()
        full = c.long_terms(xyz, q, K, weight=-1.0)
        nb = 400
        runs = np.stack([c.long_forces(xyz, q, c.sample_k(1000 + s, cfg.P)) for s in range(nb)])
    mean = runs.mean(axis=0)
    sem = runs.std(axis=0, ddof=1) / np.sqrt(nb)
    # mean squared deviation in units of the standard error: ~1 if unbiased;
    # a bias of a few per cent of the force would push it far above 2
    t = np.sum((mean - full) ** 2) / np.sum(sem ** 2)
    assert t < 2.0, t
    assert rel_rms(mean, full) < 0.1
    ref = O.long_forces(O.params(cfg.L, cfg.r_cutoff, cfg.alpha, cfg.gamma_1, cfg.gamma_2,
                                 cfg.eps_0, cfg.accuracy, cfg.N_t), xyz, q, K, full_sum=True,
                        nthreads=16)
    assert rel_rms(full, ref) < 1e-12


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_full_size_properties_slab_n100k(prec):
    # at N = 100,000: pair set = brute force on a sampled subset of atoms,
    # the short-range sum of forces obeys Newton's third law (as written the
    # pair forces cancel; only the self-image z terms remain), determinism
    cfg = CONFIGS["slab_n100k"]
    xyz, q = make_system(cfg)
    with ctx_for(cfg, precision=prec) as c:
        f1 = c.short_forces(xyz, q)
        f2 = c.short_forces(xyz, q)
        self_only = c.short_self(q, xyz)
        kidx = c.sample_k(5, cfg.P)
        l1 = c.long_forces(xyz, q, kidx)
        l2 = c.long_forces(xyz, q, kidx)
    assert np.array_equal(f1, f2) and np.array_equal(l1, l2)   # deterministic
    net = f1.sum(axis=0) - self_only.sum(axis=0)
    assert np.all(np.abs(net) < 1e-9 * np.abs(f1).sum(axis=0).max())
    assert np.all(np.isfinite(l1))


def test_neighbor_list_full_size_sampled():
    cfg = CONFIGS["slab_n100k"]
    xyz, q = make_system(cfg)
    with ctx_for(cfg, rbe_p=0) as c:
        pairs = c.neighbor_list(xyz)
    rng = np.random.default_rng(0)
    pick = rng.choice(cfg.n, 60, replace=False)
    have = {}
    for i, j in pairs.tolist():
        have.setdefault(i, set()).add(j)
        have.setdefault(j, set()).add(i)
    for a in pick:
        dx = xyz[:, 0] - xyz[a, 0]; dy = xyz[:, 1] - xyz[a, 1]
        dx -= cfg.Lx * np.rint(dx / cfg.Lx); dy -= cfg.Ly * np.rint(dy / cfg.Ly)
        ref = set(np.nonzero(dx * dx + dy * dy <= cfg.r_cutoff ** 2)[0].tolist()) - {int(a)}
        assert have.get(int(a), set()) == ref


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_full_sum_over_a_large_k_set(prec):
    # a K set of 23,408 vectors runs in passes of 8,192 that add into the
    # output (the k = 0 row with the last); same result as one pass
    cfg = scaled_config(CONFIGS["slab_n10k"], 2000)
    xyz, q = make_system(cfg, seed=8)
    sel = np.random.default_rng(0).choice(cfg.n, 300, replace=False)
    xyz, q = xyz[sel], q[sel]
    with ctx_for(cfg, precision=prec) as c:
        K, _, _ = c.kset()
        assert K.shape[0] > 2 * 8192
        full = c.long_terms(xyz, q, K, weight=-1.0)
    ref = O.long_forces(oracle_params(cfg), xyz, q, K, full_sum=True, nthreads=16)
    assert rel_rms(full, ref) < (1e-12 if prec == "fp64" else 1e-6)
