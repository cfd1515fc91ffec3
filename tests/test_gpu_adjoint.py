"""GPU adjoint: energy, dE/dq, dE/dgamma and Hessian-vector products.

References: the FP64 oracle's energy (qo_energy: pair/self J0 quadrature and
an O(N^2) k-space sum, independent of the device's structure factors) and
central finite differences (h = 1e-5) of the forces and the energy.
"""
import numpy as np
import pytest

from helpers import CONFIGS, L, O, ctx_for, make_system, max_rel, oracle_params, rel_rms, \
    scaled_config

pytestmark = pytest.mark.gpu


def _setup(name, n, seed=0, sem=0, precision="fp64"):
    cfg = scaled_config(CONFIGS[name], n) if n != CONFIGS[name].n else CONFIGS[name]
    xyz, q = make_system(cfg, seed=seed)
    c = ctx_for(cfg, semantics=sem, precision=precision, iter_period=100000)
    return cfg, xyz, q, c


@pytest.mark.parametrize("name,n", [("slab_n100", 100), ("slab_n10k", 400)])
@pytest.mark.parametrize("sem", [0, L.SEM_CORRECTED])
def test_energy_matches_oracle(name, n, sem):
    cfg, xyz, q, c = _setup(name, n, sem=sem)
    with c:
        K, P, sK = c.kset()
        kidx = c.sample_k(3, cfg.P)
        E, phi, dg = c.energy(xyz, q, kidx)
    pairs = O.neighbor_pairs(xyz, cfg.Lx, cfg.Ly, cfg.r_cutoff)
    Eo, phio, dgo = O.energy(oracle_params(cfg, semantics=sem), xyz, q, pairs, K[kidx],
                             weight=sK / cfg.P, nthreads=16)
    assert E == pytest.approx(Eo, rel=1e-12)
    assert rel_rms(phi, phio) < 1e-12
    assert np.allclose(dg, dgo, rtol=1e-11, atol=1e-14 * abs(Eo))


def test_energy_mixed_precision():
    cfg, xyz, q, c = _setup("slab_n10k", 400, precision="mixed")
    with c:
        K, P, sK = c.kset()
        kidx = c.sample_k(4, cfg.P)
        E, phi, dg = c.energy(xyz, q, kidx)
    pairs = O.neighbor_pairs(xyz, cfg.Lx, cfg.Ly, cfg.r_cutoff)
    Eo, phio, dgo = O.energy(oracle_params(cfg), xyz, q, pairs, K[kidx], weight=sK / cfg.P,
                             nthreads=16)
    assert E == pytest.approx(Eo, rel=1e-6)
    assert rel_rms(phi, phio) < 1e-6
    assert np.allclose(dg, dgo, rtol=1e-5)


@pytest.mark.parametrize("sem", [0, L.SEM_CORRECTED])
def test_hvp_matches_fd_of_oracle_forces(sem):
    cfg, xyz, q, c = _setup("slab_n100", 100, sem=sem)
    rng = np.random.default_rng(1)
    v = rng.normal(size=xyz.shape)
    with c:
        K, P, sK = c.kset()
        kidx = c.sample_k(5, cfg.P)
        hv = c.hvp(xyz, q, kidx, v)
    p = oracle_params(cfg, semantics=sem)
    pairs = O.neighbor_pairs(xyz, cfg.Lx, cfg.Ly, cfg.r_cutoff)

    def F(X):
        return (O.short_forces(p, X, q, pairs, nthreads=16)
                + O.long_forces(p, X, q, K[kidx], weight=sK / cfg.P, nthreads=16))
    h = 1e-5
    fd = -(F(xyz + h * v) - F(xyz - h * v)) / (2 * h)
    assert rel_rms(hv, fd) < 1e-7


def test_hvp_short_and_long_parts_separately():
    # P = 0 isolates the short range; a huge cutoff-free check of the long
    # range comes from the difference
    cfg, xyz, q, c = _setup("slab_n100", 100, sem=L.SEM_CORRECTED)
    rng = np.random.default_rng(2)
    v = rng.normal(size=xyz.shape)
    p = oracle_params(cfg, semantics=L.SEM_CORRECTED)
    pairs = O.neighbor_pairs(xyz, cfg.Lx, cfg.Ly, cfg.r_cutoff)
    h = 1e-5
    with c:
        hv_short = c.hvp(xyz, q, np.zeros(0, np.int64), v)
        K, P, sK = c.kset()
        kidx = c.sample_k(6, cfg.P)
        hv_all = c.hvp(xyz, q, kidx, v)
    fd_s = -(O.short_forces(p, xyz + h * v, q, pairs) - O.short_forces(p, xyz - h * v, q, pairs)) / (2 * h)
    assert rel_rms(hv_short, fd_s) < 1e-7
    fl = lambda X: O.long_forces(p, X, q, K[kidx], weight=sK / cfg.P)
    fd_l = -(fl(xyz + h * v) - fl(xyz - h * v)) / (2 * h)
    assert rel_rms(hv_all - hv_short, fd_l) < 1e-7


def test_hvp_mixed_close_to_fp64():
    cfg, xyz, q, c = _setup("slab_n10k", 1500, precision="mixed")
    rng = np.random.default_rng(3)
    v = rng.normal(size=xyz.shape)
    with c:
        kidx = c.sample_k(7, cfg.P)
        hv_m = c.hvp(xyz, q, kidx, v)
    with ctx_for(cfg, iter_period=100000) as c64:
        hv_d = c64.hvp(xyz, q, kidx, v)
    assert rel_rms(hv_m, hv_d) < 1e-6


def test_adjoint_slab_n10k_fd_and_symmetry():
    # the N = 10,000 adjoint configuration: HVP for a seeded gaussian v
    # (seed 1) against central differences (h = 1e-5) of the FP64 device
    # forces (themselves pinned to the oracle), dE/dq and dE/dgamma against
    # differences of the FP64 energy; corrected semantics (conservative force)
    cfg = CONFIGS["slab_n10k"]
    xyz, q = make_system(cfg)
    v = np.random.default_rng(1).normal(size=xyz.shape)
    h = 1e-5
    with ctx_for(cfg, semantics=L.SEM_CORRECTED, iter_period=100000) as c:
        # The inputs are unfiltered uniform positions: a few pairs nearly touch
        # (F ~ 1/rho^2) or nearly share a z (the sign of z_i - z_j flips), and
        # some pairs cross r_cutoff between x - h v and x + h v (the force is
        # cut off there: a jump, not a derivative).  A central difference is
        # not accurate across those, so their atoms are held fixed in v; their
        # forces still respond to every other atom's motion.
        # The long-range z forces jump where two atoms swap z order (the
        # k = 0 term and the sorted direct term carry sgn(z_i - z_j)), so only
        # atoms at least 1e-3 from their neighbours in z order move in z.
        order = np.argsort(xyz[:, 2], kind="stable")
        gaps = np.diff(xyz[order, 2])
        lo = np.concatenate([[np.inf], gaps]); hi = np.concatenate([gaps, [np.inf]])
        zfree = np.zeros(cfg.n, bool)
        zfree[order] = (lo > 1e-3) & (hi > 1e-3)
        v[~zfree, 2] = 0.0
        assert zfree.sum() > 0.05 * cfg.n
        pairs = c.neighbor_list(xyz)
        d = xyz[pairs[:, 0]] - xyz[pairs[:, 1]]
        d[:, :2] -= np.array([cfg.Lx, cfg.Ly]) * np.rint(d[:, :2] / [cfg.Lx, cfg.Ly])
        bad = (np.linalg.norm(d, axis=1) < 0.2) | (np.abs(d[:, 2]) < 1e-4)
        held = set(pairs[bad].ravel().tolist())
        base = set(map(tuple, pairs.tolist()))
        for _ in range(8):
            v[sorted(held)] = 0.0
            plus = set(map(tuple, c.neighbor_list(xyz + h * v).tolist()))
            minus = set(map(tuple, c.neighbor_list(xyz - h * v).tolist()))
            crossing = (plus ^ base) | (minus ^ base)
            if not crossing:
                break
            held |= {a for pr in crossing for a in pr}
        assert not crossing and len(held) < 0.02 * cfg.n
        kidx = c.sample_k(11, cfg.P)
        F = c.forces(xyz, q, kidx)          # builds the list at the unperturbed positions
        hv = c.hvp(xyz, q, kidx, v)
        fp = c.forces(xyz + h * v, q, kidx)
        fm = c.forces(xyz - h * v, q, kidx)
        E, phi, dg = c.energy(xyz, q, kidx)
        # energy along v: dE/dh = -F.v
        Ep, _, _ = c.energy(xyz + h * v, q, kidx)
        Em, _, _ = c.energy(xyz - h * v, q, kidx)
        u = np.random.default_rng(2).normal(size=xyz.shape)
        hu = c.hvp(xyz, q, kidx, u)
        # dE/dq along a random charge direction (E is quadratic in q)
        dq = np.random.default_rng(3).normal(size=q.shape)
        Eqp, _, _ = c.energy(xyz, q + 1e-3 * dq, kidx)
        Eqm, _, _ = c.energy(xyz, q - 1e-3 * dq, kidx)
    fd = -(fp - fm) / (2 * h)
    assert rel_rms(hv, fd) < 1e-6
    assert (Ep - Em) / (2 * h) == pytest.approx(-np.sum(F * v), rel=1e-6)
    assert (Eqp - Eqm) / 2e-3 == pytest.approx(np.dot(phi, dq), rel=1e-9)
    assert E == pytest.approx(0.5 * np.dot(q, phi), rel=1e-12)
    # Hessian symmetry of the conservative force (node motion of the
    # z-dependent limits breaks it only at the accuracy level)
    assert np.dot(u.ravel(), hv.ravel()) == pytest.approx(np.dot(v.ravel(), hu.ravel()), rel=1e-5)
    # dE/dgamma by differences
    for w in range(2):
        cfgp = cfg.__class__(**{**cfg.__dict__, "gamma_1": cfg.gamma_1 + (1e-5 if w == 0 else 0),
                                "gamma_2": cfg.gamma_2 + (1e-5 if w == 1 else 0)})
        cfgm = cfg.__class__(**{**cfg.__dict__, "gamma_1": cfg.gamma_1 - (1e-5 if w == 0 else 0),
                                "gamma_2": cfg.gamma_2 - (1e-5 if w == 1 else 0)})
        with ctx_for(cfgp, semantics=L.SEM_CORRECTED) as cp, ctx_for(cfgm, semantics=L.SEM_CORRECTED) as cm:
            d = (cp.energy(xyz, q, kidx)[0] - cm.energy(xyz, q, kidx)[0]) / 2e-5
        assert dg[w] == pytest.approx(d, rel=1e-6, abs=1e-9 * abs(E))


@pytest.mark.parametrize("sem", [0, L.SEM_CORRECTED])
@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_hvp_from_saved_sums(sem, precision):
    # qem_forces_energy_save keeps the pair quadrature's force-tangent node
    # sums and qem_hvp_saved combines them instead of running the tangent
    # quadrature: the forces are those of the call without saving bit for
    # bit, the product is qem_hvp's within the mixed tolerance (the sums are
    # taken at the forward's FP32 limits), and without valid sums (FP64, or
    # after a forces + energy call that does not save) qem_hvp_saved is
    # qem_hvp bit for bit
    cfg, xyz, q, c = _setup("slab_n10k", 2000, sem=sem, precision=precision)
    v = np.random.default_rng(11).normal(size=xyz.shape)
    with c:
        kidx = c.sample_k(17, cfg.P)
        F0, E0, phi0, dg0 = c.forces_energy(xyz, q, kidx)
        F, E, phi, dg = c.forces_energy(xyz, q, kidx, save=True)
        hs = c.hvp(xyz, q, kidx, v, saved=True)
        hd = c.hvp(xyz, q, kidx, v)
        c.forces_energy(xyz, q, kidx)
        hs2 = c.hvp(xyz, q, kidx, v, saved=True)
    assert np.array_equal(F, F0) and E == E0
    assert np.array_equal(phi, phi0) and np.array_equal(dg, dg0)
    assert np.array_equal(hs2, hd)
    if precision == "fp64":
        assert np.array_equal(hs, hd)
    else:
        # (not the same bits: the sums were taken at the forward's limits)
        assert not np.array_equal(hs, hd)
        assert rel_rms(hs, hd) < 1e-6


@pytest.mark.parametrize("g1,g2", [(0.0, 0.9), (0.9, -0.5), (0.95, 0.95)])
def test_mixed_hvp_against_fp64_across_gammas(g1, g2):
    # the mixed tangent loop takes its E/D-weighted node sums as 1/D-weighted
    # ones through gamma_1 gamma_2 E = D + 1; the FP64 tangent does not, so
    # the two agree only if that holds for a zero, a negative and a large
    # product of the contrasts
    base = scaled_config(CONFIGS["slab_n10k"], 1000)
    cfg = base.__class__(**{**base.__dict__, "gamma_1": g1, "gamma_2": g2})
    xyz, q = make_system(cfg, seed=4)
    v = np.random.default_rng(5).normal(size=xyz.shape)
    with ctx_for(cfg, semantics=L.SEM_CORRECTED, precision="mixed", iter_period=100000) as cm, \
            ctx_for(cfg, semantics=L.SEM_CORRECTED, precision="fp64", iter_period=100000) as cd:
        kidx = cm.sample_k(3, 50)
        hm = cm.hvp(xyz, q, kidx, v)
        hd = cd.hvp(xyz, q, kidx, v)
    assert rel_rms(hm, hd) < 1e-6


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_energy_and_hvp_beyond_one_long_range_pass(precision):
    # a batch of 8,292 k vectors runs its long-range passes in two parts of
    # at most 8,192 rows; the energy equals the oracle's for the whole batch
    # and the HVP's long range is the batch-weighted sum of its halves
    cfg, xyz, q, c = _setup("slab_n100", 100, sem=L.SEM_CORRECTED, precision=precision)
    v = np.random.default_rng(8).normal(size=xyz.shape)
    P = 8192 + 100
    with c:
        K, _, sK = c.kset()
        kidx = c.sample_k(21, P)
        E, phi, dg = c.energy(xyz, q, kidx)
        hv = c.hvp(xyz, q, kidx, v)
        hv0 = c.hvp(xyz, q, np.zeros(0, np.int64), v)          # short range alone
        hva = c.hvp(xyz, q, kidx[:5000], v)
        hvb = c.hvp(xyz, q, kidx[5000:], v)
    pairs = O.neighbor_pairs(xyz, cfg.Lx, cfg.Ly, cfg.r_cutoff)
    Eo, phio, dgo = O.energy(oracle_params(cfg, semantics=L.SEM_CORRECTED), xyz, q, pairs,
                             K[kidx], weight=sK / P, nthreads=16)
    tol = 1e-11 if precision == "fp64" else 1e-6
    assert E == pytest.approx(Eo, rel=tol)
    assert rel_rms(phi, phio) < tol
    assert np.allclose(dg, dgo, rtol=max(tol, 1e-11) * 10, atol=1e-13 * abs(Eo))
    want = hv0 + (5000 / P) * (hva - hv0) + ((P - 5000) / P) * (hvb - hv0)
    assert rel_rms(hv, want) < (1e-12 if precision == "fp64" else 1e-6)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_forces_energy_in_one_call(precision):
    # qem_forces_energy: the forces of qem_forces bit for bit and the energy,
    # dE/dq, dE/dgamma of qem_energy, in one call that advances the step
    # counter once (FP64: the two calls' own bits; mixed: one pair
    # quadrature for both, checked against the oracle's energy)
    cfg, xyz, q, c = _setup("slab_n10k", 2000, sem=L.SEM_CORRECTED, precision=precision)
    with c, ctx_for(cfg, semantics=L.SEM_CORRECTED, precision=precision,
                    iter_period=100000) as c2:
        K, _, sK = c.kset()
        kidx = c.sample_k(9, cfg.P)
        F, E, phi, dg = c.forces_energy(xyz, q, kidx)
        steps = c.short_state()[2]
        # a second call repeats every output bit for bit (fixed-order sums,
        # dE/dgamma's per-block partials included)
        F3, E3, phi3, dg3 = c.forces_energy(xyz, q, kidx)
        F2 = c2.forces(xyz, q, kidx)
        E2, phi2, dg2 = c2.energy(xyz, q, kidx)
    assert steps == 1
    assert np.array_equal(F3, F) and E3 == E
    assert np.array_equal(phi3, phi) and np.array_equal(dg3, dg)
    assert np.array_equal(F, F2)
    pairs = O.neighbor_pairs(xyz, cfg.Lx, cfg.Ly, cfg.r_cutoff)
    Eo, phio, dgo = O.energy(oracle_params(cfg, semantics=L.SEM_CORRECTED), xyz, q, pairs,
                             K[kidx], weight=sK / cfg.P, nthreads=16)
    if precision == "fp64":
        assert E == E2 and np.array_equal(phi, phi2) and np.array_equal(dg, dg2)
        assert E == pytest.approx(Eo, rel=1e-12)
    else:
        assert E == pytest.approx(Eo, rel=1e-6)
        assert rel_rms(phi, phio) < 1e-6
        assert np.allclose(dg, dgo, rtol=1e-5)
        assert E == pytest.approx(E2, rel=1e-6)


@pytest.mark.parametrize("n,P", [(2000, 40), (2000, 90), (2000, 200), (300, 90)])
def test_forces_energy_pass_b_variants(n, P):
    # the mixed fused call forms the long-range forces and potential in one
    # pass B walk; cover its row-group layouts (P = 40: one group of 40 rows;
    # 90: a last group of 26 rows, two chunks per wave; 200: a last group of
    # 8 rows, four chunks per wave) and, at n = 300, chunks of 16 atoms that
    # are not flat in e^{-kz} (the descending walk): the forces equal the
    # force call's bit for bit, the potential and E match the oracle.  E of
    # a neutral system is a small difference of terms q_i phi_i / 2 (here
    # ~1 against a sum of magnitudes ~300), so its bar is 1e-6 of that sum:
    # the mixed energy-only call misses E by up to 1.1e-6 relative too
    cfg, xyz, q, c = _setup("slab_n10k", n, sem=L.SEM_CORRECTED, precision="mixed")
    with c, ctx_for(cfg, semantics=L.SEM_CORRECTED, precision="mixed",
                    iter_period=100000) as c2:
        K, _, sK = c.kset()
        kidx = c.sample_k(13, P)
        F, E, phi, dg = c.forces_energy(xyz, q, kidx)
        F2 = c2.forces(xyz, q, kidx)
    assert np.array_equal(F, F2)
    pairs = O.neighbor_pairs(xyz, cfg.Lx, cfg.Ly, cfg.r_cutoff)
    Eo, phio, dgo = O.energy(oracle_params(cfg, semantics=L.SEM_CORRECTED), xyz, q, pairs,
                             K[kidx], weight=sK / P, nthreads=16)
    assert abs(E - Eo) <= 1e-6 * 0.5 * np.abs(q * phio).sum()
    assert rel_rms(phi, phio) < 1e-6
    assert np.allclose(dg, dgo, rtol=1e-5)
