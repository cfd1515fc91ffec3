"""Pins for the CPU FP64 oracle (CPU only).

Each check is chosen so that a plausible mistake in the restatement (a
dropped image, a wrong sign or index, a swapped operand) fails it:
  - Gauss-Legendre: exactness on polynomials (the reference's own test,
    test/Guassian_int.jl) and agreement with numpy's leggauss;
  - Green's functions: the image-charge series, the Gamma_2 / Gamma_1
    relation and dz_Gamma = d/dz_i by finite differences;
  - point-charge part of F_short_ij / F_short_i against an image-charge
    brute force (the reference's active test, test/Force_short_test.jl:1-46);
  - the full short-range pair force (gaussian part included, limits that
    bound the decay) against point-minus-smeared image charges;
  - quadrature limits against their defining decay conditions;
  - neighbour pairs against a numpy O(N^2) enumeration;
  - K set: sum(P) = 1 (test/k_gen.jl), closed-form enumeration, sum_K;
  - per-k long-range force = -grad of the image-series energy (finite
    differences), Fl_sort and the k = 0 term against O(N^2) direct sums.
"""
import numpy as np
import pytest

import pins
from helpers import O, four_atom, two_atom


# ---- Gauss-Legendre ------------------------------------------------------------

def test_gl_reference_exactness():
    # test/Guassian_int.jl:14-47: GL(20) integrates 1, x, k x^n exactly
    rng = np.random.default_rng(0)
    x, w = O.legendre(20)
    for _ in range(10):
        a, b = np.sort(rng.random(2))
        X = (b + a) / 2 + (b - a) * x / 2
        W = (b - a) * w / 2
        assert np.isclose(W.sum(), b - a, rtol=1e-14)
        assert np.isclose((W * X).sum(), (b * b - a * a) / 2, rtol=1e-13)
        k, n = rng.integers(1, 6, 2)
        assert np.isclose((W * k * X ** n).sum(), k * (b ** (n + 1) - a ** (n + 1)) / (n + 1),
                          rtol=1e-13)


@pytest.mark.parametrize("n", [1, 2, 5, 30, 40, 50, 64])
def test_gl_matches_numpy_and_degree(n):
    x, w = O.legendre(n)
    xr, wr = np.polynomial.legendre.leggauss(n)
    assert np.max(np.abs(x - xr)) < 4e-16 * max(1, n / 8)
    assert np.max(np.abs(w - wr)) < 4e-15
    if n in (30, 50):   # 40-digit values: exact to the last bit
        import mpmath as mp
        mp.mp.dps = 40
        for xi, wi in zip(x, w):
            t = mp.findroot(lambda s: mp.legendre(n, s), mp.mpf(xi))
            dp = n * (mp.legendre(n - 1, t) - t * mp.legendre(n, t)) / (1 - t * t)
            assert abs(xi - float(t)) <= 1.2e-16
            assert abs(wi - float(2 / ((1 - t * t) * dp * dp))) <= 2.3e-16 * wi
    assert np.all(np.diff(x) > 0)
    for deg in range(2 * n):   # exact up to degree 2n - 1
        exact = (1 - (-1) ** (deg + 1)) / (deg + 1)
        assert abs((w * x ** deg).sum() - exact) < 1e-13


# ---- Green's functions -----------------------------------------------------------

def test_green_kat_and_gamma2_relation():
    # green_norm_njit.gamma(1,1,1,1,1) == -1 (test/greens_funtion.jl:7): with
    # eps_1 = eps_2 = eps_3 the contrasts vanish and 1/green_d = -1
    e = O.elem_ij(0.0, 0.0, 0.4, 0.7, 1.0, 1.0, 1.0, 1e-4)
    assert O.gamma("Gamma_1", 1.0, e, 0) == pytest.approx(-0.5 * sum(
        b * np.exp(-1.0 * a) for a, b in zip(e.a, e.b)))
    rng = np.random.default_rng(1)
    for _ in range(50):
        g1, g2 = rng.uniform(-0.95, 0.95, 2)
        Lz = rng.uniform(0.5, 10)
        zi, zj = rng.uniform(0.01, 0.99, 2) * Lz
        k = rng.uniform(0.01, 5)
        e = O.elem_ij(g1, g2, zi, zj, 1.0, Lz, 1.0, 1e-6)
        f = g1 * g2 * np.exp(-2 * k * Lz)
        for l in range(5):
            assert O.gamma("Gamma_2", k, e, l) == pytest.approx(f * O.gamma("Gamma_1", k, e, l),
                                                                rel=1e-13, abs=1e-300)
            assert O.gamma("dz_Gamma_2", k, e, l) == pytest.approx(
                f * O.gamma("dz_Gamma_1", k, e, l), rel=1e-13, abs=1e-300)


def test_green_equals_image_series():
    rng = np.random.default_rng(2)
    for _ in range(40):
        g1, g2 = rng.uniform(-0.9, 0.9, 2)
        Lz = rng.uniform(1, 10)
        zi, zj = rng.uniform(0.02, 0.98, 2) * Lz
        k = rng.uniform(0.05, 3)
        e = O.elem_ij(g1, g2, zi, zj, 1.0, Lz, 1.0, 1e-6)
        ref = pins.green_series_k(k, zi, zj, g1, g2, Lz)
        assert -O.gamma("Gamma_1", k, e, 0) == pytest.approx(ref, rel=1e-12)
        es = O.elem_i(g1, g2, zi, Lz, 1.0, 1e-6)
        assert -O.gamma("Gamma_1", k, es, 0) == pytest.approx(
            pins.green_series_k(k, zi, zi, g1, g2, Lz), rel=1e-12)


def test_dz_gamma_is_z_derivative():
    rng = np.random.default_rng(3)
    h = 1e-6
    for _ in range(30):
        g1, g2 = rng.uniform(-0.9, 0.9, 2)
        Lz = rng.uniform(1, 5)
        zi, zj = rng.uniform(0.05, 0.95, 2) * Lz
        if abs(zi - zj) < 1e-3:
            continue
        k = rng.uniform(0.05, 3)
        e = O.elem_ij(g1, g2, zi, zj, 1.0, Lz, 1.0, 1e-6)
        for l in range(1, 5):
            ep = O.elem_ij(g1, g2, zi + h, zj, 1.0, Lz, 1.0, 1e-6)
            em = O.elem_ij(g1, g2, zi - h, zj, 1.0, Lz, 1.0, 1e-6)
            fd = (O.gamma("Gamma_1", k, ep, l) - O.gamma("Gamma_1", k, em, l)) / (2 * h)
            assert O.gamma("dz_Gamma_1", k, e, l) == pytest.approx(fd, rel=1e-6, abs=1e-10)


def test_quadrature_limits_bound_the_decay():
    # k_f1 solves exp(-k (2Lz + a)) = acc; k_f2 solves exp(-k^2/(4a) - k a) = acc
    rng = np.random.default_rng(4)
    for _ in range(20):
        Lz, alpha = rng.uniform(1, 20), rng.uniform(0.1, 3)
        acc = 10.0 ** rng.uniform(-12, -3)
        zi, zj = rng.uniform(0.05, 0.95, 2) * Lz
        e = O.elem_ij(0.5, 0.5, zi, zj, 1.0, Lz, alpha, acc)
        for l in range(4):
            a = e.a[l]
            assert np.exp(-e.k_f1[l] * (2 * Lz + a)) == pytest.approx(acc, rel=1e-10)
            K = e.k_f2[l]
            assert np.exp(-K * K / (4 * alpha) - K * a) == pytest.approx(acc, rel=1e-8)


# ---- short range ------------------------------------------------------------------

def test_point_part_matches_image_charges():
    # test/Force_short_test.jl:1-46 (single = true, alpha = 1, r_c = 1, N_t = 30)
    ri, rj, box = two_atom()
    rng = np.random.default_rng(5)
    for _ in range(10):
        g1, g2 = rng.random(2)
        p = O.params(box, 1.0, 1.0, g1, g2, 1.0, 1e-6, 30)
        f, _ = O.F_short_ij(p, 1, 1, ri, rj, single=True)
        fs = O.F_short_i(p, 1, ri, single=True)
        ref = pins.icm_force(ri, rj, 1, 1, g1, g2, box[2])
        ref_self = pins.icm_force(ri, ri, 1, 1, g1, g2, box[2])
        assert np.allclose(f, ref, atol=1e-5, rtol=0)          # the reference's bound
        assert np.allclose(f, ref, rtol=2e-6, atol=0)          # what the oracle reaches
        assert abs(fs[2] - ref_self[2]) < 1e-5
        assert fs[0] == 0 and fs[1] == 0


@pytest.mark.parametrize("g1,g2,alpha", [(0.0, 0.0, 1.0), (0.5, -0.4, 1.0), (0.7, 0.6, 2.5)])
def test_full_pair_force_matches_point_minus_smeared_images(g1, g2, alpha):
    # with the limits that bound each integrand's decay (SWAP_LIMITS) the
    # gaussian integrals are converged; NEWTON_IMAGES gives the j-side z force
    ri = np.array([5.0, 5.0, 0.9]); rj = np.array([5.8, 4.4, 0.5])
    p = O.params((10, 10, 2.0), 3.0, alpha, g1, g2, 1.0, 1e-14, 120, O.SEM_CORRECTED)
    f, fzj = O.F_short_ij(p, 1, 1, ri, rj)
    layers = 40 if (g1 or g2) else 0
    ref = pins.short_pair_brute(ri, rj, alpha, g1, g2, 2.0, layers)
    ref_j = pins.short_pair_brute(rj, ri, alpha, g1, g2, 2.0, layers)
    assert np.allclose(f, ref, rtol=1e-8, atol=1e-12)
    assert fzj == pytest.approx(ref_j[2], rel=1e-8, abs=1e-12)


def test_as_written_semantics_details():
    # raw direction (src/Force_short.jl:57-58): a pair across the periodic
    # boundary points along the raw difference
    p0 = O.params((10, 10, 2), 1.0, 1.0, 0.3, 0.2, 1.0, 1e-6, 30, 0)
    p2 = O.params((10, 10, 2), 1.0, 1.0, 0.3, 0.2, 1.0, 1e-6, 30, O.SEM_MINIMAGE_DIR)
    ri = np.array([0.1, 5.0, 1.0]); rj = np.array([9.8, 5.0, 1.2])   # min-image dx = +0.3
    f0, fz0 = O.F_short_ij(p0, 1, 1, ri, rj)
    f2, _ = O.F_short_ij(p2, 1, 1, ri, rj)
    assert f0[0] == pytest.approx(f2[0] * (0.1 - 9.8) / 0.3, rel=1e-12)
    assert f0[2] == pytest.approx(f2[2], rel=1e-14)
    assert fz0 == -f0[2]                      # Newton's third law on all of F_ij
    # rho == 0: no xy force, z force still there (src/Force_short.jl:46,59-62)
    f, _ = O.F_short_ij(p0, 1, -1, np.array([3.0, 3.0, 0.5]), np.array([3.0, 3.0, 1.5]))
    assert f[0] == 0 and f[1] == 0 and f[2] != 0
    # beyond the cutoff the pair contributes nothing (2-point neighbour list)
    f, fz = O.F_short_ij(p0, 1, 1, np.array([1.0, 1.0, 1.0]), np.array([2.5, 1.0, 1.0]))
    assert np.all(f == 0) and fz == 0


def test_neighbor_pairs_match_numpy():
    rng = np.random.default_rng(6)
    for n, L, rc in [(300, 20.0, 5.0), (500, 30.0, 2.5), (50, 7.0, 3.4)]:
        xyz = np.column_stack([rng.uniform(0, L, n), rng.uniform(0, L, n), rng.uniform(0.1, 1, n)])
        got = set(map(tuple, O.neighbor_pairs(xyz, L, L, rc).tolist()))
        assert got == pins.neighbor_pairs_numpy(xyz, L, L, rc)


# ---- K set -----------------------------------------------------------------------------

def test_kset_reference_sum_and_enumeration():
    K, P, sK, (mx, my) = O.kset(1.0, 1.0, 0.1, 1e-6)     # test/k_gen.jl:3-5
    assert float(np.sum(P, dtype=np.float64)) == pytest.approx(1.0, abs=1e-6)
    for (Lx, Ly, alpha) in [(20.0, 20.0, 1.0), (100.0, 37.0, 0.5), (7.0, 11.0, 2.0)]:
        K, P, sK, (mx, my) = O.kset(Lx, Ly, alpha, 1e-6)
        kc = 2 * np.sqrt(-4 * alpha * np.log(1e-6))
        assert mx == int(np.ceil(kc * Lx / (2 * np.pi))) + 1
        assert len(K) == (2 * mx + 1) * (2 * my + 1) - 1
        # closed form of the enumeration: m_x outer, m_y inner, (0,0) skipped
        M = np.array([(a, b) for a in range(-mx, mx + 1) for b in range(-my, my + 1)
                      if (a, b) != (0, 0)])
        assert np.array_equal(K[:, 0], M[:, 0] * 6.283185307179586 / Lx)
        assert np.array_equal(K[:, 1], M[:, 1] * 6.283185307179586 / Ly)
        w = np.exp(-K[:, 2] ** 2 / (4 * alpha))
        assert sK == pytest.approx(w.sum(), rel=1e-12)        # (1+S_x)(1+S_y) - 1
        assert np.allclose(P, (w / sK).astype(np.float32), rtol=0, atol=0)
        assert float(np.sum(P, dtype=np.float64)) == pytest.approx(1.0, abs=2e-6)


# ---- long range ------------------------------------------------------------------------

def _slab(n, seed, Lz=2.0):
    rng = np.random.default_rng(seed)
    X = np.column_stack([rng.uniform(0, 10, n), rng.uniform(0, 10, n), rng.uniform(0.1, Lz - 0.1, n)])
    return X, rng.normal(size=n)


@pytest.mark.parametrize("seed", [0, 1])
def test_long_per_k_is_minus_gradient_of_image_energy(seed):
    X, q = _slab(7, seed)
    rng = np.random.default_rng(10 + seed)
    g1, g2 = rng.uniform(-0.8, 0.8, 2)
    kx, ky = rng.uniform(-1.2, 1.2, 2)
    k3 = np.array([kx, ky, np.hypot(kx, ky)])
    p = O.params((10.0, 10.0, 2.0), 1.0, 1.0, g1, g2, 1.0, 1e-6, 30)
    F = O.F_l_total(k3, X, q, O.sortperm_z(X), p) * 100.0   # times Lx Ly eps_0
    G = pins.fd_gradient(lambda Y: pins.k_energy(Y, q, k3, g1, g2, 2.0), X)
    assert np.allclose(F, -G, rtol=1e-6, atol=1e-8)


def test_Fl_sort_and_k0_match_direct_sums():
    X, q = _slab(40, 3)
    zl = O.sortperm_z(X)
    k3 = np.array([0.8, -0.5, np.hypot(0.8, -0.5)])
    th = k3[0] * X[:, 0] + k3[1] * X[:, 1]
    dth = th[:, None] - th[None, :]
    dz = X[:, 2][:, None] - X[:, 2][None, :]
    E = np.exp(-k3[2] * np.abs(dz))
    np.fill_diagonal(E, 0.0)
    xy = (np.sin(dth) * E) @ q
    z = (np.cos(dth) * np.sign(dz) * E) @ q
    F = O.Fl_sort(k3, X, q, zl)
    assert np.allclose(F[:, 0], q * k3[0] / (2 * k3[2]) * xy, rtol=1e-12, atol=1e-14)
    assert np.allclose(F[:, 1], q * k3[1] / (2 * k3[2]) * xy, rtol=1e-12, atol=1e-14)
    assert np.allclose(F[:, 2], q / 2 * z, rtol=1e-12, atol=1e-14)
    F0 = O.Flz_sort_k0(q, zl)
    assert np.allclose(F0[:, 2], q * (np.sign(dz) @ q), rtol=1e-12, atol=1e-13)


def test_Fl_sort_subtractive_suffix_loses_precision():
    # The reference forms the e^{-kz} suffix sums by subtracting from the
    # total (src/Force_long.jl:83-84,97-98).  At k (z_max - z_min) ~ 40 that
    # leaves no correct digit for the top atoms; the oracle sums from the top.
    X, q = _slab(60, 4, Lz=20.0)
    zl = O.sortperm_z(X)
    k3 = np.array([2.5, 0.5, np.hypot(2.5, 0.5)])
    th = k3[0] * X[:, 0] + k3[1] * X[:, 1]
    dz = X[:, 2][:, None] - X[:, 2][None, :]
    E = np.exp(-k3[2] * np.abs(dz)); np.fill_diagonal(E, 0.0)
    z = (np.cos(th[:, None] - th[None, :]) * np.sign(dz) * E) @ q * q / 2
    stable = O.Fl_sort(k3, X, q, zl)[:, 2]
    subtr = O.Fl_sort(k3, X, q, zl, subtractive=True)[:, 2]
    assert np.allclose(stable, z, rtol=1e-10, atol=1e-12)
    assert np.max(np.abs(subtr - z)) > 1e3 * np.max(np.abs(z))


def test_four_atom_reference_geometry_zlist():
    xyz, q, box, zl = four_atom()
    assert np.array_equal(O.sortperm_z(xyz), zl)   # z_list = [2, 1, 3, 4] (1-based)


def test_sortperm_is_stable():
    X = np.array([[0, 0, 1.0], [0, 0, 0.5], [0, 0, 1.0], [0, 0, 0.5], [0, 0, 0.7]])
    assert O.sortperm_z(X).tolist() == [1, 3, 4, 0, 2]


def test_batch_weights_are_unbiased():
    # E[sum_K / P * sum_p F(k_p)] over k_p ~ P_k equals sum_k exp(-k^2/4a) F(k):
    # the batch weight sum_K times P_k reproduces exp(-k^2/(4 alpha)) up to
    # the Float32 rounding of P_k
    K, P, sK, _ = O.kset(10.0, 10.0, 1.0, 1e-6)
    w = np.exp(-K[:, 2] ** 2 / 4.0)
    normal = w / sK > 1e-37          # Float32 P_k underflows below that
    assert np.allclose(sK * P[normal].astype(np.float64), w[normal], rtol=1e-7, atol=0)
    assert np.all(P[~normal] < 1e-37) and w[~normal].sum() < 1e-30 * w.sum()
    X, q = _slab(6, 5)
    p = O.params((10.0, 10.0, 2.0), 1.0, 1.0, 0.4, 0.6, 1.0, 1e-6, 30)
    sel = np.argsort(-w)[:200]
    full = O.long_forces(p, X, q, K[sel], full_sum=True)
    batch = O.long_forces(p, X, q, K[sel], weight=1.0)  # then weight per k by sK * P
    per_k = np.zeros_like(full)
    for t in sel:
        per_k += sK * float(P[t]) * (O.long_forces(p, X, q, K[[t]], weight=1.0)
                                     - O.long_forces(p, X, q, K[:0], weight=1.0))
    k0 = O.long_forces(p, X, q, K[:0], weight=1.0)
    assert np.allclose(full, k0 + per_k, rtol=1e-6, atol=1e-9)
    assert np.isfinite(batch).all()


# ---- the short range for a subset of rows ------------------------------------

@pytest.mark.parametrize("sem", [0, O.SEM_CORRECTED])
def test_short_forces_rows_equal_the_pair_list_sum(sem):
    # qo_short_forces_rows finds each row's pairs by a scan and must give the
    # rows of qo_short_forces over the O(N^2) pair list bit for bit (same
    # pairs, same order of the adds); atoms on the periodic boundary and at
    # exactly r_cutoff included
    from helpers import CONFIGS, make_system, oracle_params, scaled_config
    cfg = scaled_config(CONFIGS["slab_n10k"], 300)
    xyz, q = make_system(cfg, seed=4)
    xyz[0, :2] = [0.0, 0.0]
    xyz[1, :2] = [cfg.r_cutoff, 0.0]
    xyz[2, :2] = [cfg.Lx - 1e-12, cfg.Ly - 1e-12]
    p = oracle_params(cfg, semantics=sem)
    pairs = O.neighbor_pairs(xyz, cfg.Lx, cfg.Ly, cfg.r_cutoff)
    assert (0, 1) in set(map(tuple, pairs.tolist()))
    full = O.short_forces(p, xyz, q, pairs, nthreads=4)
    rows = np.array([0, 1, 2, 7, 150, 299])
    got = O.short_forces_rows(p, xyz, q, rows, nthreads=4)
    assert np.array_equal(got, full[rows])
