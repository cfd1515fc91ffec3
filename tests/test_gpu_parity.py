"""GPU parity: libqem_hip.so (through its C ABI) against the FP64 oracle.

Bars: integer/index outputs bit-exact; FP64 forces within 1e-12 relative
(rel-RMS over all components, and max-abs relative for small cases);
mixed precision within 1e-6 rel-RMS of the FP64 oracle.
"""
import numpy as np
import pytest

from helpers import (CONFIGS, L, O, ctx_for, four_atom, lib_params, make_system, max_rel,
                     oracle_params, rel_rms, scaled_config, two_atom)

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
MIXED_TOL = 1e-6

SEMS = [0, L.SEM_SWAP_LIMITS, L.SEM_CORRECTED]


def _pairs_set(p):
    return set(map(tuple, np.asarray(p).tolist()))


# ---- neighbour list (bit-exact pair set) ------------------------------------

@pytest.mark.parametrize("name,n", [("slab_n100", 100), ("slab_n10k", 3000),
                                    ("slab_n100k", 4000)])
def test_neighbor_list_matches_bruteforce(name, n):
    cfg = scaled_config(CONFIGS[name], n) if n != CONFIGS[name].n else CONFIGS[name]
    xyz, q = make_system(cfg, seed=3)
    with ctx_for(cfg, rbe_p=0) as c:
        got = c.neighbor_list(xyz)
    ref = O.neighbor_pairs(xyz, cfg.Lx, cfg.Ly, cfg.r_cutoff)
    assert got.shape == ref.shape
    assert np.all(got[:, 0] < got[:, 1])
    assert _pairs_set(got) == _pairs_set(ref)


def test_neighbor_list_small_box_and_boundary():
    # fewer than 3 cells per side, atoms on the periodic boundary, exact ties
    from qem4molly_amd.workloads import SlabConfig
    cfg = SlabConfig("tiny", 40, 10.0, 7.0, 3.0, 0.3, -0.2, 1.0, 3.4, 8, 4)
    rng = np.random.default_rng(11)
    xyz = np.column_stack([rng.uniform(0, 10, 40), rng.uniform(0, 7, 40),
                           rng.uniform(0.2, 2.8, 40)])
    xyz[0, :2] = [0.0, 0.0]
    xyz[1, :2] = [10.0 - 1e-12, 7.0 - 1e-12]
    xyz[2, :2] = [3.4, 0.0]            # exactly r_c from atom 0
    xyz[3, :2] = [-1.0, 8.0]           # outside [0, L): wrapped by the minimum image
    with ctx_for(cfg, rbe_p=0) as c:
        got = c.neighbor_list(xyz)
    ref = O.neighbor_pairs(xyz, cfg.Lx, cfg.Ly, cfg.r_cutoff)
    assert _pairs_set(got) == _pairs_set(ref)
    assert (0, 2) in _pairs_set(ref)


def test_neighbor_list_empty_and_single():
    cfg = CONFIGS["slab_n100"]
    with ctx_for(cfg, rbe_p=0) as c:
        assert c.neighbor_list(np.zeros((0, 3))).shape == (0, 2)
        assert c.neighbor_list(np.array([[1.0, 1.0, 1.0]])).shape == (0, 2)


# ---- short range -------------------------------------------------------------

@pytest.mark.parametrize("sem", SEMS)
def test_short_forces_slab_n100_fp64(sem):
    cfg = CONFIGS["slab_n100"]
    xyz, q = make_system(cfg)
    p = oracle_params(cfg, semantics=sem)
    pairs = O.neighbor_pairs(xyz, cfg.Lx, cfg.Ly, cfg.r_cutoff)
    ref = O.short_forces(p, xyz, q, pairs, nthreads=8)
    with ctx_for(cfg, semantics=sem) as c:
        got = c.short_forces(xyz, q)
        got2 = c.short_forces_pairs(xyz, q, pairs)
    assert rel_rms(got, ref) < FP64_TOL
    assert max_rel(got, ref) < 1e-11
    assert rel_rms(got2, ref) < FP64_TOL


@pytest.mark.parametrize("sem", [0, L.SEM_CORRECTED])
def test_short_forces_scaled_n10k_fp64(sem):
    cfg = scaled_config(CONFIGS["slab_n10k"], 1500)
    xyz, q = make_system(cfg, seed=5)
    p = oracle_params(cfg, semantics=sem)
    pairs = O.neighbor_pairs(xyz, cfg.Lx, cfg.Ly, cfg.r_cutoff)
    ref = O.short_forces(p, xyz, q, pairs, nthreads=16)
    with ctx_for(cfg, semantics=sem) as c:
        got = c.short_forces(xyz, q)
    assert rel_rms(got, ref) < FP64_TOL


def test_F_short_ij_reference_geometry():
    ri, rj, box = two_atom()
    from qem4molly_amd.workloads import SlabConfig
    for (g1, g2) in [(0.9, 0.8), (0.3, 0.1), (-0.7, 0.95)]:
        for alpha, single in [(1.0, True), (0.1, False), (1.0, False)]:
            cfg = SlabConfig("two", 2, *box, g1, g2, alpha, 1.0, 30, 1)
            for sem in SEMS:
                ref, refj = O.F_short_ij(oracle_params(cfg, semantics=sem), 1, 1, ri, rj,
                                         single=single)
                with ctx_for(cfg, semantics=sem, rbe_p=0) as c:
                    got, gotj = c.short_pair([1.0], [1.0], [ri], [rj], single=single)
                    self_ref = O.F_short_i(oracle_params(cfg, semantics=sem), 1, ri, single)
                    self_got = c.short_self([1.0], [ri], single=single)
                assert max_rel(got[0], ref) < 1e-13
                assert abs(gotj[0] - refj) <= 1e-13 * max(abs(refj), 1e-300)
                assert max_rel(self_got[0], self_ref) < 1e-13


def test_short_pair_beyond_cutoff_is_zero():
    ri, rj, box = two_atom()
    from qem4molly_amd.workloads import SlabConfig
    cfg = SlabConfig("two", 2, *box, 0.5, 0.5, 1.0, 0.2, 30, 1)
    with ctx_for(cfg, rbe_p=0) as c:
        f, fz = c.short_pair([1.0], [1.0], [ri], [rj])
    assert np.all(f == 0.0) and fz[0] == 0.0


def test_short_step_counter_and_stale_list():
    # Q4: list rebuilt only every iter_period calls; pairs that left the
    # cutoff contribute zero, pairs that entered are missed until the rebuild
    cfg = CONFIGS["slab_n100"]
    xyz, q = make_system(cfg)
    p = oracle_params(cfg)
    rng = np.random.default_rng(9)
    with ctx_for(cfg, iter_period=3) as c:
        pairs0 = O.neighbor_pairs(xyz, cfg.Lx, cfg.Ly, cfg.r_cutoff)
        c.short_forces(xyz, q)
        assert c.short_state()[1:] == (1, 1)
        moved = xyz.copy()
        moved[:, :2] += rng.normal(scale=0.7, size=(cfg.n, 2))
        got = c.short_forces(moved, q)         # step 2: stale list
        ref = O.short_forces(p, moved, q, pairs0)
        assert c.short_state()[1:] == (1, 2)
        assert rel_rms(got, ref) < FP64_TOL
        c.short_forces(moved, q)               # step 3
        got4 = c.short_forces(moved, q)        # step 4: rebuild
        assert c.short_state()[1:] == (2, 4)
        ref4 = O.short_forces(p, moved, q, O.neighbor_pairs(moved, cfg.Lx, cfg.Ly, cfg.r_cutoff))
        assert rel_rms(got4, ref4) < FP64_TOL


# ---- long range --------------------------------------------------------------

def test_sortperm_bit_exact_with_ties():
    cfg = CONFIGS["slab_n100"]
    xyz, q = make_system(cfg, seed=4)
    xyz[10:20, 2] = 2.0          # ties resolved by index
    xyz[50:53, 2] = 0.75
    with ctx_for(cfg) as c:
        got = c.sortperm_z(xyz)
    assert np.array_equal(got, O.sortperm_z(xyz))


@pytest.mark.parametrize("part", ["nonsort", "sort", "self", "k0"])
def test_long_parts_four_atom(part):
    xyz, q, box, zl = four_atom()
    rng = np.random.default_rng(2)
    g = greens = None  # noqa: F841
    from qem4molly_amd import interactions as I
    for trial in range(5):
        g1, g2 = rng.random(2)
        kx, ky = rng.random(2) + 0.1
        k3 = np.array([kx, ky, np.hypot(kx, ky)])
        el = I.greens_element_init(g1, g2, box[2], 1.0)
        atoms = [I.Atom(charge=v) for v in q]
        assert np.array_equal(O.sortperm_z(xyz), zl)
        if part == "nonsort":
            ref = O.Fl_non_sort(k3, xyz, q, g1, g2, box[2]); got = I.Fl_non_sort(k3, atoms, xyz, el)
        elif part == "sort":
            ref = O.Fl_sort(k3, xyz, q, zl); got = I.Fl_sort(k3, atoms, zl, xyz, el)
        elif part == "self":
            ref = O.Flz_self(k3, xyz, q, g1, g2, box[2]); got = I.Flz_self(k3, atoms, xyz, el)
        else:
            ref = O.Flz_sort_k0(q, zl); got = I.Flz_sort_k0(atoms, zl, xyz, el)
        assert max_rel(got, ref) < 1e-13, (part, trial)


@pytest.mark.parametrize("precision,tol", [("fp64", FP64_TOL), ("mixed", MIXED_TOL)])
def test_long_forces_slab_n100(precision, tol):
    cfg = CONFIGS["slab_n100"]
    xyz, q = make_system(cfg)
    with ctx_for(cfg, precision=precision) as c:
        K, Pk, sK = c.kset()
        kidx = c.sample_k(0, cfg.P)
        got = c.long_forces(xyz, q, kidx)
    ref = O.long_forces(oracle_params(cfg), xyz, q, K[kidx], weight=sK / cfg.P, nthreads=8)
    assert rel_rms(got, ref) < tol


def test_long_forces_scaled_n10k_fp64():
    cfg = scaled_config(CONFIGS["slab_n10k"], 3000)
    xyz, q = make_system(cfg, seed=2)
    with ctx_for(cfg) as c:
        K, Pk, sK = c.kset()
        kidx = c.sample_k(1, cfg.P)
        got = c.long_forces(xyz, q, kidx)
    ref = O.long_forces(oracle_params(cfg), xyz, q, K[kidx], weight=sK / cfg.P, nthreads=16)
    assert rel_rms(got, ref) < FP64_TOL


@pytest.mark.parametrize("P", [10, 30, 64, 90, 150])
@pytest.mark.parametrize("n", [300, 2500])
def test_long_forces_row_groups_fp64(n, P):
    """k batches whose last group of 64 rows is full (64), holds 17-32 rows (two
    chunks per wave) or at most 16 (four chunks per wave), one group among
    several or alone; at n = 300 the chunks are walked a second time from the
    top, at n = 2500 most are flat."""
    cfg = scaled_config(CONFIGS["slab_n10k"], n)
    xyz, q = make_system(cfg, seed=3)
    with ctx_for(cfg) as c:
        K, Pk, sK = c.kset()
        kidx = c.sample_k(11, P)
        got = c.long_forces(xyz, q, kidx)
    ref = O.long_forces(oracle_params(cfg), xyz, q, K[kidx], weight=sK / P, nthreads=16)
    assert rel_rms(got, ref) < FP64_TOL


@pytest.mark.parametrize("P", [30, 90, 150])
def test_long_forces_row_groups_mixed(P):
    cfg = scaled_config(CONFIGS["slab_n100k"], 3000)
    xyz, q = make_system(cfg, seed=4)
    with ctx_for(cfg, precision="mixed") as c:
        K, Pk, sK = c.kset()
        kidx = c.sample_k(12, P)
        got = c.long_forces(xyz, q, kidx)
    ref = O.long_forces(oracle_params(cfg), xyz, q, K[kidx], weight=sK / P, nthreads=16)
    assert rel_rms(got, ref) < MIXED_TOL


# ---- both halves, mixed precision ---------------------------------------------

@pytest.mark.parametrize("name,n", [("slab_n100", 100), ("slab_n10k", 1500),
                                    ("slab_n100k", 2000)])
def test_forces_mixed_vs_fp64_oracle(name, n):
    cfg = scaled_config(CONFIGS[name], n) if n != CONFIGS[name].n else CONFIGS[name]
    xyz, q = make_system(cfg, seed=1)
    p = oracle_params(cfg)
    pairs = O.neighbor_pairs(xyz, cfg.Lx, cfg.Ly, cfg.r_cutoff)
    with ctx_for(cfg, precision="mixed") as c:
        K, Pk, sK = c.kset()
        kidx = c.sample_k(7, cfg.P)
        got = c.forces(xyz, q, kidx)
        got_s = c.short_forces(xyz, q)
    ref_s = O.short_forces(p, xyz, q, pairs, nthreads=16)
    ref_l = O.long_forces(p, xyz, q, K[kidx], weight=sK / cfg.P, nthreads=16)
    e_short = rel_rms(got_s, ref_s)
    e_total = rel_rms(got, ref_s + ref_l)
    print(f"{name} n={n}: mixed rel-RMS short {e_short:.3e} total {e_total:.3e}")
    assert e_short < MIXED_TOL
    assert e_total < MIXED_TOL


def test_forces_fp64_slab_n100_combined():
    cfg = CONFIGS["slab_n100"]
    xyz, q = make_system(cfg)
    p = oracle_params(cfg)
    pairs = O.neighbor_pairs(xyz, cfg.Lx, cfg.Ly, cfg.r_cutoff)
    with ctx_for(cfg) as c:
        K, Pk, sK = c.kset()
        kidx = c.sample_k(0, cfg.P)
        got = c.forces(xyz, q, kidx)
    ref = O.short_forces(p, xyz, q, pairs) + O.long_forces(p, xyz, q, K[kidx],
                                                           weight=sK / cfg.P)
    assert rel_rms(got, ref) < FP64_TOL


def test_bessel_table_accuracy():
    from scipy import special
    cfg = CONFIGS["slab_n10k"]
    with ctx_for(cfg, precision="mixed") as c:
        x = np.linspace(0.0, 42.0, 20001)
        j0, j1 = c.bessel01_table(x)
    xf = x.astype(np.float32).astype(np.float64)   # the kernel sees FP32 arguments
    assert np.max(np.abs(j0 - special.j0(xf))) < 4e-7
    assert np.max(np.abs(j1 - special.j1(xf))) < 4e-7


def test_bessel_table_fp64_accuracy():
    # FP64 contexts tabulate J0 with degree-7 polynomials per 1/16 (J1 from
    # the derivative)
    from scipy import special
    cfg = CONFIGS["slab_n10k"]
    with ctx_for(cfg, precision="fp64") as c:
        x = np.linspace(0.0, 42.0, 40001)
        j0, j1 = c.bessel01_table(x)
    assert np.max(np.abs(j0 - special.j0(x))) < 5e-15
    assert np.max(np.abs(j1 - special.j1(x))) < 5e-15


def test_fp64_without_table_uses_device_library():
    # a Bessel argument range beyond the FP64 table (r_c sqrt(4 alpha ln(1/acc))
    # > 64): the FP64 kernels call the device library's j0/j1 instead
    from qem4molly_amd.workloads import SlabConfig
    cfg = SlabConfig("wide", 60, 20.0, 20.0, 4.0, 0.7, -0.3, 1.0, 9.5, 20, 10)
    xyz, q = make_system(cfg, seed=4)
    p = oracle_params(cfg)
    with ctx_for(cfg, precision="fp64") as c:
        with pytest.raises(L.QEMError):
            c.bessel01_table(np.array([1.0]))
        got = c.short_forces(xyz, q)
        K, P, sK = c.kset()
        pairs = c.neighbor_list(xyz)
    ref = O.short_forces(p, xyz, q, pairs)
    assert rel_rms(got, ref) < 1e-12
