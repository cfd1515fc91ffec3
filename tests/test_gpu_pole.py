"""gamma_1 gamma_2 > 1 on the GPU: the pole-splitting rule
(src/python_lib/Force_short_norm_njit.py) against the oracle (itself pinned
against scipy's Cauchy principal value), the as-written quadrature across the
pole in FP64, and the entry points that refuse such contrasts."""
import numpy as np
import pytest

from helpers import L, O, SlabConfig, make_system, oracle_params, rel_rms, lib_params

pytestmark = pytest.mark.gpu

# gamma 2.0 / 2.0: k0 = ln 4 / (2 Lz) = 0.347 at Lz = 2; N_t = 60 gives the
# point set 6 + 10 + 6 + 48 nodes and the gaussian set 3 + 10 + 3 + 54
POLE = SlabConfig("pole_n150", 150, 20.0, 20.0, 2.0, 2.0, 2.0, 0.5, 4.0, 60, 40)
POLE_NEG = SlabConfig("pole_neg_n150", 150, 20.0, 20.0, 2.5, -2.5, -1.6, 0.4, 4.0, 60, 40)
EPS = {POLE.name: np.log(4.0) / 4.0 / 40, POLE_NEG.name: np.log(4.0) / 5.0 / 25}


def _ctx(cfg, sem, precision="fp64"):
    p = lib_params(cfg, precision=precision, semantics=sem)
    p.pole_eps = EPS[cfg.name]
    return L.Context(p)


def _oracle(cfg, sem):
    return O.params(cfg.L, cfg.r_cutoff, cfg.alpha, cfg.gamma_1, cfg.gamma_2, cfg.eps_0,
                    cfg.accuracy, cfg.N_t, sem, pole_eps=EPS[cfg.name])


@pytest.mark.parametrize("cfg", [POLE, POLE_NEG])
@pytest.mark.parametrize("sem", [L.SEM_POLE_SPLIT, L.SEM_POLE_SPLIT | L.SEM_CORRECTED])
def test_pole_split_short_forces_fp64(cfg, sem):
    xyz, q = make_system(cfg, seed=3)
    with _ctx(cfg, sem) as c:
        pairs = c.neighbor_list(xyz)
        got = c.short_forces_pairs(xyz, q, pairs)
    ref = O.short_forces(_oracle(cfg, sem), xyz, q, pairs)
    # the central difference N'(k0) (step 0.01 eps) magnifies the table's
    # 5e-15 Bessel error by 1/step; the forces still agree to 1e-11
    assert rel_rms(got, ref) < 1e-11
    assert np.all(np.isfinite(got))


@pytest.mark.parametrize("sem", [L.SEM_POLE_SPLIT, L.SEM_POLE_SPLIT | L.SEM_CORRECTED])
def test_pole_split_pair_and_self_fp64(sem):
    cfg = POLE
    rng = np.random.default_rng(4)
    m = 64
    ri = np.column_stack([rng.uniform(5, 6, m), rng.uniform(5, 6, m), rng.uniform(0.2, 1.8, m)])
    rj = np.column_stack([ri[:, 0] + rng.uniform(-2, 2, m), ri[:, 1] + rng.uniform(-2, 2, m),
                          rng.uniform(0.2, 1.8, m)])
    qi, qj = rng.normal(size=m), rng.normal(size=m)
    p = _oracle(cfg, sem)
    with _ctx(cfg, sem) as c:
        fi, fzj = c.short_pair(qi, qj, ri, rj)
        fs = c.short_self(qi, ri)
    for t in range(m):
        o, oz = O.F_short_ij(p, qi[t], qj[t], ri[t], rj[t])
        np.testing.assert_allclose(fi[t], o, rtol=1e-10, atol=1e-12 * np.abs(o).max())
        assert fzj[t] == pytest.approx(oz, rel=1e-10, abs=1e-12 * abs(oz))
        os_ = O.F_short_i(p, qi[t], ri[t])
        assert fs[t, 2] == pytest.approx(os_[2], rel=1e-11, abs=1e-14)


@pytest.mark.parametrize("cfg", [POLE, POLE_NEG])
def test_pole_split_mixed(cfg):
    sem = L.SEM_POLE_SPLIT | L.SEM_CORRECTED
    xyz, q = make_system(cfg, seed=5)
    with _ctx(cfg, sem, "mixed") as c:
        pairs = c.neighbor_list(xyz)
        got = c.short_forces_pairs(xyz, q, pairs)
    ref = O.short_forces(_oracle(cfg, sem), xyz, q, pairs)
    assert rel_rms(got, ref) < 1e-6


def test_pole_step_forces_and_refusals():
    # one whole step (short + long; the long range has no pole handling in the
    # reference and is evaluated as written), device-resident, then the
    # entry points that need |gamma_1 gamma_2| < 1
    cfg = POLE
    sem = L.SEM_POLE_SPLIT | L.SEM_CORRECTED
    xyz, q = make_system(cfg, seed=6)
    with _ctx(cfg, sem) as c:
        K, _, sK = c.kset()
        kidx = c.sample_k(2, cfg.P)
        f = c.forces(xyz, q, kidx)
        pairs = c.neighbor_list(xyz)
        with pytest.raises(L.QEMError) as e:
            c.energy(xyz, q, kidx)
        assert e.value.status == L.QEM_ERR_UNSUPPORTED
        with pytest.raises(L.QEMError) as e:
            c.hvp(xyz, q, kidx, np.ones_like(xyz))
        assert e.value.status == L.QEM_ERR_UNSUPPORTED
    p = _oracle(cfg, sem)
    ref = O.short_forces(p, xyz, q, pairs) + O.long_forces(p, xyz, q, K[kidx], weight=sK / cfg.P)
    assert rel_rms(f, ref) < 1e-11


def test_as_written_across_the_pole_fp64():
    # without the option the reference's per-image Gauss-Legendre sums run
    # across the pole; FP64 reproduces them (mixed precision refuses)
    cfg = POLE
    xyz, q = make_system(cfg, seed=7)
    with _ctx(cfg, 0) as c:
        pairs = c.neighbor_list(xyz)
        got = c.short_forces_pairs(xyz, q, pairs)
    ref = O.short_forces(_oracle(cfg, 0), xyz, q, pairs)
    assert rel_rms(got, ref) < 1e-10
    p = lib_params(cfg, precision="mixed", semantics=0)
    with pytest.raises(L.QEMError) as e:
        L.Context(p)
    assert e.value.status == L.QEM_ERR_UNSUPPORTED
