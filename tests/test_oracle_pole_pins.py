"""CPU pins for the oracle's pole splitting (gamma_1 gamma_2 > 1, the
principal-value rule of src/python_lib/Force_short_norm_njit.py).

For gamma_1 gamma_2 > 1 the denominator D(k) = g12 e^{-2 k Lz} - 1 of every
short-range integrand vanishes at k0 = ln(g12) / (2 Lz).  The reference's rule
(segments around k0, the pole term N(k0)/(2 Lz (k - k0)) added on [0, 2 k0],
the regular part of 1/D in the middle segment) is checked against the same
principal value computed another way: scipy's QUADPACK QAWC routine
(`quad(..., weight="cauchy")`) applied to the smooth N(k) (k - k0)/D(k), with
N written out here in numpy from the Julia integrands (src/Force_short.jl:4-22,
src/greens_function.jl:17-116).  A wrong sign or factor in the pole term, a
dropped middle correction or a wrong image group moves the result far beyond
the tolerances.  The force assembly (signs of every part, the j-side z force)
is pinned the same way.  The image series itself diverges for g12 > 1, so no
brute-force image sum exists to check against.
"""
import numpy as np
import pytest
from scipy import integrate, special

import oracle as O


def _images(zi, zj, Lz, g1, g2, self_term=False):
    if self_term:
        a = np.array([0.0, 2 * zi, 2 * Lz - 2 * zi, 2 * Lz])
        s = np.array([0.0, -1.0, 1.0, 0.0])
    else:
        zn, zp = abs(zi - zj), zi + zj
        a = np.array([zn, zp, 2 * Lz - zp, 2 * Lz - zn])
        sg = np.sign(zi - zj)
        s = np.array([-sg, -1.0, 1.0, sg])
    b = np.array([1.0, g1, g2, g1 * g2])
    return a, b, s


def _num(kind, k, rho, a, b, s, mask, g12, Lz, alpha):
    """the integrand times D(k), images in mask (bit l-1 = image l)"""
    sel = np.array([(mask >> i) & 1 for i in range(4)], dtype=bool)
    t = b[sel] * np.exp(-k * a[sel])
    gu = 0.5 * t.sum()
    gz = 0.5 * k * (s[sel] * t).sum()
    pt = g12 * np.exp(-2 * k * Lz)
    ga = np.exp(-k * k / (4 * alpha))
    return {"rho_point": k * pt * gu * special.j1(k * rho),
            "rho_gauss": k * ga * gu * special.j1(k * rho),
            "z_point": pt * gz * special.j0(k * rho),
            "z_gauss": ga * gz * special.j0(k * rho)}[kind]


def _pv(kind, rho, a, b, s, mask, g12, Lz, alpha, kf):
    k0 = np.log(g12) / (2 * Lz)

    def h(k):   # N(k) (k - k0) / D(k), smooth through k0
        x = 2 * Lz * (k - k0)
        r = -1.0 / (2 * Lz) if abs(x) < 1e-14 else (x / (2 * Lz)) / np.expm1(-x)
        return _num(kind, k, rho, a, b, s, mask, g12, Lz, alpha) * r

    val, err = integrate.quad(h, 0.0, kf, weight="cauchy", wvar=k0, epsabs=1e-14,
                              epsrel=1e-13, limit=400)
    return val


# g1, g2, Lz, zi, zj, rho, alpha, acc: every segment of the rule gets at
# least 3 nodes at N_t = 200, and the middle half-width is k0 / 100, where the
# rule's own error (O(eps^3) from the middle segment, plus the few nodes of
# the side segments) is below 1e-6
CASES = [
    (1.5, 1.5, 3.0, 0.9, 0.4, 0.7, 1.0, 1e-8),
    (-2.0, -1.2, 2.0, 0.3, 1.6, 1.9, 0.6, 1e-8),
    (3.0, 0.8, 4.0, 2.5, 2.45, 0.2, 0.3, 1e-9),
]
TOL = 2e-6


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("kind", ["rho_point", "rho_gauss", "z_point", "z_gauss"])
@pytest.mark.parametrize("mask", [15, 9, 6])
def test_pole_integral_is_the_principal_value(case, kind, mask):
    g1, g2, Lz, zi, zj, rho, alpha, acc = case
    g12 = g1 * g2
    a, b, s = _images(zi, zj, Lz, g1, g2)
    point = kind.endswith("point")
    kf_base = -np.log(acc) / (2 * Lz) if point else np.sqrt(-4 * alpha * np.log(acc))
    k0 = np.log(g12) / (2 * Lz)
    st, k0_o, kf, n1, n3 = O.pole_rule(g12, Lz, kf_base, k0 / 100, 200)
    assert st == 0 and k0_o == pytest.approx(k0, rel=1e-15)
    assert kf == max(kf_base, 2.1 * k0) and n1 == int(200 * k0 / kf) and n3 == 200 - 2 * n1
    assert n1 >= 3
    e = O.elem_ij(g1, g2, zi, zj, rho, Lz, alpha, acc)
    got = O.pole_integral(kind, e, mask, kf_base, k0 / 100, 200)
    ref = _pv(kind, rho, a, b, s, mask, g12, Lz, alpha, kf)
    assert abs(got - ref) <= TOL * abs(ref)


@pytest.mark.parametrize("kind", ["rho_point", "z_point", "z_gauss"])
def test_pole_rule_error_shrinks_as_eps_cubed(kind):
    # the middle segment replaces (N(k) - N(k0))/x by N'(k0)/(2 Lz): its error
    # is odd in k - k0 to first order and O(eps^3) overall
    g1, g2, Lz, zi, zj, rho, alpha, acc = CASES[0]
    a, b, s = _images(zi, zj, Lz, g1, g2)
    k0 = np.log(g1 * g2) / (2 * Lz)
    kfb = -np.log(acc) / (2 * Lz) if kind.endswith("point") else np.sqrt(-4 * alpha * np.log(acc))
    e = O.elem_ij(g1, g2, zi, zj, rho, Lz, alpha, acc)
    ref = _pv(kind, rho, a, b, s, 15, g1 * g2, Lz, alpha, max(kfb, 2.1 * k0))
    err = [abs(O.pole_integral(kind, e, 15, kfb, k0 / d, 200) - ref) for d in (10, 20, 40)]
    assert 5.0 < err[0] / err[1] < 12.0 and 5.0 < err[1] / err[2] < 12.0


def test_pole_rule_limits():
    # N1 = floor(N_t k0 / kf) < 1 has no rule (numpy's leggauss(0) raises in the
    # reference); the half-width must lie in (0, k0)
    Lz, acc = 10.0, 1e-6
    kf = -np.log(acc) / (2 * Lz)
    assert O.pole_rule(1.01, Lz, kf, 1e-5, 30)[0] != 0
    assert O.pole_rule(4.0, Lz, kf, 1e-3, 30)[0] == 0
    assert O.pole_rule(4.0, Lz, kf, 0.5, 30)[0] != 0
    assert O.pole_rule(0.9, Lz, kf, 1e-3, 30)[0] != 0
    e = O.elem_ij(1.01, 1.0, 1.0, 2.0, 0.5, Lz, 1.0, acc)
    assert np.isnan(O.pole_integral("z_point", e, 15, kf, 1e-5, 30))


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("sem", [O.SEM_POLE_SPLIT, O.SEM_POLE_SPLIT | O.SEM_CORRECTED])
def test_pair_and_self_forces_with_pole_split(case, sem):
    g1, g2, Lz, zi, zj, rho, alpha, acc = case
    g12 = g1 * g2
    N_t = 200
    ri = np.array([5.0, 5.0, zi])
    rj = np.array([5.0 + rho * 0.6, 5.0 - rho * 0.8, zj])
    k0 = np.log(g12) / (2 * Lz)
    p = O.params((10.0, 10.0, Lz), 3.0, alpha, g1, g2, 1.0, acc, N_t, sem, pole_eps=k0 / 100)
    qi, qj = 1.3, -0.7
    f, fzj = O.F_short_ij(p, qi, qj, ri, rj)
    a, b, s = _images(zi, zj, Lz, g1, g2)
    kfp = max(-np.log(acc) / (2 * Lz), 2.1 * np.log(g12) / (2 * Lz))
    kfg = max(np.sqrt(-4 * alpha * np.log(acc)), 2.1 * np.log(g12) / (2 * Lz))
    pv = lambda kind, mask, kf: _pv(kind, rho, a, b, s, mask, g12, Lz, alpha, kf)
    closed_r = 0.5 * np.sum(b * rho / (a * a + rho * rho) ** 1.5)
    closed_z = b * a * s / (a * a + rho * rho) ** 1.5
    fr = -pv("rho_point", 15, kfp) + closed_r + pv("rho_gauss", 15, kfg)
    fz = pv("z_point", 15, kfp) - 0.5 * closed_z.sum() - pv("z_gauss", 15, kfg)
    pref = qi * qj / (2 * np.pi)
    d = (ri - rj)[:2]
    want = pref * np.array([fr * d[0] / rho, fr * d[1] / rho, fz])
    np.testing.assert_allclose(f, want, rtol=TOL, atol=TOL * np.abs(want).max())
    if sem & O.SEM_NEWTON_IMAGES:
        odd = pv("z_point", 9, kfp) - 0.5 * (closed_z[0] + closed_z[3]) - pv("z_gauss", 9, kfg)
        even = pv("z_point", 6, kfp) - 0.5 * (closed_z[1] + closed_z[2]) - pv("z_gauss", 6, kfg)
        assert fzj == pytest.approx(pref * (even - odd), rel=TOL, abs=TOL * abs(pref))
    else:
        assert fzj == -f[2]
    # the self term
    a, b, s = _images(zi, zi, Lz, g1, g2, self_term=True)
    pv = lambda kind, kf: _pv(kind, 0.0, a, b, s, 15, g12, Lz, alpha, kf)
    closed = 0.5 * sum(b[l] * a[l] * s[l] / a[l] ** 3 for l in (1, 2))
    fzs = -(qi * qi) / (2 * np.pi) * (-pv("z_point", kfp) + closed + pv("z_gauss", kfg))
    got = O.F_short_i(p, qi, ri)
    assert got[2] == pytest.approx(fzs, rel=TOL, abs=1e-12)
    assert got[0] == 0.0 and got[1] == 0.0
