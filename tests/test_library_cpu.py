"""Host-side checks of libqem_hip.so and the Python mirror (no GPU needed)."""
import ctypes as C
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from qem4molly_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "qem.h")).read()
    return sorted(set(re.findall(r"^(?:int|void|const char \*)\s*(qem_\w+)\s*\(", txt, re.M)))


def test_library_exports_every_declared_symbol():
    names = _declared()
    assert len(names) >= 30
    lib = L.lib()
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(L.EXPORTS)


def test_library_is_gfx950_hip_code():
    out = subprocess.run(["readelf", "-d", L.LIB_PATH], capture_output=True, text=True).stdout
    assert "libamdhip64" in out
    blob = open(L.LIB_PATH, "rb").read()
    assert b"gfx950" in blob


def _gl_mpmath(n):
    import mpmath as mp
    mp.mp.dps = 40
    xs, ws = [], []
    for x0 in np.polynomial.legendre.leggauss(n)[0]:
        t = mp.findroot(lambda s: mp.legendre(n, s), mp.mpf(x0))
        dp = n * (mp.legendre(n - 1, t) - t * mp.legendre(n, t)) / (1 - t * t)
        xs.append(float(t))
        ws.append(float(2 / ((1 - t * t) * dp * dp)))
    return np.array(xs), np.array(ws)


def test_gauss_legendre_host_rule():
    # 40-digit reference for the orders the workloads use
    for n in (3, 30, 40, 50):
        x, w = L.gauss_legendre(n)
        xm, wm = _gl_mpmath(n)
        assert np.max(np.abs(x - xm)) <= 1.2e-16
        assert np.max(np.abs(w - wm) / wm) <= 2.3e-16
    for n in (1, 128, 256):
        x, w = L.gauss_legendre(n)
        xr, wr = np.polynomial.legendre.leggauss(n)
        assert np.max(np.abs(x - xr)) < 1e-15
        assert np.max(np.abs(w - wr) / wr) < 1e-10   # numpy's own end weights err ~1e-11
    with pytest.raises(L.QEMError):
        L.gauss_legendre(0)
    with pytest.raises(L.QEMError):
        L.gauss_legendre(257)


def test_inter_params_match_reference_factory():
    # src/Quasi_Ewald_Method.jl:5-7 (rbe) and 14-16 (non-rbe)
    a, rc, kc = L.inter_params(1000, 100.0, 100.0, 1.0, True)
    assert a == 1000 / 1e4 and rc == 1.0 / np.sqrt(a) and kc == 2 * np.sqrt(a)
    a, rc, kc = L.inter_params(1000, 100.0, 50.0, 2.0, False)
    assert a == np.sqrt(1000) / 5000 and rc == 2.0 / np.sqrt(a) and kc == 4.0 * np.sqrt(a)


@pytest.mark.parametrize("field,value,status", [
    ("Lz", 0.0, L.QEM_ERR_INVALID_ARG),
    ("r_cutoff", 11.0, L.QEM_ERR_INVALID_ARG),      # >= min(Lx, Ly)/2
    ("alpha", -1.0, L.QEM_ERR_INVALID_ARG),
    ("accuracy", 1.0, L.QEM_ERR_INVALID_ARG),
    ("N_t", 0, L.QEM_ERR_INVALID_ARG),
    ("N_t", 300, L.QEM_ERR_INVALID_ARG),
    ("iter_period", 0, L.QEM_ERR_INVALID_ARG),
    ("precision", 7, L.QEM_ERR_INVALID_ARG),
    ("semantics", 16, L.QEM_ERR_INVALID_ARG),
    ("pole_eps", float("nan"), L.QEM_ERR_INVALID_ARG),
    ("eps_0", 0.0, L.QEM_ERR_INVALID_ARG),
    ("Lx", float("nan"), L.QEM_ERR_INVALID_ARG),
])
def test_create_rejects_bad_parameters(field, value, status):
    p = L.default_params()
    p.Lx = p.Ly = 20.0
    p.Lz = 5.0
    p.r_cutoff = 5.0
    p.gamma_1 = 0.9
    p.gamma_2 = 0.9
    setattr(p, field, value)
    with pytest.raises(L.QEMError) as e:
        L.Context(p)
    assert e.value.status == status


@pytest.mark.parametrize("g1,g2,prec,sem,eps,N_t,alpha", [
    (1.1, 1.1, L.PREC_MIXED, 0, 0.0, 30, 1.0),                  # pole, mixed, no rule
    (1.005, 1.0, L.PREC_FP64, L.SEM_POLE_SPLIT, 0.0, 30, 1.0),  # floor(N_t k0 / kf) = 0
    (2.0, 2.0, L.PREC_FP64, L.SEM_POLE_SPLIT, 0.2, 60, 0.1),    # eps >= k0 = ln 4 / 10
    (2.0, 2.0, L.PREC_MIXED, L.SEM_POLE_SPLIT, 0.0, 4, 0.1),    # too few nodes
])
def test_create_rejects_pole_settings_without_a_rule(g1, g2, prec, sem, eps, N_t, alpha):
    # gamma_1 gamma_2 > 1: D(k) vanishes at k0 = ln(g1 g2) / (2 Lz)
    p = L.default_params()
    p.Lx = p.Ly = 20.0
    p.Lz = 5.0
    p.r_cutoff = 5.0
    p.gamma_1, p.gamma_2 = g1, g2
    p.precision, p.semantics, p.pole_eps, p.N_t, p.alpha = prec, sem, eps, N_t, alpha
    with pytest.raises(L.QEMError) as e:
        L.Context(p)
    assert e.value.status == L.QEM_ERR_UNSUPPORTED


def test_default_params_are_reference_defaults():
    p = L.default_params()
    # src/QEM_short.jl:19 and src/QEM_long.jl:19 keyword defaults
    assert (p.iter_period, p.N_t, p.rbe_p) == (100, 30, 30)
    assert (p.gamma_1, p.gamma_2, p.eps_0, p.accuracy, p.alpha) == (0.0, 0.0, 1.0, 1e-6, 1.0)
    assert p.precision == L.PREC_FP64 and p.semantics == 0 and p.pole_eps == 0.0


def test_null_and_status_strings():
    lib = L.lib()
    assert lib.qem_create(None, None) == L.QEM_ERR_INVALID_ARG
    assert lib.qem_destroy(None) == L.QEM_OK
    assert lib.qem_status_string(L.QEM_ERR_DOMAIN) == b"input outside the domain"
    assert lib.qem_last_error(None) == b"null context"


def test_bench_refuses_multi_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr and r.stdout == ""


def test_bench_flop_convention():
    # the roofline numerator: 4 image terms x N_t nodes x (point 62 + gaussian
    # 64) flop per pair, i.e. 25,200 at N_t = 50 (add/mul 1, FMA 2,
    # transcendental 1 plus its argument, joint J0 + J1 40)
    sys.path.insert(0, ROOT)
    import bench
    assert bench.pair_flops(1, 50) == 25_200.0
    assert bench.pair_flops(15_740_142, 50) == 15_740_142 * 25_200.0
    assert bench.parts_per_step({"phases": {"pair": (2.99, 2.99), "long": (0.32, 0.32)}}) == \
        {"pair": 2.99, "long": 0.32}


def test_workload_configs():
    from qem4molly_amd.workloads import CONFIGS, make_system
    xyz, q = make_system("slab_n100k")
    assert xyz.shape == (100_000, 3) and q.sum() == 0.0
    assert (q == 2).sum() == 33_333 and (q == -1).sum() == 66_666
    assert xyz[:, 2].min() >= 0.5 and xyz[:, 2].max() < 19.5
    xyz, q = make_system("slab_n100")
    assert q[0] == -1.0 and q[1] == 1.0 and q.sum() == 0.0   # (-1)^i, i = 1..N
    for c in CONFIGS.values():
        assert 2 * c.r_cutoff < min(c.Lx, c.Ly)


def test_fp64_exp2_coefficients_match_their_generator():
    # dev::exp2_fp64 (qem_device.h) holds the degree-10 polynomial
    # interpolating 2^f at the Chebyshev points of [-1/2, 1/2], as
    # csrc/gen_exp2_coeffs.py computes it in 50-digit arithmetic; its double
    # Horner evaluation stays within 5e-16 relative of 2^f
    import re
    import mpmath as mp
    csrc = os.path.join(ROOT, "qem4molly_amd", "csrc")
    sys.path.insert(0, csrc)
    from gen_exp2_coeffs import coefficients
    src = open(os.path.join(csrc, "qem_device.h")).read()
    body = re.search(r"kExp2C\[11\] = \{(.*?)\};", src, re.S).group(1)
    held = [float(v) for v in body.replace("\n", " ").split(",")]
    want = coefficients()
    assert held == want
    f = np.linspace(-0.5, 0.5, 2001)
    p = np.zeros_like(f)
    for v in reversed(held):
        p = p * f + v
    mp.mp.dps = 30
    ref = np.array([float(mp.power(2, mp.mpf(x))) for x in f])
    assert np.max(np.abs(p - ref) / ref) < 5e-16
