"""ctypes view of the CPU FP64 oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke()
(the comparison side) and by bench.py's cpu_baseline leg.  The product
package never imports it.  Each wrapper names the reference function it
restates (QEM4Molly, Julia; see qem_oracle.c for file:line).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_DEFAULT_LIB = os.path.join(_HERE, "liboracle.so")
_V4_LIB = os.path.join(_HERE, "liboracle_v4.so")
# QEM_ORACLE_LIB names another build of the same source (the AddressSanitizer
# build of `make -C oracle asan`, for tests/test_oracle_asan.py)
_LIB_PATH = os.environ.get("QEM_ORACLE_LIB") or _DEFAULT_LIB

SEM_SWAP_LIMITS = 1
SEM_MINIMAGE_DIR = 2
SEM_NEWTON_IMAGES = 4
SEM_CORRECTED = SEM_SWAP_LIMITS | SEM_MINIMAGE_DIR | SEM_NEWTON_IMAGES
SEM_POLE_SPLIT = 8


class Params(C.Structure):
    _fields_ = [
        ("Lx", C.c_double), ("Ly", C.c_double), ("Lz", C.c_double),
        ("r_cutoff", C.c_double), ("alpha", C.c_double),
        ("gamma_1", C.c_double), ("gamma_2", C.c_double),
        ("eps_0", C.c_double), ("accuracy", C.c_double),
        ("N_t", C.c_int32), ("semantics", C.c_uint32),
        ("pole_eps", C.c_double),
    ]


class Elem(C.Structure):
    _fields_ = [
        ("gamma_1", C.c_double), ("gamma_2", C.c_double), ("rho", C.c_double),
        ("a", C.c_double * 4), ("b", C.c_double * 4), ("s", C.c_double * 4),
        ("L_z", C.c_double), ("alpha", C.c_double),
        ("k_f1", C.c_double * 4), ("k_f2", C.c_double * 4),
    ]


def build(force: bool = False) -> str:
    """Compile liboracle.so (and liboracle_v4.so) with oracle/Makefile."""
    if force or not (os.path.exists(_DEFAULT_LIB) and os.path.exists(_V4_LIB)):
        subprocess.run(["make", "-s", "-B" if force else "-s", "-C", _HERE],
                       check=True)
    return _LIB_PATH


def host_has_avx512() -> bool:
    """Whether /proc/cpuinfo reports the x86-64-v4 extensions (AVX-512 F, BW,
    CD, DQ, VL) that liboracle_v4.so is built for."""
    want = {"avx512f", "avx512bw", "avx512cd", "avx512dq", "avx512vl"}
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("flags"):
                    return want <= set(line.split(":", 1)[1].split())
    except OSError:
        pass
    return False


def use_avx512_build() -> str:
    """For timing: load liboracle_v4.so when the host has AVX-512 (call before
    the first oracle call).  Same source and flags apart from -march, so the
    results are the same.  Returns the path that will be used."""
    global _LIB_PATH
    if _lib is not None:
        return _LIB_PATH
    if os.path.exists(_V4_LIB) and host_has_avx512() and not os.environ.get("QEM_ORACLE_LIB"):
        _LIB_PATH = _V4_LIB
    return _LIB_PATH


_lib = None
_P = C.POINTER


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if _LIB_PATH in (_DEFAULT_LIB, _V4_LIB):
            build()
        L = C.CDLL(_LIB_PATH)
        d, i32, i64, u32 = C.c_double, C.c_int32, C.c_int64, C.c_uint32
        pd, pf, pi32, pi64 = _P(d), _P(C.c_float), _P(i32), _P(i64)
        sig = {
            "qo_legendre": (C.c_int, [C.c_int, pd, pd]),
            "qo_elem_ij": (None, [d, d, d, d, d, d, d, d, _P(Elem)]),
            "qo_elem_i": (None, [d, d, d, d, d, d, _P(Elem)]),
            "qo_elem_long": (None, [d, d, d, d, _P(Elem)]),
            "qo_Gamma_1": (d, [d, _P(Elem), C.c_int]),
            "qo_Gamma_2": (d, [d, _P(Elem), C.c_int]),
            "qo_dz_Gamma_1": (d, [d, _P(Elem), C.c_int]),
            "qo_dz_Gamma_2": (d, [d, _P(Elem), C.c_int]),
            "qo_F_short_ij": (None, [_P(Params), d, d, pd, pd, C.c_int, pd, pd]),
            "qo_F_short_i": (None, [_P(Params), d, pd, C.c_int, pd]),
            "qo_neighbor_pairs": (i64, [i64, pd, d, d, d, pi32, i64]),
            "qo_short_forces": (None, [_P(Params), i64, pd, pd, i64, pi32, pd, C.c_int]),
            "qo_short_pair_forces": (None, [_P(Params), pd, pd, i64, pi32, pd, C.c_int]),
            "qo_kset_size": (i64, [d, d, d, d, pi32, pi32]),
            "qo_kset": (None, [d, d, d, d, pd, pf, pd]),
            "qo_sortperm_z": (None, [i64, pd, pi64]),
            "qo_Fl_non_sort": (None, [pd, i64, pd, pd, d, d, d, pd]),
            "qo_Fl_sort": (None, [pd, i64, pi64, pd, pd, pd]),
            "qo_Fl_sort_subtractive": (None, [pd, i64, pi64, pd, pd, pd]),
            "qo_Flz_sort_k0": (None, [i64, pi64, pd, pd]),
            "qo_Flz_self": (None, [pd, i64, pd, pd, d, d, d, pd]),
            "qo_F_l_total": (None, [pd, i64, pd, pd, pi64, _P(Params), pd]),
            "qo_long_forces": (None, [_P(Params), i64, pd, pd, pd, i64, d, C.c_int, pd, C.c_int]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _pd(a):
    return a.ctypes.data_as(_P(C.c_double))


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def default_pole_eps(gamma_1, gamma_2, Lz):
    """Middle-segment half-width when none is given: k0 / 20,
    k0 = ln(gamma_1 gamma_2) / (2 Lz) (0 when there is no pole)."""
    g12 = gamma_1 * gamma_2
    return float(np.log(g12) / (2.0 * Lz) / 20.0) if g12 > 1.0 else 0.0


def params(L, r_cutoff, alpha, gamma_1=0.0, gamma_2=0.0, eps_0=1.0,
           accuracy=1e-6, N_t=30, semantics=0, pole_eps=None) -> Params:
    """Mirror of the QEM_short/QEM_long keyword set (src/QEM_short.jl:19)."""
    Lx, Ly, Lz = L
    if pole_eps is None:
        pole_eps = default_pole_eps(gamma_1, gamma_2, Lz)
    return Params(float(Lx), float(Ly), float(Lz), float(r_cutoff),
                  float(alpha), float(gamma_1), float(gamma_2), float(eps_0),
                  float(accuracy), int(N_t), int(semantics), float(pole_eps))


def pole_rule(g12, Lz, kf_base, eps, N_t):
    """(status, k0, kf, N1, N3) of the pole-splitting rule (qo_pole_rule)."""
    L = lib()
    if not hasattr(L.qo_pole_rule, "_sig"):
        d = C.c_double
        L.qo_pole_rule.restype = C.c_int
        L.qo_pole_rule.argtypes = [d, d, d, d, C.c_int, _P(d), _P(d), _P(C.c_int), _P(C.c_int)]
        L.qo_pole_rule._sig = True
    k0, kf, n1, n3 = C.c_double(), C.c_double(), C.c_int(), C.c_int()
    st = L.qo_pole_rule(float(g12), float(Lz), float(kf_base), float(eps), int(N_t),
                        C.byref(k0), C.byref(kf), C.byref(n1), C.byref(n3))
    return st, k0.value, kf.value, n1.value, n3.value


POLE_KINDS = {"rho_gauss": 0, "rho_point": 1, "z_gauss": 2, "z_point": 3}


def pole_integral(kind, e, mask, kf_base, eps, N_t):
    """qo_pole_integral: one short-range integral by the pole-splitting rule."""
    L = lib()
    if not hasattr(L.qo_pole_integral, "_sig"):
        d = C.c_double
        L.qo_pole_integral.restype = d
        L.qo_pole_integral.argtypes = [C.c_int, _P(Elem), C.c_uint, d, d, C.c_int]
        L.qo_pole_integral._sig = True
    return L.qo_pole_integral(POLE_KINDS[kind], C.byref(e), int(mask), float(kf_base),
                              float(eps), int(N_t))


def legendre(n):
    x = np.zeros(n)
    w = np.zeros(n)
    lib().qo_legendre(n, _pd(x), _pd(w))
    return x, w


def elem_ij(g1, g2, zi, zj, rho, Lz, alpha, acc):
    e = Elem()
    lib().qo_elem_ij(g1, g2, zi, zj, rho, Lz, alpha, acc, C.byref(e))
    return e


def elem_i(g1, g2, zi, Lz, alpha, acc):
    e = Elem()
    lib().qo_elem_i(g1, g2, zi, Lz, alpha, acc, C.byref(e))
    return e


def gamma(kind, k, e, l=0):
    fn = {"Gamma_1": lib().qo_Gamma_1, "Gamma_2": lib().qo_Gamma_2,
          "dz_Gamma_1": lib().qo_dz_Gamma_1, "dz_Gamma_2": lib().qo_dz_Gamma_2}[kind]
    return fn(float(k), C.byref(e), int(l))


def F_short_ij(p: Params, qi, qj, ri, rj, single=False):
    ri = _c(ri, np.float64); rj = _c(rj, np.float64)
    out = np.zeros(3)
    fzj = np.zeros(1)
    lib().qo_F_short_ij(C.byref(p), float(qi), float(qj), _pd(ri), _pd(rj),
                        int(bool(single)), _pd(out), _pd(fzj))
    return out, float(fzj[0])


def F_short_i(p: Params, qi, ri, single=False):
    ri = _c(ri, np.float64)
    out = np.zeros(3)
    lib().qo_F_short_i(C.byref(p), float(qi), _pd(ri), int(bool(single)), _pd(out))
    return out


def neighbor_pairs(xyz, Lx, Ly, rc):
    xyz = _c(xyz, np.float64)
    n = xyz.shape[0]
    cnt = lib().qo_neighbor_pairs(n, _pd(xyz), Lx, Ly, rc, None, 0)
    pairs = np.zeros((max(cnt, 1), 2), dtype=np.int32)
    lib().qo_neighbor_pairs(n, _pd(xyz), Lx, Ly, rc,
                            pairs.ctypes.data_as(_P(C.c_int32)), cnt)
    return pairs[:cnt]


def short_forces(p: Params, xyz, q, pairs, nthreads=1):
    xyz = _c(xyz, np.float64); q = _c(q, np.float64)
    pairs = _c(pairs, np.int32).reshape(-1, 2)
    f = np.zeros_like(xyz)
    lib().qo_short_forces(C.byref(p), xyz.shape[0], _pd(xyz), _pd(q),
                          pairs.shape[0], pairs.ctypes.data_as(_P(C.c_int32)),
                          _pd(f), int(nthreads))
    return f


def short_forces_rows(p: Params, xyz, q, rows, nthreads=1):
    """Molly.forces(::QEM_short) for the given rows only (qo_short_forces_rows):
    each row's pairs by an O(N) scan, summed in qo_short_forces's order."""
    L = lib()
    if not hasattr(L.qo_short_forces_rows, "_sig"):
        d, i64 = C.c_double, C.c_int64
        L.qo_short_forces_rows.restype = None
        L.qo_short_forces_rows.argtypes = [_P(Params), i64, _P(d), _P(d), i64, _P(i64), _P(d),
                                           C.c_int]
        L.qo_short_forces_rows._sig = True
    xyz = _c(xyz, np.float64); q = _c(q, np.float64)
    rows = _c(rows, np.int64).ravel()
    f = np.zeros((rows.shape[0], 3))
    L.qo_short_forces_rows(C.byref(p), xyz.shape[0], _pd(xyz), _pd(q), rows.shape[0],
                           rows.ctypes.data_as(_P(C.c_int64)), _pd(f), int(nthreads))
    return f


def short_pair_forces(p: Params, xyz, q, pairs, nthreads=1):
    xyz = _c(xyz, np.float64); q = _c(q, np.float64)
    pairs = _c(pairs, np.int32).reshape(-1, 2)
    fp = np.zeros((pairs.shape[0], 4))
    lib().qo_short_pair_forces(C.byref(p), _pd(xyz), _pd(q), pairs.shape[0],
                               pairs.ctypes.data_as(_P(C.c_int32)), _pd(fp),
                               int(nthreads))
    return fp


def kset(Lx, Ly, alpha, acc):
    mx = C.c_int32(); my = C.c_int32()
    n = lib().qo_kset_size(Lx, Ly, alpha, acc, C.byref(mx), C.byref(my))
    K = np.zeros((n, 3))
    P = np.zeros(n, dtype=np.float32)
    sK = C.c_double()
    lib().qo_kset(Lx, Ly, alpha, acc, _pd(K), P.ctypes.data_as(_P(C.c_float)),
                  C.byref(sK))
    return K, P, sK.value, (mx.value, my.value)


def sortperm_z(xyz):
    xyz = _c(xyz, np.float64)
    z = np.zeros(xyz.shape[0], dtype=np.int64)
    lib().qo_sortperm_z(xyz.shape[0], _pd(xyz), z.ctypes.data_as(_P(C.c_int64)))
    return z


def _k3(k):
    return _c(k, np.float64)


def Fl_non_sort(k, xyz, q, g1, g2, Lz):
    xyz = _c(xyz, np.float64); q = _c(q, np.float64)
    f = np.zeros_like(xyz); k = _k3(k)
    lib().qo_Fl_non_sort(_pd(k), xyz.shape[0], _pd(xyz), _pd(q), g1, g2, Lz, _pd(f))
    return f


def Fl_sort(k, xyz, q, zlist, subtractive=False):
    xyz = _c(xyz, np.float64); q = _c(q, np.float64)
    zl = _c(zlist, np.int64); f = np.zeros_like(xyz); k = _k3(k)
    fn = lib().qo_Fl_sort_subtractive if subtractive else lib().qo_Fl_sort
    fn(_pd(k), xyz.shape[0], zl.ctypes.data_as(_P(C.c_int64)), _pd(xyz), _pd(q), _pd(f))
    return f


def Flz_sort_k0(q, zlist):
    q = _c(q, np.float64); zl = _c(zlist, np.int64)
    f = np.zeros((q.shape[0], 3))
    lib().qo_Flz_sort_k0(q.shape[0], zl.ctypes.data_as(_P(C.c_int64)), _pd(q), _pd(f))
    return f


def Flz_self(k, xyz, q, g1, g2, Lz):
    xyz = _c(xyz, np.float64); q = _c(q, np.float64)
    f = np.zeros_like(xyz); k = _k3(k)
    lib().qo_Flz_self(_pd(k), xyz.shape[0], _pd(xyz), _pd(q), g1, g2, Lz, _pd(f))
    return f


def F_l_total(k, xyz, q, zlist, p: Params):
    xyz = _c(xyz, np.float64); q = _c(q, np.float64)
    zl = _c(zlist, np.int64); f = np.zeros_like(xyz); k = _k3(k)
    lib().qo_F_l_total(_pd(k), xyz.shape[0], _pd(xyz), _pd(q),
                       zl.ctypes.data_as(_P(C.c_int64)), C.byref(p), _pd(f))
    return f


def long_forces(p: Params, xyz, q, kvecs, weight=1.0, full_sum=False, nthreads=1):
    xyz = _c(xyz, np.float64); q = _c(q, np.float64)
    kv = _c(kvecs, np.float64).reshape(-1, 3)
    f = np.zeros_like(xyz)
    lib().qo_long_forces(C.byref(p), xyz.shape[0], _pd(xyz), _pd(q), _pd(kv),
                         kv.shape[0], float(weight), int(bool(full_sum)), _pd(f),
                         int(nthreads))
    return f


def long_variance(p: Params, xyz, q, kvecs, weight, nthreads=1):
    """Per-atom batch variance of the random-batch long-range force
    (qo_long_variance); kvecs is the batch (P x 3, P >= 2)."""
    L = lib()
    if not hasattr(L.qo_long_variance, "_sig"):
        d, i64 = C.c_double, C.c_int64
        L.qo_long_variance.restype = None
        L.qo_long_variance.argtypes = [_P(Params), i64, _P(d), _P(d), _P(d), i64, d, _P(d),
                                       C.c_int]
        L.qo_long_variance._sig = True
    xyz = _c(xyz, np.float64); q = _c(q, np.float64)
    kv = _c(kvecs, np.float64).reshape(-1, 3)
    var = np.zeros(xyz.shape[0])
    L.qo_long_variance(C.byref(p), xyz.shape[0], _pd(xyz), _pd(q), _pd(kv), kv.shape[0],
                       float(weight), _pd(var), int(nthreads))
    return var


def energy(p: Params, xyz, q, pairs, kvecs=None, weight=1.0, full_sum=False,
           short=True, long=True, nthreads=1):
    """E, phi = dE/dq, (dE/dgamma_1, dE/dgamma_2) -- see qo_energy."""
    L = lib()
    if not hasattr(L.qo_energy, "_sig"):
        d, i64 = C.c_double, C.c_int64
        pd = _P(d)
        L.qo_energy.restype = None
        L.qo_energy.argtypes = [_P(Params), i64, pd, pd, i64, _P(C.c_int32), pd, i64, d,
                                C.c_int, C.c_int, C.c_int, pd, pd, pd, C.c_int]
        L.qo_energy._sig = True
    xyz = _c(xyz, np.float64); q = _c(q, np.float64)
    pairs = _c(pairs, np.int32).reshape(-1, 2)
    kv = _c(np.zeros((0, 3)) if kvecs is None else kvecs, np.float64).reshape(-1, 3)
    E = C.c_double()
    phi = np.zeros(xyz.shape[0])
    dg = np.zeros(2)
    L.qo_energy(C.byref(p), xyz.shape[0], _pd(xyz), _pd(q), pairs.shape[0],
                pairs.ctypes.data_as(_P(C.c_int32)), _pd(kv), kv.shape[0], float(weight),
                int(bool(full_sum)), int(bool(short)), int(bool(long)), C.byref(E), _pd(phi),
                _pd(dg), int(nthreads))
    return E.value, phi, dg


def substrate_lj(xyz, cutoff=1.0, sigma=0.5, eps=1.0, sub_pos=0.0, direction=1, sem=0):
    """(F, status): qo_substrate_lj with accumulate = 0; status 1 = the
    reference's @assert failed (F is zeros then)."""
    L = lib()
    if not hasattr(L.qo_substrate_lj, "_sig"):
        d = C.c_double
        L.qo_substrate_lj.restype = C.c_int
        L.qo_substrate_lj.argtypes = [d, d, d, d, C.c_int, C.c_uint, C.c_int64, _P(d), _P(d),
                                      C.c_int]
        L.qo_substrate_lj._sig = True
    xyz = _c(xyz, np.float64).reshape(-1, 3)
    f = np.zeros_like(xyz)
    st = L.qo_substrate_lj(float(cutoff), float(sigma), float(eps), float(sub_pos),
                           int(direction), int(sem), xyz.shape[0], _pd(xyz), _pd(f), 0)
    return f, st
