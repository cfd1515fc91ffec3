"""ctypes binding of libqem_hip.so (the C ABI declared in include/qem.h).

The library is loaded from this package directory.  There is no fallback:
if it is missing or fails to load, every entry point raises, so nothing can
silently run on the CPU.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libqem_hip.so")

QEM_OK = 0
QEM_ERR_INVALID_ARG = 1
QEM_ERR_DOMAIN = 2
QEM_ERR_HIP = 3
QEM_ERR_NOMEM = 4
QEM_ERR_UNSUPPORTED = 5
QEM_ERR_CAPACITY = 6

PREC_FP64 = 0
PREC_MIXED = 1

SEM_SWAP_LIMITS = 1
SEM_MINIMAGE_DIR = 2
SEM_NEWTON_IMAGES = 4
SEM_CORRECTED = 7
SEM_POLE_SPLIT = 8
# stream argument of the *_dev calls selecting HIP's null (default) stream;
# None selects the context's own stream
STREAM_NULL = 1

LONG_NONSORT = 1
LONG_SORT = 2
LONG_SELF = 4
LONG_K0 = 8
LONG_ALL = 15

# every symbol include/qem.h declares
EXPORTS = (
    "qem_default_params", "qem_create", "qem_destroy", "qem_last_error",
    "qem_status_string", "qem_get_params", "qem_inter_params", "qem_kset_size",
    "qem_kset_copy", "qem_sample_k", "qem_neighbor_list", "qem_short_forces",
    "qem_short_forces_pairs", "qem_short_pair", "qem_short_self",
    "qem_sortperm_z", "qem_long_forces", "qem_long_terms", "qem_forces",
    "qem_short_forces_dev", "qem_long_forces_dev", "qem_forces_dev",
    "qem_short_state", "qem_gauss_legendre", "qem_bessel01_table",
    "qem_dev_malloc", "qem_dev_free", "qem_memcpy_htod", "qem_memcpy_dtoh",
    "qem_device_synchronize", "qem_dev_copy_kernel", "qem_timing_enable",
    "qem_timing_read", "qem_energy", "qem_energy_dev", "qem_hvp", "qem_hvp_dev",
    "qem_device_bytes", "qem_substrate_lj", "qem_substrate_lj_dev", "qem_general_forces_dev",
    "qem_long_variance", "qem_long_variance_dev", "qem_long_terms_dev",
    "qem_long_forces_noalloc_dev", "qem_forces_energy_dev", "qem_forces_energy",
    "qem_forces_energy_save_dev", "qem_forces_energy_save", "qem_hvp_saved_dev", "qem_hvp_saved",
)

NOALLOC_MAX_N = 131072
NOALLOC_MAX_P = 512

PHASES = ("neighbor", "pair", "gather", "long")


class QEMError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"[{status}] {message}")
        self.status = status


class Params(C.Structure):
    """struct qem_params (include/qem.h)."""
    _fields_ = [
        ("Lx", C.c_double), ("Ly", C.c_double), ("Lz", C.c_double),
        ("r_cutoff", C.c_double), ("alpha", C.c_double),
        ("gamma_1", C.c_double), ("gamma_2", C.c_double),
        ("eps_0", C.c_double), ("accuracy", C.c_double),
        ("N_t", C.c_int32), ("iter_period", C.c_int32), ("rbe_p", C.c_int32),
        ("precision", C.c_int32), ("semantics", C.c_uint32), ("device", C.c_int32),
        ("pole_eps", C.c_double),
    ]


class Wall(C.Structure):
    """struct qem_wall (include/qem.h): one Substrate_LJ wall."""
    _fields_ = [
        ("cutoff", C.c_double), ("sigma", C.c_double), ("eps", C.c_double),
        ("sub_pos", C.c_double), ("direction", C.c_int32), ("semantics", C.c_uint32),
    ]


WALL_SEM_ABS_CUTOFF = 1
WALL_SEM_TRUE_FORCE = 2


def wall(cutoff=1.0, sigma=0.5, eps=1.0, sub_pos=0.0, direction=1, semantics=0) -> Wall:
    """A Wall with the reference constructor's defaults (src/Substrate_LJ.jl:15-17)."""
    return Wall(float(cutoff), float(sigma), float(eps), float(sub_pos), int(direction),
                int(semantics))


_P = C.POINTER
_d, _f = C.c_double, C.c_float
_i32, _i64, _u32, _u64 = C.c_int32, C.c_int64, C.c_uint32, C.c_uint64
_pd, _pf, _pi32, _pi64 = _P(_d), _P(_f), _P(_i32), _P(_i64)
_vp = C.c_void_p

_SIGS = {
    "qem_default_params": (None, [_P(Params)]),
    "qem_create": (C.c_int, [_P(Params), _P(_vp)]),
    "qem_destroy": (C.c_int, [_vp]),
    "qem_last_error": (C.c_char_p, [_vp]),
    "qem_status_string": (C.c_char_p, [C.c_int]),
    "qem_get_params": (C.c_int, [_vp, _P(Params)]),
    "qem_inter_params": (C.c_int, [_i64, _d, _d, _d, _i32, _pd, _pd, _pd]),
    "qem_kset_size": (C.c_int, [_vp, _pi64]),
    "qem_kset_copy": (C.c_int, [_vp, _pd, _pf, _pd]),
    "qem_sample_k": (C.c_int, [_vp, _u64, _i64, _pi64]),
    "qem_neighbor_list": (C.c_int, [_vp, _i64, _pd, _i64, _pi32, _pi64]),
    "qem_short_forces": (C.c_int, [_vp, _i64, _pd, _pd, _pd]),
    "qem_short_forces_pairs": (C.c_int, [_vp, _i64, _pd, _pd, _i64, _pi32, _pd]),
    "qem_short_pair": (C.c_int, [_vp, _i64, _pd, _pd, _pd, _pd, _i32, _pd, _pd]),
    "qem_short_self": (C.c_int, [_vp, _i64, _pd, _pd, _i32, _pd]),
    "qem_sortperm_z": (C.c_int, [_vp, _i64, _pd, _pi64]),
    "qem_long_forces": (C.c_int, [_vp, _i64, _pd, _pd, _pi64, _i64, _pd]),
    "qem_long_terms": (C.c_int, [_vp, _i64, _pd, _pd, _pd, _i64, _d, _u32, _pd]),
    "qem_forces": (C.c_int, [_vp, _i64, _pd, _pd, _pi64, _i64, _pd]),
    "qem_long_terms_dev": (C.c_int, [_vp, _i64, _vp, _vp, _pd, _i64, _d, _u32, _vp, _i32, _vp]),
    "qem_long_variance": (C.c_int, [_vp, _i64, _pd, _pd, _pi64, _i64, _pd, _pd]),
    "qem_long_variance_dev": (C.c_int, [_vp, _i64, _vp, _vp, _pi64, _i64, _vp, _vp, _vp]),
    "qem_short_forces_dev": (C.c_int, [_vp, _i64, _vp, _vp, _vp, _vp]),
    "qem_long_forces_dev": (C.c_int, [_vp, _i64, _vp, _vp, _pi64, _i64, _vp, _vp]),
    "qem_forces_dev": (C.c_int, [_vp, _i64, _vp, _vp, _pi64, _i64, _vp, _vp]),
    "qem_short_state": (C.c_int, [_vp, _pi64, _pi64, _pi64]),
    "qem_gauss_legendre": (C.c_int, [_i32, _pd, _pd]),
    "qem_bessel01_table": (C.c_int, [_vp, _i64, _pd, _pd, _pd]),
    "qem_dev_malloc": (C.c_int, [_P(_vp), C.c_size_t]),
    "qem_dev_free": (C.c_int, [_vp]),
    "qem_memcpy_htod": (C.c_int, [_vp, _vp, C.c_size_t]),
    "qem_memcpy_dtoh": (C.c_int, [_vp, _vp, C.c_size_t]),
    "qem_device_synchronize": (C.c_int, []),
    "qem_dev_copy_kernel": (C.c_int, [_vp, _vp, C.c_size_t, _vp]),
    "qem_timing_enable": (C.c_int, [_vp, _i32]),
    "qem_timing_read": (C.c_int, [_vp, _pd, _pi64, _i32]),
    "qem_energy": (C.c_int, [_vp, _i64, _pd, _pd, _pi64, _i64, _pd, _pd, _pd]),
    "qem_energy_dev": (C.c_int, [_vp, _i64, _vp, _vp, _pi64, _i64, _vp, _pd, _pd, _vp]),
    "qem_hvp": (C.c_int, [_vp, _i64, _pd, _pd, _pi64, _i64, _pd, _pd]),
    "qem_hvp_dev": (C.c_int, [_vp, _i64, _vp, _vp, _pi64, _i64, _vp, _vp, _vp]),
    "qem_device_bytes": (C.c_int, [_vp, _pi64]),
    "qem_substrate_lj": (C.c_int, [_vp, _P(Wall), _i64, _pd, _pd, _i32]),
    "qem_substrate_lj_dev": (C.c_int, [_vp, _P(Wall), _i64, _vp, _vp, _i32, _vp]),
    "qem_general_forces_dev": (C.c_int, [_vp, _i64, _vp, _vp, _pi64, _i64, _P(Wall), _i32,
                                         _vp, _vp]),
    "qem_long_forces_noalloc_dev": (C.c_int, [_vp, _i64, _vp, _vp, _pi64, _i64, _vp, _vp]),
    "qem_forces_energy_dev": (C.c_int, [_vp, _i64, _vp, _vp, _pi64, _i64, _vp, _vp, _pd, _pd,
                                        _vp]),
    "qem_forces_energy": (C.c_int, [_vp, _i64, _pd, _pd, _pi64, _i64, _pd, _pd, _pd, _pd]),
    "qem_forces_energy_save_dev": (C.c_int, [_vp, _i64, _vp, _vp, _pi64, _i64, _vp, _vp, _pd,
                                             _pd, _vp]),
    "qem_forces_energy_save": (C.c_int, [_vp, _i64, _pd, _pd, _pi64, _i64, _pd, _pd, _pd, _pd]),
    "qem_hvp_saved": (C.c_int, [_vp, _i64, _pd, _pd, _pi64, _i64, _pd, _pd]),
    "qem_hvp_saved_dev": (C.c_int, [_vp, _i64, _vp, _vp, _pi64, _i64, _vp, _vp, _vp]),
}

_lib = None
_lock = threading.Lock()
# torch ships its own HIP runtime (DT_NEEDED "libamdhip64.so"); this library
# needs "libamdhip64.so.7".  Loaded in that order the process shares torch's
# runtime; the other order would map two runtimes, so torch users must import
# torch first (qem4molly_amd.torch_qem checks this).
_loaded_with_torch = False


def lib() -> C.CDLL:
    """Load libqem_hip.so (raises if it is not built)."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise RuntimeError(
                    f"{LIB_PATH} is not built; run `python -m qem4molly_amd.build` "
                    "(or __graft_entry__.build())")
            global _loaded_with_torch
            import sys
            _loaded_with_torch = "torch" in sys.modules
            L = C.CDLL(LIB_PATH)
            for name, (res, args) in _SIGS.items():
                fn = getattr(L, name)
                fn.restype = res
                fn.argtypes = args
            _lib = L
    return _lib


def f64(a, shape=None) -> np.ndarray:
    out = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None:
        out = out.reshape(shape)
    return out


def ptr(a: np.ndarray):
    if a.dtype == np.float64:
        return a.ctypes.data_as(_pd)
    if a.dtype == np.float32:
        return a.ctypes.data_as(_pf)
    if a.dtype == np.int32:
        return a.ctypes.data_as(_pi32)
    if a.dtype == np.int64:
        return a.ctypes.data_as(_pi64)
    raise TypeError(a.dtype)


def default_params() -> Params:
    p = Params()
    lib().qem_default_params(C.byref(p))
    return p


class Context:
    """Owns one qem_ctx (include/qem.h); thin, checked wrappers per entry point."""

    def __init__(self, params: Params):
        h = _vp()
        st = lib().qem_create(C.byref(params), C.byref(h))
        if st != QEM_OK:
            msg = lib().qem_status_string(st).decode()
            raise QEMError(st, f"qem_create: {msg}")
        self._h = h
        self.params = params
        # counts the forces + energy calls: each one replaces (save) or drops
        # the kept tangent sums, so a caller that recorded the count after its
        # saving call knows the sums are still its own while the count holds
        self.sums_gen = 0

    def close(self):
        if getattr(self, "_h", None):
            lib().qem_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self):
        return self._h

    def _check(self, st: int, what: str):
        if st != QEM_OK:
            err = lib().qem_last_error(self._h)
            raise QEMError(st, f"{what}: {err.decode() if err else ''}")

    # ---- K set -----------------------------------------------------------
    def kset(self):
        n = _i64()
        self._check(lib().qem_kset_size(self._h, C.byref(n)), "qem_kset_size")
        K = np.zeros((n.value, 3))
        P = np.zeros(n.value, dtype=np.float32)
        sK = _d()
        self._check(lib().qem_kset_copy(self._h, ptr(K), ptr(P), C.byref(sK)), "qem_kset_copy")
        return K, P, sK.value

    def sample_k(self, seed: int, P: int) -> np.ndarray:
        out = np.zeros(int(P), dtype=np.int64)
        self._check(lib().qem_sample_k(self._h, _u64(seed & 0xFFFFFFFFFFFFFFFF), int(P), ptr(out)),
                    "qem_sample_k")
        return out

    # ---- short range ------------------------------------------------------
    def neighbor_list(self, xyz) -> np.ndarray:
        xyz = f64(xyz, (-1, 3))
        n = xyz.shape[0]
        cnt = _i64()
        # first call learns the count
        st = lib().qem_neighbor_list(self._h, n, ptr(xyz), 0, None, C.byref(cnt))
        if st not in (QEM_OK, QEM_ERR_CAPACITY):
            self._check(st, "qem_neighbor_list")
        pairs = np.zeros((max(cnt.value, 1), 2), dtype=np.int32)
        self._check(lib().qem_neighbor_list(self._h, n, ptr(xyz), cnt.value, ptr(pairs),
                                            C.byref(cnt)), "qem_neighbor_list")
        return pairs[:cnt.value]

    def short_forces(self, xyz, q) -> np.ndarray:
        xyz = f64(xyz, (-1, 3)); q = f64(q)
        f = np.zeros_like(xyz)
        self._check(lib().qem_short_forces(self._h, xyz.shape[0], ptr(xyz), ptr(q), ptr(f)),
                    "qem_short_forces")
        return f

    def short_forces_pairs(self, xyz, q, pairs) -> np.ndarray:
        xyz = f64(xyz, (-1, 3)); q = f64(q)
        pairs = np.ascontiguousarray(pairs, dtype=np.int32).reshape(-1, 2)
        f = np.zeros_like(xyz)
        self._check(lib().qem_short_forces_pairs(self._h, xyz.shape[0], ptr(xyz), ptr(q),
                                                 pairs.shape[0], ptr(pairs), ptr(f)),
                    "qem_short_forces_pairs")
        return f

    def short_pair(self, qi, qj, ri, rj, single=False):
        ri = f64(ri, (-1, 3)); rj = f64(rj, (-1, 3))
        m = ri.shape[0]
        qi = f64(np.broadcast_to(qi, (m,))); qj = f64(np.broadcast_to(qj, (m,)))
        fi = np.zeros((m, 3)); fzj = np.zeros(m)
        self._check(lib().qem_short_pair(self._h, m, ptr(qi), ptr(qj), ptr(ri), ptr(rj),
                                         int(bool(single)), ptr(fi), ptr(fzj)), "qem_short_pair")
        return fi, fzj

    def short_self(self, q, xyz, single=False) -> np.ndarray:
        xyz = f64(xyz, (-1, 3)); q = f64(q)
        f = np.zeros_like(xyz)
        self._check(lib().qem_short_self(self._h, xyz.shape[0], ptr(q), ptr(xyz),
                                         int(bool(single)), ptr(f)), "qem_short_self")
        return f

    def device_bytes(self) -> int:
        b = _i64()
        self._check(lib().qem_device_bytes(self._h, C.byref(b)), "qem_device_bytes")
        return b.value

    def short_state(self):
        a, b, c = _i64(), _i64(), _i64()
        self._check(lib().qem_short_state(self._h, C.byref(a), C.byref(b), C.byref(c)),
                    "qem_short_state")
        return a.value, b.value, c.value

    # ---- long range -------------------------------------------------------
    def sortperm_z(self, xyz) -> np.ndarray:
        xyz = f64(xyz, (-1, 3))
        z = np.zeros(xyz.shape[0], dtype=np.int64)
        self._check(lib().qem_sortperm_z(self._h, xyz.shape[0], ptr(xyz), ptr(z)),
                    "qem_sortperm_z")
        return z

    def long_forces(self, xyz, q, k_idx) -> np.ndarray:
        xyz = f64(xyz, (-1, 3)); q = f64(q)
        k_idx = np.ascontiguousarray(k_idx, dtype=np.int64)
        f = np.zeros_like(xyz)
        self._check(lib().qem_long_forces(self._h, xyz.shape[0], ptr(xyz), ptr(q), ptr(k_idx),
                                          k_idx.shape[0], ptr(f)), "qem_long_forces")
        return f

    def long_terms(self, xyz, q, kvecs, weight=1.0, parts=LONG_ALL) -> np.ndarray:
        xyz = f64(xyz, (-1, 3)); q = f64(q)
        kv = f64(kvecs, (-1, 3))
        f = np.zeros_like(xyz)
        self._check(lib().qem_long_terms(self._h, xyz.shape[0], ptr(xyz), ptr(q), ptr(kv),
                                         kv.shape[0], float(weight), int(parts), ptr(f)),
                    "qem_long_terms")
        return f

    def long_terms_dev(self, n, d_xyz, d_q, kvecs, weight, parts, d_f, accumulate=False,
                       stream=None):
        kv = f64(kvecs, (-1, 3))
        self._check(lib().qem_long_terms_dev(self._h, int(n), d_xyz, d_q, ptr(kv), kv.shape[0],
                                             float(weight), int(parts), d_f,
                                             int(bool(accumulate)), stream),
                    "qem_long_terms_dev")

    def long_variance(self, xyz, q, k_idx):
        """(F, var): the batch's long-range force (N x 3) and the per-atom
        batch estimate of its variance (N)."""
        xyz = f64(xyz, (-1, 3)); q = f64(q)
        k_idx = np.ascontiguousarray(k_idx, dtype=np.int64)
        f = np.zeros_like(xyz)
        var = np.zeros(xyz.shape[0])
        self._check(lib().qem_long_variance(self._h, xyz.shape[0], ptr(xyz), ptr(q), ptr(k_idx),
                                            k_idx.shape[0], ptr(f), ptr(var)),
                    "qem_long_variance")
        return f, var

    def long_variance_dev(self, n, d_xyz, d_q, k_idx, d_var, d_f=None, stream=None):
        k_idx = np.ascontiguousarray(k_idx, dtype=np.int64)
        self._check(lib().qem_long_variance_dev(self._h, int(n), d_xyz, d_q, ptr(k_idx),
                                                k_idx.shape[0], d_f, d_var, stream),
                    "qem_long_variance_dev")

    def forces(self, xyz, q, k_idx) -> np.ndarray:
        xyz = f64(xyz, (-1, 3)); q = f64(q)
        k_idx = np.ascontiguousarray(k_idx, dtype=np.int64)
        f = np.zeros_like(xyz)
        self._check(lib().qem_forces(self._h, xyz.shape[0], ptr(xyz), ptr(q), ptr(k_idx),
                                     k_idx.shape[0], ptr(f)), "qem_forces")
        return f

    def substrate_lj(self, xyz, w: "Wall", f=None) -> np.ndarray:
        """One wall's forces (N x 3); added to f when f is given."""
        xyz = f64(xyz, (-1, 3))
        acc = f is not None
        out = np.ascontiguousarray(f, dtype=np.float64).copy() if acc else np.zeros_like(xyz)
        self._check(lib().qem_substrate_lj(self._h, C.byref(w), xyz.shape[0], ptr(xyz),
                                           ptr(out), int(acc)), "qem_substrate_lj")
        return out

    # ---- device-pointer variants -----------------------------------------
    def substrate_lj_dev(self, n, w: "Wall", d_xyz, d_f, accumulate=False, stream=None):
        self._check(lib().qem_substrate_lj_dev(self._h, C.byref(w), int(n), d_xyz, d_f,
                                               int(bool(accumulate)), stream),
                    "qem_substrate_lj_dev")

    def general_forces_dev(self, n, d_xyz, d_q, k_idx, walls, d_f, stream=None):
        k_idx = np.ascontiguousarray(k_idx, dtype=np.int64)
        arr = (Wall * max(1, len(walls)))(*walls)
        self._check(lib().qem_general_forces_dev(self._h, int(n), d_xyz, d_q, ptr(k_idx),
                                                 k_idx.shape[0], arr, len(walls), d_f, stream),
                    "qem_general_forces_dev")

    def forces_dev(self, n, d_xyz, d_q, k_idx, d_f, stream=None):
        k_idx = np.ascontiguousarray(k_idx, dtype=np.int64)
        self._check(lib().qem_forces_dev(self._h, int(n), d_xyz, d_q, ptr(k_idx),
                                         k_idx.shape[0], d_f, stream), "qem_forces_dev")

    def short_forces_dev(self, n, d_xyz, d_q, d_f, stream=None):
        self._check(lib().qem_short_forces_dev(self._h, int(n), d_xyz, d_q, d_f, stream),
                    "qem_short_forces_dev")

    def long_forces_dev(self, n, d_xyz, d_q, k_idx, d_f, stream=None):
        k_idx = np.ascontiguousarray(k_idx, dtype=np.int64)
        self._check(lib().qem_long_forces_dev(self._h, int(n), d_xyz, d_q, ptr(k_idx),
                                              k_idx.shape[0], d_f, stream),
                    "qem_long_forces_dev")

    def long_forces_noalloc_dev(self, n, d_xyz, d_q, k_idx, d_f, stream=None):
        """qem_long_forces_noalloc_dev: qem_long_forces_dev with no device
        allocation (work buffers in the library's fixed per-GPU array)."""
        k_idx = np.ascontiguousarray(k_idx, dtype=np.int64)
        self._check(lib().qem_long_forces_noalloc_dev(self._h, int(n), d_xyz, d_q, ptr(k_idx),
                                                      k_idx.shape[0], d_f, stream),
                    "qem_long_forces_noalloc_dev")

    # ---- adjoint ------------------------------------------------------------
    def energy(self, xyz, q, k_idx):
        """(E, dE/dq [N], (dE/dgamma_1, dE/dgamma_2))."""
        xyz = f64(xyz, (-1, 3)); q = f64(q)
        k_idx = np.ascontiguousarray(k_idx, dtype=np.int64)
        E = _d(); phi = np.zeros(xyz.shape[0]); dg = np.zeros(2)
        self._check(lib().qem_energy(self._h, xyz.shape[0], ptr(xyz), ptr(q), ptr(k_idx),
                                     k_idx.shape[0], C.byref(E), ptr(phi), ptr(dg)), "qem_energy")
        return E.value, phi, dg

    def forces_energy(self, xyz, q, k_idx, save=False):
        """(F [N x 3], E, dE/dq [N], (dE/dgamma_1, dE/dgamma_2)) in one call; save
        keeps the pair quadrature's tangent sums for hvp(..., saved=True)."""
        xyz = f64(xyz, (-1, 3)); q = f64(q)
        k_idx = np.ascontiguousarray(k_idx, dtype=np.int64)
        f = np.zeros_like(xyz); E = _d(); phi = np.zeros(xyz.shape[0]); dg = np.zeros(2)
        name = "qem_forces_energy_save" if save else "qem_forces_energy"
        self.sums_gen += 1
        self._check(getattr(lib(), name)(self._h, xyz.shape[0], ptr(xyz), ptr(q), ptr(k_idx),
                                         k_idx.shape[0], ptr(f), C.byref(E), ptr(phi),
                                         ptr(dg)), name)
        return f, E.value, phi, dg

    def forces_energy_dev(self, n, d_xyz, d_q, k_idx, d_f, d_phi=None, stream=None,
                          save=False):
        """(E, dE/dgamma); forces into d_f, dE/dq into d_phi (device)."""
        k_idx = np.ascontiguousarray(k_idx, dtype=np.int64)
        E = _d(); dg = np.zeros(2)
        name = "qem_forces_energy_save_dev" if save else "qem_forces_energy_dev"
        self.sums_gen += 1
        self._check(getattr(lib(), name)(self._h, int(n), d_xyz, d_q, ptr(k_idx),
                                         k_idx.shape[0], d_f, d_phi, C.byref(E), ptr(dg),
                                         stream), name)
        return E.value, dg

    def hvp(self, xyz, q, k_idx, v, saved=False) -> np.ndarray:
        """-(dF/dx) v of the force as computed (N x 3); saved uses the sums of the
        last forces_energy(..., save=True) at the same xyz, q, k_idx."""
        xyz = f64(xyz, (-1, 3)); q = f64(q); v = f64(v, (-1, 3))
        k_idx = np.ascontiguousarray(k_idx, dtype=np.int64)
        hv = np.zeros_like(xyz)
        name = "qem_hvp_saved" if saved else "qem_hvp"
        self._check(getattr(lib(), name)(self._h, xyz.shape[0], ptr(xyz), ptr(q), ptr(k_idx),
                                         k_idx.shape[0], ptr(v), ptr(hv)), name)
        return hv

    def energy_dev(self, n, d_xyz, d_q, k_idx, d_phi=None, stream=None):
        k_idx = np.ascontiguousarray(k_idx, dtype=np.int64)
        E = _d(); dg = np.zeros(2)
        self._check(lib().qem_energy_dev(self._h, int(n), d_xyz, d_q, ptr(k_idx), k_idx.shape[0],
                                         d_phi, C.byref(E), ptr(dg), stream), "qem_energy_dev")
        return E.value, dg

    def hvp_dev(self, n, d_xyz, d_q, k_idx, d_v, d_hv, stream=None, saved=False):
        k_idx = np.ascontiguousarray(k_idx, dtype=np.int64)
        name = "qem_hvp_saved_dev" if saved else "qem_hvp_dev"
        self._check(getattr(lib(), name)(self._h, int(n), d_xyz, d_q, ptr(k_idx),
                                         k_idx.shape[0], d_v, d_hv, stream), name)

    def timing_enable(self, on=True):
        self._check(lib().qem_timing_enable(self._h, int(bool(on))), "qem_timing_enable")

    def timing_read(self, reset=True):
        ms = np.zeros(4); calls = np.zeros(4, dtype=np.int64)
        self._check(lib().qem_timing_read(self._h, ptr(ms), ptr(calls), int(bool(reset))),
                    "qem_timing_read")
        return {ph: (float(ms[i]), int(calls[i])) for i, ph in enumerate(PHASES)}

    def bessel01_table(self, x):
        """J0, J1 from the context's table (FP32 cubic or FP64 degree 7)."""
        x = f64(x)
        j0 = np.zeros_like(x); j1 = np.zeros_like(x)
        self._check(lib().qem_bessel01_table(self._h, x.shape[0], ptr(x), ptr(j0), ptr(j1)),
                    "qem_bessel01_table")
        return j0, j1


def gauss_legendre(n: int):
    x = np.zeros(int(n)); w = np.zeros(int(n))
    st = lib().qem_gauss_legendre(int(n), ptr(x), ptr(w))
    if st != QEM_OK:
        raise QEMError(st, "qem_gauss_legendre")
    return x, w


def inter_params(n_atoms, Lx, Ly, s=1.0, rbe_mode=True):
    a, rc, kc = _d(), _d(), _d()
    st = lib().qem_inter_params(int(n_atoms), float(Lx), float(Ly), float(s), int(bool(rbe_mode)),
                                C.byref(a), C.byref(rc), C.byref(kc))
    if st != QEM_OK:
        raise QEMError(st, "qem_inter_params")
    return a.value, rc.value, kc.value


class DeviceArray:
    """A device allocation made through the library (no torch needed)."""

    def __init__(self, host: np.ndarray):
        host = np.ascontiguousarray(host)
        self.shape, self.dtype, self.nbytes = host.shape, host.dtype, host.nbytes
        p = _vp()
        st = lib().qem_dev_malloc(C.byref(p), host.nbytes)
        if st != QEM_OK:
            raise QEMError(st, "qem_dev_malloc")
        self.ptr = p
        if host.nbytes:
            self.upload(host)

    def upload(self, host: np.ndarray):
        host = np.ascontiguousarray(host, dtype=self.dtype)
        st = lib().qem_memcpy_htod(self.ptr, host.ctypes.data_as(_vp), host.nbytes)
        if st != QEM_OK:
            raise QEMError(st, "qem_memcpy_htod")

    def download(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.dtype)
        st = lib().qem_memcpy_dtoh(out.ctypes.data_as(_vp), self.ptr, self.nbytes)
        if st != QEM_OK:
            raise QEMError(st, "qem_memcpy_dtoh")
        return out

    def free(self):
        if self.ptr:
            lib().qem_dev_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def dev_copy(dst: "DeviceArray", src: "DeviceArray", nbytes: int):
    """Plain load/store copy kernel on the default stream (traffic calibration)."""
    if nbytes > min(dst.nbytes, src.nbytes):
        raise ValueError("dev_copy: nbytes exceeds a buffer")
    st = lib().qem_dev_copy_kernel(dst.ptr, src.ptr, nbytes, None)
    if st != QEM_OK:
        raise QEMError(st, "qem_dev_copy_kernel")


def synchronize():
    st = lib().qem_device_synchronize()
    if st != QEM_OK:
        raise QEMError(st, "qem_device_synchronize")
