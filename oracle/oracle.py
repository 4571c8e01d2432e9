"""ctypes binding of the CPU oracle (``libpartia_oracle.so``).

Test infrastructure only: tests/, ``__graft_entry__.smoke()`` and bench.py's
``cpu_baseline`` leg use it as the checker / CPU baseline.  The product package never
imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libpartia_oracle.so")
_lib = None

KERNELS = {"M4": 0, "M5": 1, "M6": 2, "C2": 3, "C4": 4, "C6": 5}
STRATEGIES = {"gather": 0, "scatter": 1, "symmetric": 2}

_d = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_u64 = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
_i64 = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_i32 = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_vp = C.c_void_p


class _Tree(C.Structure):
    _fields_ = [("n", C.c_int64), ("root", C.c_int32)] + [
        (name, C.c_void_p)
        for name in ("codes", "order", "left", "right", "escape", "parent", "lx", "ly", "lz",
                     "lh", "nminx", "nminy", "nminz", "nmaxx", "nmaxy", "nmaxz", "nhmax")
    ]


class _Catalog(C.Structure):
    _fields_ = [("n_scalar", C.c_int), ("scalar_slots", C.c_void_p), ("scalar_shepard", C.c_void_p),
                ("n_grad", C.c_int), ("grad_slots", C.c_void_p),
                ("n_div", C.c_int), ("div_slots", C.c_void_p),
                ("n_curl", C.c_int), ("curl_slots", C.c_void_p)]


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc, OpenMP, -ffp-contract=off)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-B" if force else "all"], cwd=_HERE, check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.orc_expand_bits3.restype = C.c_uint64
        L.orc_expand_bits3.argtypes = [C.c_uint64]
        L.orc_compact_bits3.restype = C.c_uint64
        L.orc_compact_bits3.argtypes = [C.c_uint64]
        L.orc_encode3.restype = C.c_uint64
        L.orc_encode3.argtypes = [C.c_uint64] * 3
        L.orc_quantize_axis.argtypes = [C.c_int64, _d, _u64]
        L.orc_morton_codes.argtypes = [C.c_int64, _d, _d, _d, _u64]
        L.orc_linear_bvh.restype = C.POINTER(_Tree)
        L.orc_linear_bvh.argtypes = [C.c_int64, _d, _d, _d, _d, _d, _d, C.c_int,
                                     C.POINTER(C.c_void_p), C.c_int]
        L.orc_lbvh_free.argtypes = [C.POINTER(_Tree)]
        for f in ("orc_kernel_f", "orc_kernel_df"):
            getattr(L, f).restype = C.c_double
            getattr(L, f).argtypes = [C.c_int, C.c_double]
        L.orc_jl_pow.restype = C.c_double
        L.orc_jl_pow.argtypes = [C.c_double, C.c_int64]
        for f in ("orc_kernel_norm3", "orc_kernel_support"):
            getattr(L, f).restype = C.c_double
            getattr(L, f).argtypes = [C.c_int]
        L.orc_W.restype = C.c_double
        L.orc_W.argtypes = [C.c_int, C.c_double, C.c_double]
        L.orc_line_W.restype = C.c_double
        L.orc_line_W.argtypes = [C.c_int, C.c_double, C.c_double]
        L.orc_gradW.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, _d]
        L.orc_line_table.argtypes = [C.c_int, _d, _d]
        L.orc_nearest.argtypes = [C.POINTER(_Tree), C.c_int64, _d, _d, _d, _i64, _d, C.c_int]
        L.orc_point_neighbors.restype = C.c_int64
        L.orc_point_neighbors.argtypes = [C.POINTER(_Tree), C.c_int, C.c_int, C.c_int64, _d, _d,
                                          _d, _vp, _vp, _vp, C.c_int64]
        L.orc_line_neighbors.restype = C.c_int64
        L.orc_line_neighbors.argtypes = [C.POINTER(_Tree), C.c_int, C.c_int, C.c_int64, _d, _d,
                                         _d, _d, _d, _d, _vp, _vp, _vp, C.c_int64]
        L.orc_point_visits.argtypes = [C.POINTER(_Tree), C.c_int, C.c_int, C.c_int64, _d, _d, _d,
                                       _i64, _i64, _i64, C.c_int]
        L.orc_line_visits.argtypes = [C.POINTER(_Tree), C.c_int, C.c_int64, _d, _d, _d, _d, _d,
                                      _d, _i64, _i64, _i64, C.c_int]
        L.orc_point_samples.restype = C.c_int
        L.orc_point_samples.argtypes = [C.POINTER(_Tree), _d, _d, C.c_int, C.POINTER(C.c_void_p),
                                        C.c_int, C.c_int, C.POINTER(_Catalog), C.c_int64, _d, _d,
                                        _d, C.POINTER(C.c_void_p), C.c_int]
        L.orc_line_samples.restype = C.c_int
        L.orc_line_samples.argtypes = [C.POINTER(_Tree), _d, _d, C.c_int, C.POINTER(C.c_void_p),
                                       C.c_int, C.c_int, _vp, C.c_int, _i32, _vp, C.c_int64,
                                       _d, _d, _d, _d, _d, _d, C.POINTER(C.c_void_p), C.c_int]
        _lib = L
    return _lib


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _ptrs(arrs):
    return (C.c_void_p * max(1, len(arrs)))(*[a.ctypes.data for a in arrs])


def kernel_id(k) -> int:
    return KERNELS[k] if isinstance(k, str) else int(k)


def strategy_id(s) -> int:
    return STRATEGIES[s] if isinstance(s, str) else int(s)


# ---------------------------------------------------------------- Morton ----
def expand_bits3(x: int) -> int:
    return lib().orc_expand_bits3(x)


def compact_bits3(x: int) -> int:
    return lib().orc_compact_bits3(x)


def encode3(ix: int, iy: int, iz: int) -> int:
    return lib().orc_encode3(ix, iy, iz)


def quantize_axis(v) -> np.ndarray:
    v = _f64(v)
    out = np.empty(v.size, np.uint64)
    lib().orc_quantize_axis(v.size, v, out)
    return out


def morton_codes(x, y, z) -> np.ndarray:
    x, y, z = _f64(x), _f64(y), _f64(z)
    out = np.empty(x.size, np.uint64)
    lib().orc_morton_codes(x.size, x, y, z, out)
    return out


# --------------------------------------------------------------- kernels ----
def kernel_f(k, q):
    return lib().orc_kernel_f(kernel_id(k), float(q))


def kernel_df(k, q):
    return lib().orc_kernel_df(kernel_id(k), float(q))


def kernel_norm3(k):
    return lib().orc_kernel_norm3(kernel_id(k))


def kernel_support(k):
    return lib().orc_kernel_support(kernel_id(k))


def jl_pow(x, n):
    return lib().orc_jl_pow(float(x), int(n))


def W(k, r, h):
    return lib().orc_W(kernel_id(k), float(r), float(h))


def gradW(k, dx, dy, dz, h):
    g = np.zeros(3)
    lib().orc_gradW(kernel_id(k), float(dx), float(dy), float(dz), float(h), g)
    return g


def line_W(k, r, h):
    return lib().orc_line_W(kernel_id(k), float(r), float(h))


def line_table(k):
    Q = np.empty(4096)
    I = np.empty(4096)
    lib().orc_line_table(kernel_id(k), Q, I)
    return Q, I


# ------------------------------------------------------------------ LBVH ----
class LBVH:
    """LinearBVH!(input, Val(3)) on FP64 columns.  Like the reference, the columns passed
    in are permuted into Morton order in place (InterpolationInput.jl:211-219)."""

    def __init__(self, x, y, z, m, h, rho, quant=(), nthreads=0):
        cols = [x, y, z, m, h, rho] + list(quant)
        for c in cols:
            if not (isinstance(c, np.ndarray) and c.dtype == np.float64 and c.flags.c_contiguous):
                raise TypeError("oracle LBVH columns must be C-contiguous float64 arrays")
        n = x.size
        self._cols = cols
        self.n = n
        qp = _ptrs(list(quant))
        self._t = lib().orc_linear_bvh(n, x, y, z, m, h, rho, len(quant), qp, int(nthreads))
        if not self._t:
            raise ValueError(f"oracle LinearBVH!: invalid particle count {n}")
        t = self._t.contents
        self.root = t.root

        def arr(ptr, count, dt):
            return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(dt)), shape=(count,)).copy()

        self.codes = arr(t.codes, n, C.c_uint64).astype(np.uint64)
        self.order = arr(t.order, n, C.c_int64)
        nt = 2 * n
        self.left = arr(t.left, nt, C.c_int32)[1:]
        self.right = arr(t.right, nt, C.c_int32)[1:]
        self.escape = arr(t.escape, nt, C.c_int32)[1:]
        self.parent = arr(t.parent, nt, C.c_int32)[1:]
        self.leaf = [arr(getattr(t, f), n, C.c_double) for f in ("lx", "ly", "lz", "lh")]
        self.node_min = [arr(getattr(t, f), n, C.c_double)[1:] for f in ("nminx", "nminy", "nminz")]
        self.node_max = [arr(getattr(t, f), n, C.c_double)[1:] for f in ("nmaxx", "nmaxy", "nmaxz")]
        self.node_hmax = arr(t.nhmax, n, C.c_double)[1:]

    def __del__(self):
        if getattr(self, "_t", None):
            lib().orc_lbvh_free(self._t)
            self._t = None

    # ---- queries ----
    def nearest(self, qx, qy, qz, nthreads=0):
        qx, qy, qz = _f64(qx), _f64(qy), _f64(qz)
        idx = np.empty(qx.size, np.int64)
        h = np.empty(qx.size)
        lib().orc_nearest(self._t, qx.size, qx, qy, qz, idx, h, int(nthreads))
        return idx, h

    def point_neighbors(self, kernel, strategy, qx, qy, qz, ha=None):
        qx, qy, qz = _f64(qx), _f64(qy), _f64(qz)
        ha_p = None if ha is None else _f64(ha)
        L = lib()
        args = (self._t, kernel_id(kernel), strategy_id(strategy), qx.size, qx, qy, qz,
                None if ha_p is None else ha_p.ctypes.data)
        total = L.orc_point_neighbors(*args, None, None, 0)
        off = np.empty(qx.size + 1, np.int64)
        idx = np.empty(max(total, 1), np.int32)
        L.orc_point_neighbors(*args, off.ctypes.data, idx.ctypes.data, total)
        return off, idx[:total]

    def line_neighbors(self, kernel, strategy, o, d, ha=None):
        o = [_f64(a) for a in o]
        d = [_f64(a) for a in d]
        ha_p = None if ha is None else _f64(ha)
        L = lib()
        args = (self._t, kernel_id(kernel), strategy_id(strategy), o[0].size, *o, *d,
                None if ha_p is None else ha_p.ctypes.data)
        total = L.orc_line_neighbors(*args, None, None, 0)
        off = np.empty(o[0].size + 1, np.int64)
        idx = np.empty(max(total, 1), np.int32)
        L.orc_line_neighbors(*args, off.ctypes.data, idx.ctypes.data, total)
        return off, idx[:total]

    def point_visits(self, kernel, strategy, qx, qy, qz, nthreads=0):
        qx, qy, qz = _f64(qx), _f64(qy), _f64(qz)
        n = qx.size
        ni, nl, na = np.empty(n, np.int64), np.empty(n, np.int64), np.empty(n, np.int64)
        lib().orc_point_visits(self._t, kernel_id(kernel), strategy_id(strategy), n, qx, qy, qz,
                               ni, nl, na, int(nthreads))
        return ni, nl, na

    def line_visits(self, kernel, o, d, nthreads=0):
        o = [_f64(a) for a in o]
        d = [_f64(a) for a in d]
        n = o[0].size
        ni, nl, na = np.empty(n, np.int64), np.empty(n, np.int64), np.empty(n, np.int64)
        lib().orc_line_visits(self._t, kernel_id(kernel), n, *o, *d, ni, nl, na, int(nthreads))
        return ni, nl, na

    # ---- interpolation ----
    def point_samples(self, kernel, strategy, catalog, qx, qy, qz, nthreads=0):
        """catalog: dict(scalar_slots, scalar_shepard, grad_slots, div_slots, curl_slots),
        0-based quant slots, grad slot -1 = density gradient.  Returns L output arrays."""
        x, y, z, m, h, rho = self._cols[:6]
        quant = self._cols[6:]
        qx, qy, qz = _f64(qx), _f64(qy), _f64(qz)
        cat, keep = make_catalog(catalog)
        L = cat.n_scalar + 3 * cat.n_grad + cat.n_div + 3 * cat.n_curl
        outs = [np.empty(qx.size) for _ in range(L)]
        rc = lib().orc_point_samples(self._t, m, rho, len(quant), _ptrs(quant), kernel_id(kernel),
                                     strategy_id(strategy), C.byref(cat), qx.size, qx, qy, qz,
                                     _ptrs(outs), int(nthreads))
        if rc != 0:
            raise ValueError("oracle point_samples: catalog too large")
        del keep
        return outs

    def line_samples(self, kernel, strategy, slots, shepard, o, d, ha=None, nthreads=0):
        x, y, z, m, h, rho = self._cols[:6]
        quant = self._cols[6:]
        o = [_f64(a) for a in o]
        d = [_f64(a) for a in d]
        slots = np.ascontiguousarray(slots, np.int32)
        shep = np.ascontiguousarray(shepard, np.uint8)
        ha_p = None if ha is None else _f64(ha)
        outs = [np.empty(o[0].size) for _ in range(slots.size)]
        rc = lib().orc_line_samples(self._t, m, rho, len(quant), _ptrs(quant), kernel_id(kernel),
                                    strategy_id(strategy),
                                    None if ha_p is None else ha_p.ctypes.data, slots.size, slots,
                                    shep.ctypes.data, o[0].size, *o, *d, _ptrs(outs),
                                    int(nthreads))
        if rc != 0:
            raise ValueError("oracle line_samples: too many columns")
        return outs


def make_catalog(c):
    keep = {}

    def arr(name, dt, default):
        a = np.ascontiguousarray(c.get(name, default), dtype=dt).ravel()
        keep[name] = a
        return a

    ss = arr("scalar_slots", np.int32, [])
    sh = arr("scalar_shepard", np.uint8, [1] * ss.size)
    gs = arr("grad_slots", np.int32, [])
    ds = arr("div_slots", np.int32, [])
    cs = arr("curl_slots", np.int32, [])
    cat = _Catalog(ss.size, ss.ctypes.data, sh.ctypes.data, gs.size, gs.ctypes.data,
                   ds.size // 3, ds.ctypes.data, cs.size // 3, cs.ctypes.data)
    return cat, keep
