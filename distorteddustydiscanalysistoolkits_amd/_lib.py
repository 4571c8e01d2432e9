"""ctypes binding of libpartia_mi355x.so (include/partia_mi355x.h).

The shared library is the product: there is no Python or CPU fallback.  If it is missing
or cannot be loaded, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpartia_mi355x.so")

PMX_OK, PMX_ERR_ARGUMENT, PMX_ERR_DIMENSION, PMX_ERR_STATE, PMX_ERR_HIP, PMX_ERR_NOMEM, PMX_ERR_CAPACITY = (
    0, -1, -2, -3, -4, -5, -6)
PMX_ERR_FORMAT, PMX_ERR_IO, PMX_ERR_INTERNAL = -7, -8, -9
PMX_I8, PMX_I16, PMX_I32, PMX_I64 = 10, 11, 12, 13
PMX_F32, PMX_F64 = 0, 1
PMX_MEM_HOST, PMX_MEM_DEVICE = 0, 1
PMX_GATHER, PMX_SCATTER, PMX_SYMMETRIC = 0, 1, 2
PMX_PREC_F64, PMX_PREC_F32C = 0, 1
PMX_BUILD_WRITEBACK = 1
PMX_LINE_INV_H2 = 0x100     # OR'd into a line call's precision: weight Inorm / h^2 (default Inorm / h)

EXPORTS = ("pmx_create", "pmx_destroy", "pmx_last_error", "pmx_version", "pmx_build_lbvh",
           "pmx_interpolate_points", "pmx_interpolate_cartesian_grid", "pmx_interpolate_lines",
           "pmx_interpolate_lines_noalloc",
           "pmx_point_samples_interpolation",
           "pmx_line_samples_interpolation", "pmx_export_tree", "pmx_export_nodes",
           "pmx_export_particles", "pmx_nearest", "pmx_point_neighbors", "pmx_line_neighbors",
           "pmx_kernel_eval", "pmx_line_kernel_table", "pmx_last_timings", "pmx_point_traversal_stats",
           "pmx_line_traversal_stats", "pmx_line_neighbor_digest", "pmx_batch_create", "pmx_batch_run",
           "pmx_batch_interpolate", "pmx_batch_tune", "pmx_batch_info", "pmx_batch_last_error",
           "pmx_batch_destroy",
           # include/partia_phantom.h
           "pmx_phantom_open", "pmx_phantom_open_error", "pmx_phantom_close", "pmx_phantom_last_error",
           "pmx_phantom_info", "pmx_phantom_file_identifier", "pmx_phantom_block_size", "pmx_phantom_column_info",
           "pmx_phantom_header_var", "pmx_phantom_read_column", "pmx_phantom_density")
PMX_JOB_POINTS, PMX_JOB_LINES = 0, 1


class PmxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[pmx {code}] {msg}")
        self.code = code
        self.msg = msg


class ArgumentError(PmxError, ValueError):
    """≙ Julia ArgumentError."""


class DimensionMismatch(PmxError, ValueError):
    """≙ Julia DimensionMismatch."""


class FormatError(PmxError):
    """≙ the ErrorException read_phantom raises for a file that is not a Phantom dump."""


class PmxIOError(PmxError, OSError):
    """≙ Julia SystemError (the file cannot be opened or read)."""


def _raise(code, msg):
    if code == PMX_ERR_ARGUMENT:
        raise ArgumentError(code, msg)
    if code == PMX_ERR_DIMENSION:
        raise DimensionMismatch(code, msg)
    if code == PMX_ERR_FORMAT:
        raise FormatError(code, msg)
    if code == PMX_ERR_IO:
        raise PmxIOError(code, msg)
    raise PmxError(code, msg)


class Particles(C.Structure):
    _fields_ = [("dtype", C.c_int32), ("mem", C.c_int32), ("n", C.c_int64),
                ("x", C.c_void_p), ("y", C.c_void_p), ("z", C.c_void_p),
                ("m", C.c_void_p), ("h", C.c_void_p), ("rho", C.c_void_p),
                ("ncol", C.c_int32), ("quant", C.POINTER(C.c_void_p)), ("kernel", C.c_int32)]


class Catalog(C.Structure):
    _fields_ = [("n_scalar", C.c_int32), ("scalar_slots", C.c_void_p), ("scalar_shepard", C.c_void_p),
                ("n_grad", C.c_int32), ("grad_slots", C.c_void_p),
                ("n_div", C.c_int32), ("div_slots", C.c_void_p),
                ("n_curl", C.c_int32), ("curl_slots", C.c_void_p)]


class Points(C.Structure):
    _fields_ = [("mem", C.c_int32), ("n", C.c_int64), ("x", C.c_void_p), ("y", C.c_void_p), ("z", C.c_void_p)]


class Axes(C.Structure):
    _fields_ = [("mem", C.c_int32), ("n", C.c_int64 * 3), ("a", C.c_void_p * 3)]


class Lines(C.Structure):
    _fields_ = [("mem", C.c_int32), ("n", C.c_int64), ("ox", C.c_void_p), ("oy", C.c_void_p),
                ("oz", C.c_void_p), ("dx", C.c_void_p), ("dy", C.c_void_p), ("dz", C.c_void_p)]


class Output(C.Structure):
    _fields_ = [("dtype", C.c_int32), ("mem", C.c_int32), ("cols", C.POINTER(C.c_void_p))]


class BatchJob(C.Structure):
    _fields_ = [("particles", Particles), ("kind", C.c_int32), ("strategy", C.c_int32), ("points", Points),
                ("lines", Lines), ("catalog", Catalog), ("out", Output)]


_lib = None


def lib():
    """Load the HIP library; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with "
                               "`python -m distorteddustydiscanalysistoolkits_amd.build`")
        L = C.CDLL(LIB_PATH)
        vp, i32, i64 = C.c_void_p, C.c_int32, C.c_int64
        L.pmx_create.argtypes = [C.POINTER(vp), i32]
        L.pmx_destroy.argtypes = [vp]
        L.pmx_destroy.restype = None
        L.pmx_last_error.argtypes = [vp]
        L.pmx_last_error.restype = C.c_char_p
        L.pmx_version.restype = C.c_char_p
        L.pmx_build_lbvh.argtypes = [vp, C.POINTER(Particles), i32, vp, vp]
        L.pmx_interpolate_points.argtypes = [vp, i32, i32, C.POINTER(Catalog), C.POINTER(Points), C.POINTER(Output)]
        L.pmx_interpolate_lines.argtypes = [vp, i32, i32, C.POINTER(Catalog), C.POINTER(Lines), C.POINTER(Output)]
        L.pmx_interpolate_lines_noalloc.argtypes = [vp, i32, i32, C.POINTER(Catalog), C.POINTER(Lines),
                                                    C.POINTER(Output)]
        L.pmx_interpolate_cartesian_grid.argtypes = [vp, i32, i32, C.POINTER(Catalog), C.POINTER(Axes),
                                                     C.POINTER(Output)]
        L.pmx_point_samples_interpolation.argtypes = [vp, C.POINTER(Particles), i32, i32, C.POINTER(Catalog),
                                                      C.POINTER(Points), C.POINTER(Output)]
        L.pmx_line_samples_interpolation.argtypes = [vp, C.POINTER(Particles), i32, i32, C.POINTER(Catalog),
                                                     C.POINTER(Lines), C.POINTER(Output)]
        L.pmx_export_tree.argtypes = [vp, vp, vp, vp, vp]
        L.pmx_export_nodes.argtypes = [vp, vp, vp, vp]
        L.pmx_export_particles.argtypes = [vp, C.POINTER(vp), i32]
        L.pmx_nearest.argtypes = [vp, i32, C.POINTER(Points), vp]
        L.pmx_point_neighbors.argtypes = [vp, i32, i32, C.POINTER(Points), vp, vp, vp, vp, i64]
        L.pmx_line_neighbors.argtypes = [vp, i32, i32, C.POINTER(Lines), vp, vp, vp, vp, i64]
        L.pmx_kernel_eval.argtypes = [vp, i32, i64, vp, vp, vp]
        L.pmx_line_kernel_table.argtypes = [i32, vp, vp]
        L.pmx_last_timings.argtypes = [vp, vp, i32]
        L.pmx_point_traversal_stats.argtypes = [vp, i32, C.POINTER(Points), vp]
        L.pmx_line_traversal_stats.argtypes = [vp, C.POINTER(Lines), vp]
        L.pmx_line_neighbor_digest.argtypes = [vp, C.POINTER(Lines), vp, C.POINTER(i32)]
        L.pmx_batch_create.argtypes = [C.POINTER(vp), i32, i32, C.POINTER(BatchJob)]
        L.pmx_batch_run.argtypes = [vp, i32]
        L.pmx_batch_interpolate.argtypes = [vp]
        L.pmx_batch_tune.argtypes = [vp]
        L.pmx_batch_info.argtypes = [vp, vp, i32]
        L.pmx_batch_last_error.argtypes = [vp]
        L.pmx_batch_last_error.restype = C.c_char_p
        L.pmx_phantom_open.argtypes = [C.POINTER(vp), C.c_char_p]
        L.pmx_phantom_open_error.argtypes = []
        L.pmx_phantom_open_error.restype = C.c_char_p
        L.pmx_phantom_close.argtypes = [vp]
        L.pmx_phantom_close.restype = None
        L.pmx_phantom_last_error.argtypes = [vp]
        L.pmx_phantom_last_error.restype = C.c_char_p
        L.pmx_phantom_info.argtypes = [vp, vp, i32]
        L.pmx_phantom_file_identifier.argtypes = [vp, C.c_char_p, i32]
        L.pmx_phantom_block_size.argtypes = [vp, i32, C.POINTER(i64)]
        L.pmx_phantom_column_info.argtypes = [vp, i32, C.c_char_p, C.POINTER(i32), C.POINTER(i32)]
        L.pmx_phantom_header_var.argtypes = [vp, i32, C.c_char_p, C.POINTER(i32), C.POINTER(C.c_double),
                                             C.POINTER(i64)]
        L.pmx_phantom_read_column.argtypes = [vp, i32, i32, i32, vp, vp, i64]
        L.pmx_phantom_density.argtypes = [i32, i32, i32, i64, vp, C.c_double, C.c_double, vp]
        L.pmx_batch_destroy.argtypes = [vp]
        L.pmx_batch_destroy.restype = None
        _lib = L
    return _lib


def _ptr(a):
    """Address of a numpy array or a torch tensor (device memory)."""
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        return a.ctypes.data
    return a.data_ptr()


def _mem_of(arrs):
    import sys
    torch = sys.modules.get("torch")
    if torch is not None and any(isinstance(a, torch.Tensor) for a in arrs):
        if not all(isinstance(a, torch.Tensor) and a.is_cuda for a in arrs):
            raise ValueError("mixing host arrays and tensors, or CPU tensors, is not supported")
        return PMX_MEM_DEVICE
    return PMX_MEM_HOST


def _dtype_code(a):
    s = str(a.dtype)
    if s in ("float64", "torch.float64"):
        return PMX_F64
    if s in ("float32", "torch.float32"):
        return PMX_F32
    raise TypeError(f"unsupported element type {a.dtype}")


def line_kernel_table(kernel: int):
    Q = np.empty(4096)
    I = np.empty(4096)
    rc = lib().pmx_line_kernel_table(int(kernel), Q.ctypes.data, I.ctypes.data)
    if rc != PMX_OK:
        _raise(rc, "pmx_line_kernel_table")
    return Q, I


def catalog_struct(scalar_slots=(), scalar_shepard=None, grad_slots=(), div_slots=(), curl_slots=()):
    """pmx_catalog from 0-based slots; keeps the backing arrays alive on the struct."""
    ss = np.ascontiguousarray(scalar_slots, np.int32).ravel()
    sh = np.ascontiguousarray([1] * ss.size if scalar_shepard is None else scalar_shepard, np.uint8).ravel()
    gs = np.ascontiguousarray(grad_slots, np.int32).ravel()
    ds = np.ascontiguousarray(div_slots, np.int32).ravel()
    cs = np.ascontiguousarray(curl_slots, np.int32).ravel()
    cat = Catalog(ss.size, ss.ctypes.data, sh.ctypes.data, gs.size, gs.ctypes.data, ds.size // 3,
                  ds.ctypes.data, cs.size // 3, cs.ctypes.data)
    cat._keep = (ss, sh, gs, ds, cs)
    return cat


class Context:
    """One pmx_ctx: a HIP stream plus the device-resident LBVH and particle columns."""

    def __init__(self, device: int = 0):
        L = lib()
        h = C.c_void_p()
        rc = L.pmx_create(C.byref(h), int(device))
        if rc != PMX_OK:
            raise PmxError(rc, "pmx_create failed (no usable HIP device?)")
        self._h = h
        self.n = 0
        self.ncol = 0

    def close(self):
        if getattr(self, "_h", None) and _lib is not None:     # (module globals are gone at exit)
            _lib.pmx_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def _check(self, rc):
        if rc != PMX_OK:
            _raise(rc, lib().pmx_last_error(self._h).decode())

    # ---- build ----
    def build(self, x, y, z, m, h, rho, quant=(), kernel=0, writeback=False, want_codes=False,
              want_order=False):
        cols = [x, y, z, m, h, rho] + list(quant)
        mem = _mem_of(cols)
        dt = _dtype_code(x)
        n = int(x.shape[0])
        for c in cols:
            if _dtype_code(c) != dt:
                raise TypeError("all particle columns must share one element type")
        qp = (C.c_void_p * max(1, len(quant)))(*[_ptr(q) for q in quant])
        p = Particles(dt, mem, n, _ptr(x), _ptr(y), _ptr(z), _ptr(m), _ptr(h), _ptr(rho), len(quant),
                      C.cast(qp, C.POINTER(C.c_void_p)), int(kernel))
        codes = np.empty(n, np.uint64) if want_codes else None
        order = np.empty(n, np.int64) if want_order else None
        rc = lib().pmx_build_lbvh(self._h, C.byref(p), PMX_BUILD_WRITEBACK if writeback else 0,
                                  _ptr(codes), _ptr(order))
        self._check(rc)
        self.n, self.ncol = n, len(quant)
        return codes, order

    # ---- interpolation ----
    def _points(self, qx, qy, qz):
        mem = _mem_of([qx, qy, qz])
        return Points(mem, int(qx.shape[0]), _ptr(qx), _ptr(qy), _ptr(qz))

    def _lines(self, origin, direction):
        arrs = list(origin) + list(direction)
        mem = _mem_of(arrs)
        return Lines(mem, int(origin[0].shape[0]), *[_ptr(a) for a in arrs])

    def interpolate_points(self, strategy, precision, cat, qx, qy, qz, out_cols):
        pts = self._points(qx, qy, qz)
        mem = _mem_of(out_cols) if out_cols else PMX_MEM_HOST
        dt = _dtype_code(out_cols[0]) if out_cols else PMX_F64
        cp = (C.c_void_p * max(1, len(out_cols)))(*[_ptr(o) for o in out_cols])
        out = Output(dt, mem, C.cast(cp, C.POINTER(C.c_void_p)))
        self._check(lib().pmx_interpolate_points(self._h, int(strategy), int(precision), C.byref(cat),
                                                 C.byref(pts), C.byref(out)))

    def interpolate_cartesian_grid(self, strategy, precision, cat, xa, ya, za, out_cols):
        """Structured Cartesian grid from its axes; samples flattened on the device in column-major
        order (x fastest); out_cols hold len(xa) * len(ya) * len(za) values in that order."""
        ax = Axes(_mem_of([xa, ya, za]), (C.c_int64 * 3)(*(int(a.shape[0]) for a in (xa, ya, za))),
                  (C.c_void_p * 3)(*(_ptr(a) for a in (xa, ya, za))))
        mem = _mem_of(out_cols) if out_cols else PMX_MEM_HOST
        dt = _dtype_code(out_cols[0]) if out_cols else PMX_F64
        cp = (C.c_void_p * max(1, len(out_cols)))(*[_ptr(o) for o in out_cols])
        out = Output(dt, mem, C.cast(cp, C.POINTER(C.c_void_p)))
        self._check(lib().pmx_interpolate_cartesian_grid(self._h, int(strategy), int(precision), C.byref(cat),
                                                         C.byref(ax), C.byref(out)))

    def interpolate_lines(self, strategy, precision, cat, origin, direction, out_cols):
        ln = self._lines(origin, direction)
        mem = _mem_of(out_cols) if out_cols else PMX_MEM_HOST
        dt = _dtype_code(out_cols[0]) if out_cols else PMX_F64
        cp = (C.c_void_p * max(1, len(out_cols)))(*[_ptr(o) for o in out_cols])
        out = Output(dt, mem, C.cast(cp, C.POINTER(C.c_void_p)))
        self._check(lib().pmx_interpolate_lines(self._h, int(strategy), int(precision), C.byref(cat),
                                                C.byref(ln), C.byref(out)))

    def interpolate_lines_noalloc(self, strategy, precision, cat, origin, direction, out_cols):
        """pmx_interpolate_lines_noalloc: device lines and outputs (torch tensors), no allocation."""
        ln = self._lines(origin, direction)
        mem = _mem_of(out_cols) if out_cols else PMX_MEM_HOST
        dt = _dtype_code(out_cols[0]) if out_cols else PMX_F32
        cp = (C.c_void_p * max(1, len(out_cols)))(*[_ptr(o) for o in out_cols])
        out = Output(dt, mem, C.cast(cp, C.POINTER(C.c_void_p)))
        self._check(lib().pmx_interpolate_lines_noalloc(self._h, int(strategy), int(precision), C.byref(cat),
                                                        C.byref(ln), C.byref(out)))

    # ---- inspection ----
    def export_tree(self):
        t = 2 * self.n - 1
        arrs = [np.empty(t, np.int32) for _ in range(4)]
        self._check(lib().pmx_export_tree(self._h, *[a.ctypes.data for a in arrs]))
        return arrs  # left, right, parent, escape

    def export_nodes(self):
        m = max(self.n - 1, 0)
        mn, mx, hm = np.empty((3, m)), np.empty((3, m)), np.empty(m)
        self._check(lib().pmx_export_nodes(self._h, mn.ctypes.data, mx.ctypes.data, hm.ctypes.data))
        return mn, mx, hm

    def export_particles(self):
        k = 6 + self.ncol
        arrs = [np.empty(self.n) for _ in range(k)]
        cp = (C.c_void_p * k)(*[a.ctypes.data for a in arrs])
        self._check(lib().pmx_export_particles(self._h, cp, k))
        return arrs

    def nearest(self, qx, qy, qz, precision=PMX_PREC_F64):
        pts = self._points(qx, qy, qz)
        idx = np.empty(int(qx.shape[0]), np.int64)
        self._check(lib().pmx_nearest(self._h, int(precision), C.byref(pts), idx.ctypes.data))
        return idx

    def point_neighbors(self, strategy, qx, qy, qz, ha=None, precision=PMX_PREC_F64):
        pts = self._points(qx, qy, qz)
        n = int(qx.shape[0])
        counts = np.empty(n, np.int64)
        hap = None if ha is None else np.ascontiguousarray(ha, np.float64)
        self._check(lib().pmx_point_neighbors(self._h, int(strategy), int(precision), C.byref(pts), _ptr(hap),
                                              counts.ctypes.data, None, None, 0))
        total = int(counts.sum())
        off = np.empty(n + 1, np.int64)
        idx = np.empty(max(total, 1), np.int32)
        self._check(lib().pmx_point_neighbors(self._h, int(strategy), int(precision), C.byref(pts), _ptr(hap),
                                              counts.ctypes.data, off.ctypes.data, idx.ctypes.data, total))
        return off, idx[:total]

    def line_neighbors(self, strategy, origin, direction, ha=None, precision=PMX_PREC_F64):
        ln = self._lines(origin, direction)
        n = int(origin[0].shape[0])
        counts = np.empty(n, np.int64)
        hap = None if ha is None else np.ascontiguousarray(ha, np.float64)
        self._check(lib().pmx_line_neighbors(self._h, int(strategy), int(precision), C.byref(ln), _ptr(hap),
                                             counts.ctypes.data, None, None, 0))
        total = int(counts.sum())
        off = np.empty(n + 1, np.int64)
        idx = np.empty(max(total, 1), np.int32)
        self._check(lib().pmx_line_neighbors(self._h, int(strategy), int(precision), C.byref(ln), _ptr(hap),
                                             counts.ctypes.data, off.ctypes.data, idx.ctypes.data, total))
        return off, idx[:total]

    def point_stats(self, strategy, qx, qy, qz):
        pts = self._points(qx, qy, qz)
        st = np.zeros(5, np.int64)
        self._check(lib().pmx_point_traversal_stats(self._h, int(strategy), C.byref(pts), st.ctypes.data))
        return st

    def line_stats(self, origin, direction):
        ln = self._lines(origin, direction)
        st = np.zeros(5, np.int64)
        self._check(lib().pmx_line_traversal_stats(self._h, C.byref(ln), st.ctypes.data))
        return st

    def line_digest(self, origin, direction):
        """(count, index sum, index-square sum) per line from the tiled line-map path, or None
        when the lines do not take that path."""
        ln = self._lines(origin, direction)
        n = int(ln.n)
        dg = np.zeros(3 * max(n, 1), np.uint64)
        used = C.c_int32(0)
        self._check(lib().pmx_line_neighbor_digest(self._h, C.byref(ln), dg.ctypes.data, C.byref(used)))
        if not used.value:
            return None
        return dg[:n], dg[n:2 * n], dg[2 * n:3 * n]

    def kernel_eval(self, kernel, q):
        q = np.ascontiguousarray(q, np.float64)
        f, df = np.empty_like(q), np.empty_like(q)
        self._check(lib().pmx_kernel_eval(self._h, int(kernel), q.size, q.ctypes.data, f.ctypes.data,
                                          df.ctypes.data))
        return f, df

    def timings(self):
        """Phase times (ms) of the last calls, then the split counts of the last FP32c
        gather / symmetric point pass ([8] tiles split, [9] items queued)."""
        t = np.zeros(10)
        lib().pmx_last_timings(self._h, t.ctypes.data, 10)
        return t


class Batch:
    """pmx_batch: many snapshots (device particle columns + device samples), one resident LBVH
    each, interpolated by one persistent kernel (optionally replayed as a hipGraph).

    jobs: list of dicts with keys
      cols (x, y, z, m, h, rho), quant (list), kernel, kind ("points" | "lines"),
      strategy, samples ((qx, qy, qz) or ((ox, oy, oz), (dx, dy, dz))), scalar_slots,
      scalar_shepard, grad_slots / div_slots / curl_slots (point jobs only, optional), out (list
      of FP32 device tensors, one per output column in catalog order: scalars, 3 per gradient,
      1 per divergence, 3 per curl)."""

    def __init__(self, jobs, device: int = 0):
        self._keep = []
        arr = (BatchJob * len(jobs))()
        for j, J in enumerate(jobs):
            cols = list(J["cols"])
            quant = list(J.get("quant", ()))
            qp = (C.c_void_p * max(1, len(quant)))(*[_ptr(q) for q in quant])
            part = Particles(_dtype_code(cols[0]), _mem_of(cols + quant), int(cols[0].shape[0]),
                             *[_ptr(c) for c in cols], len(quant), C.cast(qp, C.POINTER(C.c_void_p)),
                             int(J.get("kernel", 0)))
            cat = catalog_struct(scalar_slots=J.get("scalar_slots", ()), scalar_shepard=J.get("scalar_shepard"),
                                 grad_slots=J.get("grad_slots", ()), div_slots=J.get("div_slots", ()),
                                 curl_slots=J.get("curl_slots", ()))
            outs = list(J["out"])
            L = cat.n_scalar + 3 * cat.n_grad + cat.n_div + 3 * cat.n_curl
            if len(outs) != L:
                raise DimensionMismatch(PMX_ERR_DIMENSION, f"batch job {j}: {len(outs)} output columns for {L} outputs")
            op = (C.c_void_p * max(1, len(outs)))(*[_ptr(o) for o in outs])
            out = Output(PMX_F32, PMX_MEM_DEVICE, C.cast(op, C.POINTER(C.c_void_p)))
            job = arr[j]
            job.particles = part
            job.catalog = cat
            job.out = out
            if J["kind"] == "points":
                qx, qy, qz = J["samples"]
                job.kind = PMX_JOB_POINTS
                job.strategy = int(J.get("strategy", PMX_SYMMETRIC))
                job.points = Points(_mem_of([qx, qy, qz]), int(qx.shape[0]), _ptr(qx), _ptr(qy), _ptr(qz))
            else:
                o, d = J["samples"]
                job.kind = PMX_JOB_LINES
                job.strategy = PMX_SCATTER
                job.lines = Lines(_mem_of(list(o) + list(d)), int(o[0].shape[0]), *[_ptr(a) for a in list(o) + list(d)])
            self._keep.append((qp, cat, op, J))
        self._h = C.c_void_p()
        rc = lib().pmx_batch_create(C.byref(self._h), int(device), len(jobs), arr)
        if rc != PMX_OK:
            msg = lib().pmx_batch_last_error(self._h).decode() if self._h else "pmx_batch_create"
            self.close()
            _raise(rc, msg)

    def _check(self, rc):
        if rc != PMX_OK:
            _raise(rc, lib().pmx_batch_last_error(self._h).decode())

    def run(self, graph: bool = False):
        self._check(lib().pmx_batch_run(self._h, 1 if graph else 0))

    def interpolate(self):
        self._check(lib().pmx_batch_interpolate(self._h))

    def tune(self):
        self._check(lib().pmx_batch_tune(self._h))

    def info(self):
        v = np.zeros(9)
        lib().pmx_batch_info(self._h, v.ctypes.data, 9)
        return dict(run_ms=v[0], build_ms=v[1], interp_ms=v[2], tiles=int(v[3]), waves_per_cu=int(v[4]),
                    bucket_points=int(v[5]), bucket_lines=int(v[6]), heavy_first=int(v[7]),
                    split_budget=int(v[8]))

    def close(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.pmx_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- Phantom dump reader (include/partia_phantom.h) ----
_NP_OF = {PMX_I8: np.int8, PMX_I16: np.int16, PMX_I32: np.int32, PMX_I64: np.int64, PMX_F32: np.float32,
          PMX_F64: np.float64}
_CODE_OF = {np.dtype(v): k for k, v in _NP_OF.items()}


def dtype_code_np(dt) -> int:
    """PMX element-type code of a numpy / torch dtype."""
    s = str(dt).replace("torch.", "")
    return _CODE_OF[np.dtype(s)]


class Phantom:
    """An open Phantom dump: its structure (parsed at open) and column reads into host arrays or
    device tensors."""

    def __init__(self, path: str):
        L = lib()
        h = C.c_void_p()
        rc = L.pmx_phantom_open(C.byref(h), os.fsencode(path))
        if rc != PMX_OK:
            _raise(rc, L.pmx_phantom_open_error().decode())
        self._h = h
        v = np.zeros(5, np.int64)
        self._check(L.pmx_phantom_info(h, v.ctypes.data, 5))
        self.int_bytes, self.real_bytes, nblocks, ncols, nvars = (int(x) for x in v)
        buf = C.create_string_buffer(128)
        self._check(L.pmx_phantom_file_identifier(h, buf, 128))
        self.file_identifier = buf.value.decode(errors="replace")
        self.block_sizes = []
        for b in range(nblocks):
            n = C.c_int64()
            self._check(L.pmx_phantom_block_size(h, b, C.byref(n)))
            self.block_sizes.append(int(n.value))
        self.columns = []            # (name, block, dtype code) in file order
        for i in range(ncols):
            nm, blk, dt = C.create_string_buffer(17), C.c_int32(), C.c_int32()
            self._check(L.pmx_phantom_column_info(h, i, nm, C.byref(blk), C.byref(dt)))
            self.columns.append((nm.value.decode(errors="replace"), int(blk.value), int(dt.value)))
        self.header = {}             # name -> numpy scalar of the stored type
        for i in range(nvars):
            nm, dt, fv, iv = C.create_string_buffer(24), C.c_int32(), C.c_double(), C.c_int64()
            self._check(L.pmx_phantom_header_var(h, i, nm, C.byref(dt), C.byref(fv), C.byref(iv)))
            t = _NP_OF[int(dt.value)]
            val = t(fv.value) if int(dt.value) in (PMX_F32, PMX_F64) else t(iv.value)
            self.header[nm.value.decode(errors="replace")] = val

    def _check(self, rc):
        if rc != PMX_OK:
            _raise(rc, lib().pmx_phantom_last_error(self._h).decode())

    def close(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.pmx_phantom_close(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def read_column(self, col: int, out=None, dtype=None, keep=None):
        """Column `col` into `out` (a numpy array, or a device tensor for a straight-to-HBM read),
        or into a new host array of `dtype` (default: the stored type).  keep: increasing 0-based
        rows to keep (int64), or None for all rows."""
        name, blk, code = self.columns[col]
        n = self.block_sizes[blk] if keep is None else len(keep)
        if out is None:
            out = np.empty(n, _NP_OF[code] if dtype is None else dtype)
        if out.shape[0] != n:
            raise DimensionMismatch(PMX_ERR_DIMENSION, f"column {name}: {n} rows, destination {out.shape[0]}")
        mem = _mem_of([out])
        kp = None if keep is None else np.ascontiguousarray(keep, np.int64)
        self._check(lib().pmx_phantom_read_column(self._h, int(col), dtype_code_np(out.dtype), mem, _ptr(out),
                                                  None if kp is None else kp.ctypes.data,
                                                  0 if kp is None else int(kp.size)))
        return out


def density_from_h(h, m, hfact, dim=3, out=None):
    """add_rho!'s _density on one column (host array or device tensor; m, hfact in h's type)."""
    if out is None:
        out = h.new_empty(h.shape) if not isinstance(h, np.ndarray) else np.empty_like(h)
    mem = _mem_of([h, out])
    rc = lib().pmx_phantom_density(_dtype_code(h), int(dim), mem, int(h.shape[0]), _ptr(h), float(m), float(hfact),
                                   _ptr(out))
    if rc != PMX_OK:
        _raise(rc, "pmx_phantom_density")
    return out
