"""Full-size parity of the column-map workload (C3) and of the point workload (C2, below): a warped disc of 1.6e7 particles stored in FP32
(r ~ U[1, 100], H = 0.05 r^1.25, tilt beta = 0.2 r / 100 about gamma = 0.5 ln r), Wendland C4
kernel, face-on 1024 x 1024 map of the column density (rho, not Shepard-normalised) and the
column-averaged vz, with the reference's Inorm / h weight (line_integrated_kernel.jl:132-137,
line_integrated_scalar_interpolation.jl:123-166, kernels/C4_Wendland.jl:9-32).

The device computes the whole map on both paths; the oracle evaluates an evenly strided subset of
4096 lines (every 256th).  Bars on the subset: per-line neighbour digests of the tiled path
(count, index sum, index-square sum) equal to the oracle's accepted sets; FP64 fields <= 1e-13
and FP32c fields <= 1e-6 normwise relative L-inf and L2 per field.  Size-independent checks on
the whole map: both paths give the same NaN pattern (empty columns) and agree to 1e-6."""
import time

import numpy as np
import pytest

from distorteddustydiscanalysistoolkits_amd import _lib, disc

from .helpers import oracle_tree, rel_norms

pytestmark = pytest.mark.gpu
N = 16_000_000
MAP = 1024
STRIDE = 256
F32C_TOL = 1e-6
F64_TOL = 1e-13


@pytest.fixture(scope="module")
def c3():
    t0 = time.time()
    d = disc.make_disc(N, seed=0, warped=True)
    cols = [np.ascontiguousarray(d[k], np.float32) for k in ("x", "y", "z", "m", "h", "rho")]
    quant = [np.ascontiguousarray(d[k], np.float32) for k in ("rho", "vz")]
    del d
    print(f"[c3] disc of {N} particles drawn in {time.time() - t0:.1f}s", flush=True)
    o, dv = disc.faceon_lines(MAP)
    ctx = _lib.Context(0)
    print(f"[c3] device build n={N} kernel=C4", flush=True)
    ctx.build(*cols, quant=quant, kernel=4)
    cat = _lib.catalog_struct(scalar_slots=[0, 1], scalar_shepard=[0, 1])
    maps = {}
    for name, prec in (("f32c", _lib.PMX_PREC_F32C), ("f64", _lib.PMX_PREC_F64)):
        outs = [np.empty(MAP * MAP) for _ in range(2)]
        print(f"[c3] device {name} map {MAP}x{MAP}", flush=True)
        ctx.interpolate_lines(_lib.PMX_SCATTER, prec, cat, o, dv, outs)
        maps[name] = outs
    print("[c3] device digest of the tiled path", flush=True)
    digest = ctx.line_digest(o, dv)
    ctx.close()
    t0 = time.time()
    tree, _, _ = oracle_tree(cols, quant)
    sub = np.arange(0, MAP * MAP, STRIDE)
    so = [a[sub].copy() for a in o]
    sd = [a[sub].copy() for a in dv]
    ref = tree.line_samples("C4", "scatter", [0, 1], [0, 1], so, sd)
    roff, ridx = tree.line_neighbors("C4", "scatter", so, sd)
    print(f"[c3] oracle build + {sub.size} lines in {time.time() - t0:.1f}s, "
          f"{roff[-1]} accepted pairs on the subset", flush=True)
    return dict(maps=maps, digest=digest, sub=sub, ref=ref, roff=roff, ridx=ridx)


def test_c3_digest_subset_bitexact(c3):
    dg = c3["digest"]
    assert dg is not None, "the face-on map must take the tiled path"
    roff, ridx = c3["roff"], c3["ridx"]
    cnt = np.diff(roff).astype(np.uint64)
    seg = np.repeat(np.arange(cnt.size), np.diff(roff))
    ii = ridx.astype(np.uint64)
    s1 = np.zeros(cnt.size, np.uint64)
    s2 = np.zeros(cnt.size, np.uint64)
    np.add.at(s1, seg, ii)
    np.add.at(s2, seg, ii * ii)
    sub = c3["sub"]
    for g, r in zip(dg, (cnt, s1, s2)):
        assert np.array_equal(g[sub], r)
    assert cnt.sum() > 1_000_000


@pytest.mark.parametrize("path, tol", [("f32c", F32C_TOL), ("f64", F64_TOL)])
def test_c3_fields_subset(c3, path, tol):
    for j, (g, r) in enumerate(zip(c3["maps"][path], c3["ref"])):
        g = g[c3["sub"]]
        assert np.array_equal(np.isnan(g), np.isnan(r)), f"field {j}: NaN pattern differs"
        linf, l2 = rel_norms(g, r)
        print(f"[c3 {path}] field {j}: rel Linf {linf:.3e} L2 {l2:.3e}", flush=True)
        assert linf <= tol and l2 <= tol


def test_c3_whole_map_paths_agree(c3):
    for j, (a, b) in enumerate(zip(c3["maps"]["f32c"], c3["maps"]["f64"])):
        assert np.array_equal(np.isnan(a), np.isnan(b))
        linf, l2 = rel_norms(a, b)
        print(f"[c3 whole map] field {j}: f32c vs f64 rel Linf {linf:.3e} L2 {l2:.3e}", flush=True)
        assert linf <= F32C_TOL and l2 <= F32C_TOL


# ---- points: C2 at full size ---------------------------------------------------------------
# A flat disc of 4e6 particles stored in FP32, M4 kernel, the cylindrical 256 x 256 x 64 grid
# (s in [1, 100], phi in [0, 2 pi), z in [-5, 5]) with rho, grad rho, div v, curl v, itpSymmetric
# (general_interpolation.jl:446-669, LinearBVHQueries.jl:118-141 for the nearest-h).  The device
# computes the whole grid on the FP32c path; the oracle evaluates every 1024th sample (4096
# samples).  Bars on the subset: neighbour sets of both device paths bit-equal to the oracle's,
# FP32c fields <= 1e-6 and FP64 fields <= 1e-13 normwise relative L-inf and L2 per field.
N2 = 4_000_000
GRID2 = (256, 256, 64)
STRIDE2 = 1024


@pytest.fixture(scope="module")
def c2():
    from .helpers import CAT_RHO_V
    t0 = time.time()
    d = disc.make_disc(N2, seed=0, warped=False)
    cols = [np.ascontiguousarray(d[k], np.float32) for k in ("x", "y", "z", "m", "h", "rho")]
    quant = [np.ascontiguousarray(d[k], np.float32) for k in ("rho", "vx", "vy", "vz")]
    del d
    sa, pa, za = disc.cylindrical_axes((1.0, 100.0, GRID2[0]), (0.0, 2 * np.pi, GRID2[1]), (-5.0, 5.0, GRID2[2]))
    q = disc.flatten_cylindrical(sa, pa, za)
    sub = np.arange(0, q[0].size, STRIDE2)
    qs = [np.ascontiguousarray(a[sub]) for a in q]
    print(f"[c2] disc of {N2} particles and {q[0].size} samples in {time.time() - t0:.1f}s", flush=True)
    ctx = _lib.Context(0)
    print(f"[c2] device build n={N2} kernel=M4", flush=True)
    ctx.build(*cols, quant=quant, kernel=0)
    cat = _lib.catalog_struct(**CAT_RHO_V)
    print(f"[c2] device f32c grid of {q[0].size} samples", flush=True)
    full = [np.empty(q[0].size) for _ in range(8)]
    ctx.interpolate_points(_lib.PMX_SYMMETRIC, _lib.PMX_PREC_F32C, cat, *q, full)
    print("[c2] device f64 fields and both paths' neighbour sets on the subset", flush=True)
    f64 = [np.empty(sub.size) for _ in range(8)]
    ctx.interpolate_points(_lib.PMX_SYMMETRIC, _lib.PMX_PREC_F64, cat, *qs, f64)
    nb = {name: ctx.point_neighbors(_lib.PMX_SYMMETRIC, *qs, precision=prec)
          for name, prec in (("f32c", _lib.PMX_PREC_F32C), ("f64", _lib.PMX_PREC_F64))}
    ctx.close()
    t0 = time.time()
    tree, _, _ = oracle_tree(cols, quant)
    ref = tree.point_samples("M4", "symmetric", CAT_RHO_V, *qs)
    roff, ridx = tree.point_neighbors("M4", "symmetric", *qs)
    print(f"[c2] oracle build + {sub.size} samples in {time.time() - t0:.1f}s, {roff[-1]} accepted pairs",
          flush=True)
    return dict(fields={"f32c": [a[sub] for a in full], "f64": f64}, nb=nb, ref=ref, roff=roff, ridx=ridx)


@pytest.mark.parametrize("path", ["f32c", "f64"])
def test_c2_neighbour_sets_subset_bitexact(c2, path):
    off, idx = c2["nb"][path]
    assert np.array_equal(off, c2["roff"]) and np.array_equal(idx, c2["ridx"])
    assert c2["roff"][-1] > 500_000


@pytest.mark.parametrize("path, tol", [("f32c", F32C_TOL), ("f64", F64_TOL)])
def test_c2_fields_subset(c2, path, tol):
    for j, (g, r) in enumerate(zip(c2["fields"][path], c2["ref"])):
        assert np.array_equal(np.isnan(g), np.isnan(r)), f"field {j}: NaN pattern differs"
        linf, l2 = rel_norms(g, r)
        print(f"[c2 {path}] field {j}: rel Linf {linf:.3e} L2 {l2:.3e}", flush=True)
        assert linf <= tol and l2 <= tol
