"""Full-size parity of the column-map workload: a warped disc of 1.6e7 particles stored in FP32
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
