"""Every SPH kernel (M4, M5, M6, C2, C4, C6; kernels/*.jl) through both interpolation paths
against the oracle on a 20 000-particle warped disc stored in FP32.

The FP32-compensated point path evaluates rewritten shape forms (f'(q)/q without the
reference's cancelling differences) and its own FP32 weights; the line paths look up each
kernel's own Inorm table (line_integrated_kernel_tables.jl) with the reference's 1/h weight.
Bars: neighbour sets and tiled-line digests bit-exact; plain FP64 fields <= 1e-13 and FP32c
fields <= 1e-6 normwise relative L-inf and L2 per field."""
import numpy as np
import pytest

from distorteddustydiscanalysistoolkits_amd import _lib, disc
from oracle import oracle as O

from .helpers import CAT_RHO_V, cyl_points, disc_columns, oracle_tree, rel_norms

pytestmark = pytest.mark.gpu
N = 20000
F32C_TOL = 1e-6
F64_TOL = 1e-13
KNAMES = ["M4", "M5", "M6", "C2", "C4", "C6"]
LINE_CAT = dict(slots=[0, 3], shep=[0, 1])


@pytest.fixture(scope="module", params=range(6), ids=KNAMES)
def kbuilt(request):
    k = request.param
    cols, quant = disc_columns(N, seed=13, warped=True, dtype=np.float32)
    ctx = _lib.Context(0)
    print(f"[kernels] build n={N} kernel={KNAMES[k]}", flush=True)
    ctx.build(*cols, quant=quant, kernel=k)
    tree, _, _ = oracle_tree(cols, quant)
    yield dict(ctx=ctx, tree=tree, k=k, name=KNAMES[k])
    ctx.close()


def _check(outs, ref, tol, what):
    worst = []
    for j, (g, r) in enumerate(zip(outs, ref)):
        assert np.array_equal(np.isnan(g), np.isnan(r)), f"{what} field {j}: NaN pattern differs"
        worst.append(rel_norms(g, r))
    print(f"[{what}] rel (Linf, L2) per field: {[(float('%.2e' % a), float('%.2e' % b)) for a, b in worst]}",
          flush=True)
    for j, (linf, l2) in enumerate(worst):
        assert linf <= tol and l2 <= tol, f"{what} field {j}: {linf} {l2}"


@pytest.mark.parametrize("strategy", ["gather", "scatter", "symmetric"])
@pytest.mark.parametrize("prec", [_lib.PMX_PREC_F64, _lib.PMX_PREC_F32C], ids=["f64", "f32c"])
def test_point_fields(kbuilt, strategy, prec):
    qx, qy, qz = cyl_points(18, 18, 10)
    outs = [np.empty(qx.size) for _ in range(8)]
    kbuilt["ctx"].interpolate_points(O.STRATEGIES[strategy], prec, _lib.catalog_struct(**CAT_RHO_V), qx, qy, qz, outs)
    ref = kbuilt["tree"].point_samples(kbuilt["name"], strategy, CAT_RHO_V, qx, qy, qz)
    _check(outs, ref, F64_TOL if prec == _lib.PMX_PREC_F64 else F32C_TOL,
           f"{kbuilt['name']} points {strategy} prec={prec}")


@pytest.mark.parametrize("strategy", ["gather", "symmetric"])
def test_point_neighbor_sets(kbuilt, strategy):
    """The radius rule scales with each kernel's support (2, 2.5, 3): sets bit-exact on both paths."""
    qx, qy, qz = cyl_points(12, 12, 6)
    roff, ridx = kbuilt["tree"].point_neighbors(kbuilt["name"], strategy, qx, qy, qz)
    for prec in (_lib.PMX_PREC_F64, _lib.PMX_PREC_F32C):
        off, idx = kbuilt["ctx"].point_neighbors(O.STRATEGIES[strategy], qx, qy, qz, precision=prec)
        assert np.array_equal(off, roff) and np.array_equal(idx, ridx), (strategy, prec)
    assert roff[-1] > 0


@pytest.mark.parametrize("geometry", ["faceon", "nonparallel"])
@pytest.mark.parametrize("prec", [_lib.PMX_PREC_F64, _lib.PMX_PREC_F32C], ids=["f64", "f32c"])
def test_line_fields(kbuilt, geometry, prec):
    if geometry == "faceon":
        o, d = disc.faceon_lines(32)
    else:
        rng = np.random.default_rng(17)
        m = 24 * 24
        o = [rng.uniform(-60.0, 60.0, m), rng.uniform(-60.0, 60.0, m), rng.uniform(-3.0, 3.0, m)]
        tilt = rng.uniform(-0.2, 0.2, (2, m))
        nrm = np.sqrt(1.0 + tilt[0] ** 2 + tilt[1] ** 2)
        d = [tilt[0] / nrm, tilt[1] / nrm, 1.0 / nrm]
    outs = [np.empty(o[0].size) for _ in range(2)]
    kbuilt["ctx"].interpolate_lines(_lib.PMX_SCATTER, prec,
                                    _lib.catalog_struct(scalar_slots=LINE_CAT["slots"], scalar_shepard=LINE_CAT["shep"]),
                                    o, d, outs)
    ref = kbuilt["tree"].line_samples(kbuilt["name"], "scatter", LINE_CAT["slots"], LINE_CAT["shep"], o, d)
    _check(outs, ref, F64_TOL if prec == _lib.PMX_PREC_F64 else F32C_TOL,
           f"{kbuilt['name']} lines {geometry} prec={prec}")


def test_tiled_line_digest(kbuilt):
    o, d = disc.faceon_lines(40)
    dg = kbuilt["ctx"].line_digest(o, d)
    assert dg is not None
    roff, ridx = kbuilt["tree"].line_neighbors(kbuilt["name"], "scatter", o, d)
    cnt = np.diff(roff).astype(np.uint64)
    seg = np.repeat(np.arange(cnt.size), np.diff(roff))
    s1 = np.zeros(cnt.size, np.uint64)
    s2 = np.zeros(cnt.size, np.uint64)
    ii = ridx.astype(np.uint64)
    np.add.at(s1, seg, ii)
    np.add.at(s2, seg, ii * ii)
    for g, r in zip(dg, (cnt, s1, s2)):
        assert np.array_equal(g, r)
    assert cnt.sum() > 0


@pytest.mark.parametrize("strategy", ["scatter", "symmetric"])
def test_dust_species_fields(kbuilt, strategy):
    """The thin dust layer (H_d = 0.3 H): rho, div v and curl v of the FP32c path with this kernel
    (nearly solenoidal velocity field, divergence far below its pair terms)."""
    d = disc.make_dust(N, seed=5, warped=True)
    cols = [np.ascontiguousarray(d[k], np.float32) for k in ("x", "y", "z", "m", "h", "rho")]
    quant = [np.ascontiguousarray(d[k], np.float32) for k in ("rho", "vx", "vy", "vz")]
    cat_d = dict(scalar_slots=[0], scalar_shepard=[0], div_slots=[1, 2, 3], curl_slots=[1, 2, 3])
    qx, qy, qz = cyl_points(20, 20, 10, zr=2.0)
    outs = [np.empty(qx.size) for _ in range(5)]
    c2 = _lib.Context(0)
    try:
        c2.build(*cols, quant=quant, kernel=kbuilt["k"])
        c2.interpolate_points(O.STRATEGIES[strategy], _lib.PMX_PREC_F32C, _lib.catalog_struct(**cat_d), qx, qy, qz,
                              outs)
    finally:
        c2.close()
    tree, _, _ = oracle_tree(cols, quant)
    ref = tree.point_samples(kbuilt["name"], strategy, cat_d, qx, qy, qz)
    _check(outs, ref, F32C_TOL, f"{kbuilt['name']} dust {strategy}")
