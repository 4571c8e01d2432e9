"""PMX_LINE_INV_H2: the line-integrated weight Inorm(q) / h^2 instead of the reference's
Inorm(q) / h (an extension, off by default).

The reference has no such option, so the flag is pinned by what the mathematics fixes: with every
particle at one smoothing length h0, the extra factor is the constant 1/h0, so an unnormalised
column equals the default one divided by h0 and a Shepard-normalised column is unchanged.  That
holds on every line path (plain FP64, FP32c tiled for parallel lines, FP32c union walk for
non-parallel lines).  With varying h, the FP32c result agrees with the FP64 one to the FP32c
tolerance, and the default path is untouched by the flag's existence (bit-equal to a call
without it)."""
import numpy as np
import pytest

from distorteddustydiscanalysistoolkits_amd import _lib, disc

from .helpers import disc_columns, rel_norms

pytestmark = pytest.mark.gpu
N = 20000
H0 = 1.75


def _lines(kind):
    if kind == "parallel":
        return disc.faceon_lines(32)
    o, d = disc.faceon_lines(24)
    rng = np.random.default_rng(5)
    tilt = rng.uniform(-0.2, 0.2, (2, o[0].size))
    nrm = np.sqrt(1.0 + tilt[0] ** 2 + tilt[1] ** 2)
    return o, [tilt[0] / nrm, tilt[1] / nrm, 1.0 / nrm]


def _run(ctx, prec, o, d):
    cat = _lib.catalog_struct(scalar_slots=[0, 3], scalar_shepard=[0, 1])    # rho column, vz Shepard
    outs = [np.empty(o[0].size) for _ in range(2)]
    ctx.interpolate_lines(_lib.PMX_SCATTER, prec, cat, o, d, outs)
    return outs


@pytest.mark.parametrize("kind", ["parallel", "nonparallel"])
@pytest.mark.parametrize("prec", [_lib.PMX_PREC_F64, _lib.PMX_PREC_F32C], ids=["f64", "f32c"])
def test_uniform_h_scales_by_one_over_h(ctx, kind, prec):
    cols, quant = disc_columns(N, seed=8, warped=True, dtype=np.float64)
    cols[4] = np.full(N, H0)                                  # h
    ctx.build(*cols, quant=quant, kernel=0)
    o, d = _lines(kind)
    base = _run(ctx, prec, o, d)
    h2 = _run(ctx, prec | _lib.PMX_LINE_INV_H2, o, d)
    tol = 1e-13 if prec == _lib.PMX_PREC_F64 else 1e-6
    assert np.array_equal(np.isnan(base[0]), np.isnan(h2[0]))
    assert np.isfinite(base[0]).sum() > 100
    assert rel_norms(h2[0], base[0] / H0)[0] <= tol            # unnormalised column: / h0
    assert rel_norms(h2[1], base[1])[0] <= tol                 # Shepard column: unchanged


@pytest.mark.parametrize("kind", ["parallel", "nonparallel"])
def test_f32c_matches_f64_with_flag(ctx, kind):
    cols, quant = disc_columns(N, seed=9, warped=True, dtype=np.float32)
    ctx.build(*cols, quant=quant, kernel=4)
    o, d = _lines(kind)
    ref = _run(ctx, _lib.PMX_PREC_F64 | _lib.PMX_LINE_INV_H2, o, d)
    got = _run(ctx, _lib.PMX_PREC_F32C | _lib.PMX_LINE_INV_H2, o, d)
    plain = _run(ctx, _lib.PMX_PREC_F64, o, d)
    for g, r, p in zip(got, ref, plain):
        assert np.array_equal(np.isnan(g), np.isnan(r))
        assert max(rel_norms(g, r)) <= 1e-6
    # h varies, so the flag changes the unnormalised column by more than rounding
    assert rel_norms(ref[0], plain[0])[0] > 1e-3
