"""Error paths through the C-ABI on the device, with the reference's error classes and texts:
empty and oversized inputs (BinaryRadixTree.jl:85-90), a NaN coordinate (UInt64(floor(NaN)) in
_quantize_coords, MortonCode.jl:473-498), a non-scatter line strategy
(LineSamples_interpolation.jl:48-52), catalog slots outside the input's columns
(InterpolationCatalog.jl:141-188), an exception inside an entry point (PMX_TEST_THROW) -- each
leaving the context usable -- and the in-place Morton-order write-back of the caller's host and
device columns (LinearBVH!, InterpolationInput.jl:211-219)."""
import ctypes as C

import numpy as np
import pytest

from distorteddustydiscanalysistoolkits_amd import _lib, disc

from .helpers import disc_columns, oracle_tree, rel_norms

pytestmark = pytest.mark.gpu


def _particles(n, arrs, dtype=_lib.PMX_F64, ncol=0):
    qp = (C.c_void_p * 1)(None)
    return _lib.Particles(dtype, _lib.PMX_MEM_HOST, n, *[a.ctypes.data for a in arrs], ncol,
                          C.cast(qp, C.POINTER(C.c_void_p)), 0)


def _still_usable(ctx):
    cols, quant = disc_columns(2000, seed=9)
    ctx.build(*cols, quant=quant)
    qx, qy, qz = [np.ascontiguousarray(a) for a in (cols[0][:50], cols[1][:50], cols[2][:50])]
    outs = [np.empty(50)]
    ctx.interpolate_points(_lib.PMX_SYMMETRIC, _lib.PMX_PREC_F32C, _lib.catalog_struct(scalar_slots=[0]), qx, qy, qz,
                           outs)
    assert np.isfinite(outs[0]).all()


def test_empty_and_oversized_inputs(ctx):
    a = [np.zeros(4) for _ in range(6)]
    for n, text in ((0, b"non-empty"), ((1 << 30) + 1, b"Int32 node-id capacity")):
        p = _particles(n, a)
        rc = _lib.lib().pmx_build_lbvh(ctx._h, C.byref(p), 0, None, None)
        assert rc == _lib.PMX_ERR_ARGUMENT, n
        assert text in _lib.lib().pmx_last_error(ctx._h)
    _still_usable(ctx)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_nan_coordinate(ctx, dtype):
    cols, quant = disc_columns(3000, seed=4, dtype=dtype)
    cols[1][1234] = np.nan
    with pytest.raises(_lib.ArgumentError) as e:
        ctx.build(*cols, quant=quant)
    assert "NaN" in str(e.value)
    _still_usable(ctx)


def test_line_strategy_and_catalog_slots(ctx):
    cols, quant = disc_columns(3000, seed=6)
    ctx.build(*cols, quant=quant)
    o, d = disc.faceon_lines(8)
    outs = [np.empty(64)]
    for s in (_lib.PMX_GATHER, _lib.PMX_SYMMETRIC):
        with pytest.raises(_lib.ArgumentError) as e:
            ctx.interpolate_lines(s, _lib.PMX_PREC_F32C, _lib.catalog_struct(scalar_slots=[0]), o, d, outs)
        assert "only supports itpScatter" in str(e.value)
    with pytest.raises(_lib.DimensionMismatch):
        ctx.interpolate_lines(_lib.PMX_SCATTER, _lib.PMX_PREC_F32C, _lib.catalog_struct(scalar_slots=[4]), o, d, outs)
    with pytest.raises(_lib.ArgumentError):       # line maps take scalars only
        ctx.interpolate_lines(_lib.PMX_SCATTER, _lib.PMX_PREC_F64, _lib.catalog_struct(grad_slots=[0]), o, d,
                              [np.empty(64)] * 3)
    qx, qy, qz = (np.zeros(5), np.zeros(5), np.zeros(5))
    bad = [dict(scalar_slots=[7]), dict(grad_slots=[4]), dict(div_slots=[1, 2, 9]), dict(curl_slots=[-1, 1, 2])]
    for cat in bad:
        L = len(cat.get("scalar_slots", [])) + 3 * len(cat.get("grad_slots", [])) + len(cat.get("div_slots", [])) // 3 \
            + len(cat.get("curl_slots", []))
        for prec in (_lib.PMX_PREC_F64, _lib.PMX_PREC_F32C):
            with pytest.raises(_lib.DimensionMismatch):
                ctx.interpolate_points(_lib.PMX_SYMMETRIC, prec, _lib.catalog_struct(**cat), qx, qy, qz,
                                       [np.empty(5) for _ in range(L)])
    with pytest.raises(_lib.ArgumentError):
        ctx.interpolate_points(3, _lib.PMX_PREC_F64, _lib.catalog_struct(scalar_slots=[0]), qx, qy, qz, [np.empty(5)])
    _still_usable(ctx)


@pytest.mark.parametrize("fault, code", [("bad_alloc", _lib.PMX_ERR_NOMEM), ("exception", _lib.PMX_ERR_INTERNAL)])
def test_exception_in_entry_point(ctx, monkeypatch, fault, code):
    cols, quant = disc_columns(2000, seed=2)
    monkeypatch.setenv("PMX_TEST_THROW", fault)
    with pytest.raises(_lib.PmxError) as e:
        ctx.build(*cols, quant=quant, writeback=True)
    assert e.value.code == code
    monkeypatch.delenv("PMX_TEST_THROW")
    _still_usable(ctx)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_writeback_permutes_host_columns(ctx, dtype):
    cols, quant = disc_columns(5000, seed=8, warped=True, dtype=dtype)
    tree, ocols, oquant = oracle_tree(cols, quant)
    mine = [c.copy() for c in cols]
    mq = [q.copy() for q in quant]
    ctx.build(*mine, quant=mq, writeback=True)
    order = tree.order - 1
    for got, orig, ref in zip(mine + mq, cols + quant, ocols + oquant):
        assert got.dtype == orig.dtype
        assert np.array_equal(got, orig[order])                 # the permutation of LinearBVH!
        assert np.array_equal(got.astype(np.float64), ref)      # the oracle's Morton-ordered columns
    # the permuted columns describe the same particles: a second build on them is the identity
    _, order2 = ctx.build(*[c.copy() for c in mine], quant=[q.copy() for q in mq], want_order=True)
    assert np.array_equal(order2, np.arange(order2.size))


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_writeback_permutes_device_columns(ctx, dtype):
    import torch
    cols, quant = disc_columns(7000, seed=11, warped=True, dtype=dtype)
    tree, ocols, oquant = oracle_tree(cols, quant)
    dev = [torch.from_numpy(c.copy()).to("cuda:0") for c in cols]
    dq = [torch.from_numpy(q.copy()).to("cuda:0") for q in quant]
    ctx.build(*dev, quant=dq, writeback=True)
    order = tree.order - 1
    for got, orig, ref in zip(dev + dq, cols + quant, ocols + oquant):
        g = got.cpu().numpy()
        assert g.dtype == orig.dtype
        assert np.array_equal(g, orig[order])
        assert np.array_equal(g.astype(np.float64), ref)
    # without the flag the caller's device columns are left as they are
    dev2 = [torch.from_numpy(c.copy()).to("cuda:0") for c in cols]
    ctx.build(*dev2, quant=[torch.from_numpy(q.copy()).to("cuda:0") for q in quant])
    for got, orig in zip(dev2, cols):
        assert np.array_equal(got.cpu().numpy(), orig)
