"""The Phantom dump reader straight into device memory, and the whole input side on the GPU:
read_phantom(device) -> add_rho -> build_input -> PointSamples / LineSamples interpolation.

Device reads must equal host reads bit for bit (same conversion, same kept rows, chunk
boundaries crossed); the device add_rho must equal the host one bit for bit (IEEE division and
products in the mass type); the interpolation of device-resident inputs must equal that of the
host-resident copy (FP64 path: bit for bit) and the oracle (FP64 rounding of FP32 inputs)."""
import numpy as np
import pytest
import torch

from distorteddustydiscanalysistoolkits_amd import _lib, disc
from distorteddustydiscanalysistoolkits_amd import particleio as PIO
from distorteddustydiscanalysistoolkits_amd import phantomdump as W
from distorteddustydiscanalysistoolkits_amd import partia as P

from .helpers import oracle_tree, rel_norms

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def test_device_reads_equal_host_reads(tmp_path):
    n = (1 << 22) + 777
    rng = np.random.default_rng(8)
    a = rng.standard_normal(n).astype(np.float32)
    b = rng.integers(-100, 100, n).astype(np.int8)
    c = rng.standard_normal(n)
    path = W.write_dump(str(tmp_path / "big"), header={},
                        blocks=[(n, [(W.I8, "b", b), (W.DEF_REAL, "a", a), (W.R8, "c", c)])])
    keep = np.unique(np.concatenate([rng.choice(n, 200000, replace=False), [0, (1 << 22) - 1, 1 << 22, n - 1]]))
    keep = keep.astype(np.int64)
    with _lib.Phantom(path) as ph:
        for col, dt in ((0, torch.int8), (0, torch.float32), (1, torch.float32), (1, torch.float64), (2, torch.float64),
                        (2, torch.float32)):
            host_dt = np.dtype(str(dt).replace("torch.", ""))
            for kp in (None, keep):
                m = n if kp is None else kp.size
                out = torch.empty(m, dtype=dt, device=DEV)
                ph.read_column(col, out=out, keep=kp)
                ref = ph.read_column(col, dtype=host_dt, keep=kp)
                assert np.array_equal(out.cpu().numpy(), ref), (col, dt, kp is None)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("dim", [2, 3])
def test_density_device_equals_host(dtype, dim):
    rng = np.random.default_rng(9)
    npdt = np.float32 if dtype == torch.float32 else np.float64
    h = rng.uniform(0.01, 5.0, 300001).astype(npdt)
    m, hfact = npdt(2.5e-7), npdt(1.2)
    ref = _lib.density_from_h(h, m, hfact, dim=dim)
    got = _lib.density_from_h(torch.from_numpy(h).to(DEV), m, hfact, dim=dim)
    torch.cuda.synchronize()
    assert np.array_equal(got.cpu().numpy(), ref)


@pytest.fixture(scope="module")
def disc_dump(tmp_path_factory):
    d = disc.make_disc(30000, seed=12, warped=True)
    path = W.disc_dump(str(tmp_path_factory.mktemp("phantom") / "disc_00000"), d)
    return path


def _prepare(path, device):
    (gas,) = PIO.read_phantom(path, device=device)
    gas.params["mass"] = gas.params["massoftype"]
    PIO.add_rho(gas)
    return PIO.build_input(gas, PIO.MassFromParams("massoftype"), scalars=("vz",), divergences=("v",),
                           curls=("v",), smoothed_kernel=P.M4_spline)


def _host_copy(inp):
    cols = [inp.coord[0], inp.coord[1], inp.coord[2], inp.m, inp.h, inp.ρ]
    return (tuple(np.ascontiguousarray(c.cpu().numpy()) for c in cols),
            tuple(np.ascontiguousarray(q.cpu().numpy()) for q in inp.quant))


@pytest.mark.parametrize("precision", ["f64", "f32c"])
def test_point_samples_from_device_resident_dump(disc_dump, precision):
    inp_d, cat = _prepare(disc_dump, DEV)
    assert inp_d.coord[0].is_cuda and str(inp_d.dtype) == "torch.float32"
    (cols, quant) = _host_copy(inp_d)
    inp_h, cat_h = _prepare(disc_dump, None)
    # the host pipeline yields the same columns (the device one is not permuted by the build)
    for a, b in zip(cols, [inp_h.coord[0], inp_h.coord[1], inp_h.coord[2], inp_h.m, inp_h.h, inp_h.ρ]):
        assert np.array_equal(a, b)
    grid = P.StructuredGrid.build(P.Cylindrical, (1.0, 100.0, 12), (0.0, 2 * np.pi, 12), (-4.0, 4.0, 6))
    flat = P.flatten(P.Cylindrical, grid)
    backend = P.MI355XComputeBackend(precision=precision)
    b_dev = P.PointSamples_interpolation(backend, flat, inp_d, cat, P.itpSymmetric)
    b_host = P.PointSamples_interpolation(backend, flat, inp_h, cat_h, P.itpSymmetric)
    assert b_dev.names == b_host.names
    for gd, gh in zip(b_dev.grids, b_host.grids):
        if precision == "f64":
            assert np.array_equal(gd.grid, gh.grid, equal_nan=True)
        else:
            assert np.array_equal(np.isnan(gd.grid), np.isnan(gh.grid))
            assert rel_norms(gd.grid.astype(np.float64), gh.grid.astype(np.float64))[0] <= 1e-6
    # against the oracle on the file-order columns
    tree, _, _ = oracle_tree(list(cols), list(quant))
    cat_o = dict(scalar_slots=[0], scalar_shepard=[1], grad_slots=[], div_slots=[1, 2, 0], curl_slots=[1, 2, 0])
    ref = tree.point_samples("M4", "symmetric", cat_o, *flat.coor)
    tol = 1e-7 if precision == "f64" else 1e-6
    for g, r in zip(b_dev.grids, ref):
        g = np.asarray(g.grid, np.float64)
        assert np.array_equal(np.isnan(g), np.isnan(r))
        assert max(rel_norms(g, r)) <= tol


@pytest.mark.parametrize("separate", ["sinks", "all"])
def test_read_phantom_device_equals_host_with_inactive(tmp_path, separate):
    rng = np.random.default_rng(21)
    n = 50000
    g = {k: rng.standard_normal(n).astype(np.float32) for k in ("x", "y", "z", "vx")}
    g["h"] = rng.uniform(0.1, 1.0, n).astype(np.float32)
    g["h"][rng.choice(n, 500, replace=False)] = -1.0
    g["h"][rng.choice(n, 50, replace=False)] = np.nan
    itype = np.where(rng.random(n) < 0.3, 2, 1).astype(np.int8)
    arrays = [(W.I8, "itype", itype)] + [(W.DEF_REAL, k, g[k]) for k in ("x", "y", "z", "h")] + [(W.R4, "vx", g["vx"])]
    sinks = [(W.R8, k, np.arange(3.0) + j) for j, k in enumerate(("x", "y", "z", "m", "h"))]
    header = {W.DEF_REAL: [("hfact", 1.2), ("massoftype", 1e-6), ("massoftype_2", 2e-6)]}
    path = W.write_dump(str(tmp_path / "mixed"), header=header, blocks=[(n, arrays), (3, sinks)])
    fd = PIO.read_phantom(path, separate_types=separate, device=DEV)
    fh = PIO.read_phantom(path, separate_types=separate)
    assert len(fd) == len(fh) == (3 if separate == "all" else 2)
    for a, b in zip(fd, fh):
        assert a.names() == b.names()
        for k in a.names():
            assert a[k].is_cuda and np.array_equal(a[k].cpu().numpy(), b[k], equal_nan=True), k
    assert PIO.get_npart(fh[0]) + (PIO.get_npart(fh[1]) if separate == "all" else 0) == int((g["h"] > 0).sum())
