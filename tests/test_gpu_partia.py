"""The host mirror of Partia's interface driving the library on the GPU, end to end:
StructuredGrid_interpolation (cylindrical and spherical: host flatten, PointSamples_interpolation, reshape;
Cartesian: samples flattened on the device from the axes), PointSamples_interpolation and
LineSamples_interpolation, each against the oracle on copies of the same input (the calls permute
the caller's columns into Morton order, as LinearBVH! does).  The device-flattened Cartesian grid
must reproduce the host-flattened points exactly on the FP64 path."""
import numpy as np
import pytest

from distorteddustydiscanalysistoolkits_amd import disc
from distorteddustydiscanalysistoolkits_amd import partia as P

from .helpers import CAT_RHO_V, oracle_tree, rel_norms

pytestmark = pytest.mark.gpu
N = 20000
F32C_TOL = 1e-6
F64_TOL = 1e-7      # FP32 input columns give FP32 outputs (Partia keeps the input type): rounding only
NAMES = ("x", "y", "z", "m", "h", "rho")
QUANT = ("rho", "vx", "vy", "vz")


@pytest.fixture(scope="module")
def columns():
    d = disc.make_disc(N, seed=11, warped=True)
    cols = {k: np.ascontiguousarray(d[k], np.float32) for k in NAMES + QUANT[1:]}
    tree, _, _ = oracle_tree([cols[k] for k in NAMES], [cols[k] for k in QUANT])
    return cols, tree


def _input(cols):
    """Fresh copies: the interpolation calls permute their input columns in place."""
    return P.InterpolationInput(*(cols[k].copy() for k in NAMES), quant=tuple(cols[k].copy() for k in QUANT),
                                smoothed_kernel=P.M4_spline)


def _catalog():
    return P.InterpolationCatalog(QUANT, scalars=("rho",), gradients=("rho",), divergences=("v",), curls=("v",))


def _check(got, ref, tol, tag):
    for j, (g, r) in enumerate(zip(got, ref)):
        g = np.asarray(g, np.float64)
        assert np.array_equal(np.isnan(g), np.isnan(r)), f"{tag} field {j}: NaN pattern differs"
        assert max(rel_norms(g, r)) <= tol, f"{tag} field {j}: {rel_norms(g, r)}"


@pytest.mark.parametrize("precision", ["f32c", "f64"])
def test_structured_grid_cylindrical(columns, precision):
    cols, tree = columns
    grid = P.StructuredGrid.build(P.Cylindrical, (1.0, 100.0, 16), (0.0, 2 * np.pi, 16), (-5.0, 5.0, 8))
    backend = P.MI355XComputeBackend(precision=precision)
    cat = _catalog()
    bundle = P.StructuredGrid_interpolation(backend, P.Cylindrical, grid, _input(cols), cat, P.itpSymmetric)
    assert bundle.names == cat.ordered_names
    assert all(g.grid.shape == grid.size for g in bundle.grids)
    ref = tree.point_samples("M4", "symmetric", CAT_RHO_V, *disc.flatten_cylindrical(*grid.axes))
    _check([np.ravel(g.grid, order="F") for g in bundle.grids], ref, F32C_TOL if precision == "f32c" else F64_TOL,
           "cylindrical")


@pytest.mark.parametrize("precision", ["f32c", "f64"])
@pytest.mark.parametrize("shape", ["ball", "shell"])
def test_structured_grid_spherical(columns, precision, shape):
    """StructuredGrid(Spherical, ...) through the host flatten (transform.jl:68-71, _sph2cart),
    PointSamples_interpolation and the reshape (restore_struct), against the oracle."""
    cols, tree = columns
    r = (2.0, 90.0, 10) if shape == "ball" else 40.0
    grid = P.StructuredGrid.build(P.Spherical, r, (0.0, 2 * np.pi, 12), (0.2, np.pi - 0.2, 9))
    backend = P.MI355XComputeBackend(precision=precision)
    cat = _catalog()
    bundle = P.StructuredGrid_interpolation(backend, P.Spherical, grid, _input(cols), cat, P.itpSymmetric)
    assert bundle.names == cat.ordered_names
    assert all(g.grid.shape == grid.size for g in bundle.grids)
    ref = tree.point_samples("M4", "symmetric", CAT_RHO_V, *disc.flatten_spherical(*grid.axes))
    _check([np.ravel(g.grid, order="F") for g in bundle.grids], ref, F32C_TOL if precision == "f32c" else F64_TOL,
           "spherical")


@pytest.mark.parametrize("precision", ["f32c", "f64"])
@pytest.mark.parametrize("strategy", ["gather", "symmetric"])
def test_structured_grid_cartesian_device_flatten(columns, precision, strategy):
    cols, tree = columns
    grid = P.StructuredGrid.build(P.Cartesian, (-60.0, 60.0, 23), (-60.0, 60.0, 19), (-2.0, 2.0, 5))
    backend = P.MI355XComputeBackend(precision=precision)
    itp = {"gather": P.itpGather, "symmetric": P.itpSymmetric}[strategy]
    b_dev = P.StructuredGrid_interpolation(backend, P.Cartesian, grid, _input(cols), _catalog(), itp)
    flat = P.flatten(P.Cartesian, grid)
    b_host = P.PointSamples_interpolation(backend, flat, _input(cols), _catalog(), itp)
    for gd, gh in zip(b_dev.grids, b_host.grids):
        a, b = np.ravel(gd.grid, order="F"), gh.grid
        if precision == "f64":
            assert np.array_equal(a, b, equal_nan=True)
        else:
            assert np.array_equal(np.isnan(a), np.isnan(b))
            assert rel_norms(a, b)[0] <= F32C_TOL
    ref = tree.point_samples("M4", strategy, CAT_RHO_V, *flat.coor)
    _check([np.ravel(g.grid, order="F") for g in b_dev.grids], ref, F32C_TOL if precision == "f32c" else F64_TOL,
           "cartesian")


def test_line_samples(columns):
    cols, tree = columns
    o, d = disc.faceon_lines(40)
    ls = P.LineSamples(np.zeros(o[0].size), o, d)
    cat = P.InterpolationCatalog(QUANT, scalars=("rho", "vz"))
    bundle = P.LineSamples_interpolation(P.MI355XComputeBackend(precision="f32c"), ls, _input(cols), cat,
                                         P.itpScatter)
    ref = tree.line_samples("M4", "scatter", [0, 3], [0, 1], o, d)
    _check([g.grid for g in bundle.grids], ref, F32C_TOL, "lines")


def test_grid_bundle_to_hdf5_and_back(columns, tmp_path):
    """StructuredGrid_interpolation on the device, then write_GridBundle and read_GridDataset
    (write_grids.jl:181-215, read_grids.jl:29-86): the file holds the bundle bit for bit."""
    import os
    from distorteddustydiscanalysistoolkits_amd import griddataset as G
    if not os.path.exists(G.LIB):
        pytest.skip("libpartia_griddataset.so not built (no HDF5)")
    cols, _ = columns
    grid = P.StructuredGrid.build(P.Cylindrical, (1.0, 100.0, 16), (0.0, 2 * np.pi, 16), (-5.0, 5.0, 8))
    bundle = P.StructuredGrid_interpolation(P.MI355XComputeBackend(precision="f32c"), P.Cylindrical, grid,
                                            _input(cols), _catalog(), P.itpSymmetric)
    fn = str(tmp_path / "c2_small.h5")
    P.write_GridBundle(bundle, fn, code_units=dict(umass=1.0, udist=1.0, utime=1.0, umagfd=1.0),
                       operation_name="cylindrical")
    gd = P.read_GridDataset(fn)
    assert gd.data.names == bundle.names
    assert gd.params["grid_type"] == "StructuredGrid" and gd.params["dtype"] == "Float32"
    for g, r in zip(gd.data.grids, bundle.grids):
        assert g.size == grid.size
        assert np.array_equal(g.grid, r.grid, equal_nan=True)
