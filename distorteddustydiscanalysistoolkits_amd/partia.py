"""Host-side mirror of Partia.jl's interpolation interface over libpartia_mi355x.

Same names, argument meaning and error behaviour as the reference, so code (and tests)
written against Partia's API read the same here:

  InterpolationInput          interpolation_setup/InterpolationInput.jl:43-118
  InterpolationCatalog        interpolation_setup/InterpolationCatalog.jl:71-225
  itpGather/itpScatter/itpSymmetric   interpolation_setup/InterpolationStrategy.jl:1-5
  M4_spline ... C6_Wendland   kernel_function/kernels/*.jl
  PointSamples, LineSamples, StructuredGrid, GridBundle   P/src/julia/Grids/
  PointSamples_interpolation  grid_interpolation/PointSamples_interpolation.jl:34
                              (GPU method: ext/CUDAExt/.../PointSamples_interpolation.jl:12)
  LineSamples_interpolation   grid_interpolation/LineSamples_interpolation.jl:47
  StructuredGrid_interpolation  grid_interpolation/StructuredGrid_interpolation.jl:46
  GridDataset, write_GridDataset, write_GridBundle, read_GridDataset
                              Grids/griddataset/GridDataset.jl, IO/gridsIO/write_grids.jl,
                              IO/gridsIO/read_grids.jl (HDF5, libpartia_griddataset.so)

The compute runs only in the HIP library; a missing library raises.  As in the reference,
interpolation permutes the caller's particle columns into Morton order in place
(InterpolationInput.jl:211-219) and returns a GridBundle whose grids share geometry.
"""
from __future__ import annotations

import numpy as np

from . import _lib
from ._lib import ArgumentError, DimensionMismatch  # noqa: F401  (re-exported error classes)

# --------------------------------------------------------------------------- kernels ----


class AbstractSPHKernel:
    id = -1
    valid = 2.0
    nneigh = 0


class M4_spline(AbstractSPHKernel):
    id, valid, nneigh = 0, 2.0, 57


class M5_spline(AbstractSPHKernel):
    id, valid, nneigh = 1, 2.5, 113


class M6_spline(AbstractSPHKernel):
    id, valid, nneigh = 2, 3.0, 112


class C2_Wendland(AbstractSPHKernel):
    id, valid, nneigh = 3, 2.0, 92


class C4_Wendland(AbstractSPHKernel):
    id, valid, nneigh = 4, 2.0, 137


class C6_Wendland(AbstractSPHKernel):
    id, valid, nneigh = 5, 2.0, 356


def KernelFunctionValid(K) -> float:
    """Support radius in units of h (e.g. M4_spline.jl:62)."""
    return float(K.valid)


def KernelFunctionNneigh(K) -> int:
    return int(K.nneigh)


# ------------------------------------------------------------------------- strategies ---


class AbstractInterpolationStrategy:
    code = -1


class itpGather(AbstractInterpolationStrategy):
    code = _lib.PMX_GATHER


class itpScatter(AbstractInterpolationStrategy):
    code = _lib.PMX_SCATTER


class itpSymmetric(AbstractInterpolationStrategy):
    code = _lib.PMX_SYMMETRIC


# --------------------------------------------------------------------------- backend ----


class AbstractExecutionBackend:
    pass


class MI355XComputeBackend(AbstractExecutionBackend):
    """Backend tag of the MI355X library (≙ CUDAComputeBackend, CUDAExtDummy.jlh:12).

    precision: "f32c" (FP32 compensated, default) or "f64" (plain FP64)."""

    def __init__(self, device: int = 0, precision: str = "f32c"):
        if precision not in ("f32c", "f64"):
            raise ArgumentError(_lib.PMX_ERR_ARGUMENT, f"unknown precision {precision!r}")
        self.device = device
        self.precision = _lib.PMX_PREC_F32C if precision == "f32c" else _lib.PMX_PREC_F64
        self._ctx = None

    @property
    def ctx(self) -> _lib.Context:
        if self._ctx is None:
            self._ctx = _lib.Context(self.device)
        return self._ctx


# ------------------------------------------------------------------------ input/catalog -


class InterpolationInput:
    """Particle columns of one species (InterpolationInput.jl:43-118)."""

    def __init__(self, x, y, z, m, h, rho, quant=(), smoothed_kernel=M5_spline):
        cols = [x, y, z, m, h, rho] + list(quant)
        npart = len(m)
        for d, c in enumerate((x, y, z)):
            if len(c) != npart:
                raise DimensionMismatch(_lib.PMX_ERR_DIMENSION, f"coord[{d + 1}] length {len(c)} != Nparticles {npart}")
        if len(h) != npart:
            raise DimensionMismatch(_lib.PMX_ERR_DIMENSION, f"h length {len(h)} != Nparticles {npart}")
        if len(rho) != npart:
            raise DimensionMismatch(_lib.PMX_ERR_DIMENSION, f"ρ length {len(rho)} != Nparticles {npart}")
        for j, q in enumerate(quant):
            if len(q) != npart:
                raise DimensionMismatch(_lib.PMX_ERR_DIMENSION, f"quant[{j + 1}] length {len(q)} != Nparticles {npart}")
        dts = {str(c.dtype) for c in cols}
        if len(dts) != 1 or dts.pop() not in ("float32", "float64", "torch.float32", "torch.float64"):
            raise TypeError("all particle columns must share one element type, Float32 or Float64")
        self.Npart = npart
        self.coord = (x, y, z)
        self.m, self.h, self.ρ = m, h, rho
        self.quant = tuple(quant)
        self.smoothed_kernel = smoothed_kernel

    @property
    def rho(self):
        return self.ρ

    def __len__(self):
        return self.Npart

    @property
    def dtype(self):
        return self.m.dtype


_SUFFIX = ("ˣ", "ʸ", "ᶻ")


class InterpolationCatalog:
    """Requested outputs and their column slots (InterpolationCatalog.jl:9-225).

    Construct from column names like the reference:
        InterpolationCatalog(("rho", "vx", "vy", "vz"), scalars=("rho",), gradients=("rho",),
                             divergences=("v",), curls=("v",))
    Slots are stored 0-based; the density-gradient slot of the reference's explicit
    constructor (slot 0) is -1 here."""

    def __init__(self, column_names=None, scalars=(), gradients=(), divergences=(), curls=(), *, _slots=None):
        if _slots is not None:
            (self.scalar_names, self.scalar_slots, self.grad_names, self.grad_slots, self.div_names,
             self.div_slots, self.curl_names, self.curl_slots) = _slots
        else:
            if len(set(column_names)) != len(column_names):
                raise ArgumentError(_lib.PMX_ERR_ARGUMENT, "column_names must be unique")
            index = {nm: i for i, nm in enumerate(column_names)}

            def slot(nm):
                if nm not in index:
                    raise KeyError(nm)
                return index[nm]

            self.scalar_names = tuple(scalars)
            self.scalar_slots = tuple(slot(s) for s in scalars)
            self.grad_names = tuple(gradients)
            self.grad_slots = tuple(slot(s) for s in gradients)
            self.div_names = tuple(divergences)
            self.div_slots = tuple(tuple(slot(f"{v}{c}") for c in "xyz") for v in divergences)
            self.curl_names = tuple(curls)
            self.curl_slots = tuple(tuple(slot(f"{v}{c}") for c in "xyz") for v in curls)
        # Shepard normalisation is off only for the names ρ / rho (InterpolationCatalog.jl:222-225)
        self.scalar_snormalization = tuple(nm not in ("ρ", "rho") for nm in self.scalar_names)
        names = list(self.scalar_names)
        for g in self.grad_names:
            names += [f"∇{g}{s}" for s in _SUFFIX]
        names += [f"∇⋅{d}" for d in self.div_names]
        for c in self.curl_names:
            names += [f"∇×{c}{s}" for s in _SUFFIX]
        self.ordered_names = tuple(names)

    @classmethod
    def from_slots(cls, scalar_names=(), scalar_slots=(), grad_names=(), grad_slots=(), div_names=(), div_slots=(),
                   curl_names=(), curl_slots=()):
        """Explicit-slot constructor (InterpolationCatalog.jl:71-102), 0-based slots."""
        for a, b, what in ((scalar_names, scalar_slots, "scalar"), (grad_names, grad_slots, "grad"),
                           (div_names, div_slots, "div"), (curl_names, curl_slots, "curl")):
            if len(a) != len(b):
                raise DimensionMismatch(_lib.PMX_ERR_DIMENSION, f"{what}_slots length {len(b)} != {what}_names length {len(a)}")
        return cls(_slots=(tuple(scalar_names), tuple(scalar_slots), tuple(grad_names), tuple(grad_slots),
                           tuple(div_names), tuple(map(tuple, div_slots)), tuple(curl_names), tuple(map(tuple, curl_slots))))

    @property
    def L(self):
        return len(self.ordered_names)

    def struct(self):
        return _lib.catalog_struct(
            scalar_slots=list(self.scalar_slots), scalar_shepard=[int(b) for b in self.scalar_snormalization],
            grad_slots=list(self.grad_slots), div_slots=[s for t in self.div_slots for s in t],
            curl_slots=[s for t in self.curl_slots for s in t])


# ----------------------------------------------------------------------------- samples ---


class PointSamples:
    """Values at explicit Cartesian sample points (PointSamples.jl:25-39)."""

    def __init__(self, grid, coor):
        self.grid = grid
        self.coor = tuple(coor)

    def __len__(self):
        return len(self.grid)

    def similar(self, dtype=None):
        return PointSamples(np.empty(len(self.grid), dtype or self.grid.dtype), self.coor)


class LineSamples:
    """Values along sample lines, origin + unit direction (LineSamples.jl:31-47)."""

    def __init__(self, grid, origin, direction):
        self.grid = grid
        self.origin = tuple(origin)
        self.direction = tuple(direction)

    def __len__(self):
        return len(self.grid)

    def similar(self, dtype=None):
        return LineSamples(np.empty(len(self.grid), dtype or self.grid.dtype), self.origin, self.direction)


class Cartesian:
    pass


class Cylindrical:
    pass


class Spherical:
    pass


class StructuredGrid:
    """Tensor-product grid (StructuredGrid.jl:42); values column-major over the axes."""

    def __init__(self, grid, axes):
        self.axes = tuple(np.asarray(a, np.float64) for a in axes)
        self.size = tuple(len(a) for a in self.axes)
        self.grid = grid

    @classmethod
    def build(cls, COORD, *params):
        from .disc import cylindrical_axes, linrange, spherical_axes
        if COORD is Cartesian:
            axes = [linrange(a, b, n) for a, b, n in params]
        elif COORD is Cylindrical:
            axes = list(cylindrical_axes(*params))
        elif COORD is Spherical:
            # StructuredGrid(Spherical, rparams | rconst, phiparams, thetaparams), StructuredGrid.jl:427-518
            if len(params) != 3:
                raise ArgumentError(_lib.PMX_ERR_ARGUMENT, "Spherical grids take (r | rconst, phi, theta)")
            r, phi, theta = params
            shell = np.isscalar(r)
            nr = 1 if shell else r[2]
            if not (nr >= 1 and phi[2] >= 1 and theta[2] >= 1):
                raise ArgumentError(_lib.PMX_ERR_ARGUMENT, "nr, nϕ, nθ must be ≥ 1")
            if shell and not r >= 0.0:
                raise ArgumentError(_lib.PMX_ERR_ARGUMENT, "radial range must satisfy r ≥ 0 ")
            if not shell and not (r[0] >= 0.0 and r[1] > r[0]):
                raise ArgumentError(_lib.PMX_ERR_ARGUMENT, "radial range must satisfy 0 ≤ rmin < rmax")
            if not (phi[0] >= 0.0 and (phi[1] <= 2 * np.pi or np.isclose(phi[1], 2 * np.pi)) and phi[1] > phi[0]):
                raise ArgumentError(_lib.PMX_ERR_ARGUMENT, "angular range must satisfy 0 ≤ ϕmin < ϕmax ≤ 2π")
            if not (theta[0] >= 0.0 and theta[1] <= np.pi and theta[1] > theta[0]):
                raise ArgumentError(_lib.PMX_ERR_ARGUMENT, "polar range must satisfy 0 ≤ θmin < θmax ≤ π")
            axes = list(spherical_axes(r, phi, theta))
        else:
            raise ArgumentError(_lib.PMX_ERR_ARGUMENT, f"unsupported coordinate system {COORD}")
        return cls(np.zeros([len(a) for a in axes], order="F"), axes)


def flatten(COORD, grid: StructuredGrid) -> PointSamples:
    """transform.jl:53-72: Cartesian sample coordinates in column-major order (Cartesian,
    Cylindrical, Spherical)."""
    from .disc import flatten_cartesian, flatten_cylindrical, flatten_spherical
    if COORD is Cartesian:
        coor = flatten_cartesian(*grid.axes)
    elif COORD is Cylindrical:
        coor = flatten_cylindrical(*grid.axes)
    elif COORD is Spherical:
        coor = flatten_spherical(*grid.axes)
    else:
        raise ArgumentError(_lib.PMX_ERR_ARGUMENT, f"unsupported coordinate system {COORD}")
    return PointSamples(np.ravel(grid.grid, order="F").copy(), coor)


def restore_struct(COORD, samples: PointSamples, axes, atol=1e-8, rtol=1e-8) -> StructuredGrid:
    """transform.jl:107-116: check the coordinates against the axes, then reshape."""
    tmpl = StructuredGrid(np.zeros([len(a) for a in axes], order="F"), axes)
    expect = flatten(COORD, tmpl).coor
    for a, b in zip(samples.coor, expect):
        if not np.allclose(a, b, atol=atol, rtol=rtol):
            raise ArgumentError(_lib.PMX_ERR_ARGUMENT,
                                f"PointSamples coordinates do not match the provided axes under {COORD.__name__} dispatch")
    return StructuredGrid(np.reshape(samples.grid, tmpl.size, order="F"), axes)


class GridBundle:
    """Named output grids (GridBundle.jl:1-4)."""

    def __init__(self, grids, names):
        self.grids = tuple(grids)
        self.names = tuple(names)

    def __getitem__(self, name):
        return self.grids[self.names.index(name)]


# --------------------------------------------------------------- GridDataset (HDF5) ----

PARTIA_VERSION = "0.1.0"      # the version of Partia.jl whose interface this mirrors (Project.toml)


class GridDataset:
    """A GridBundle with its metadata (GridDataset.jl:3-6): params maps names to str, int, bool or
    float values."""

    def __init__(self, data: GridBundle, params: dict):
        self.data = data
        self.params = dict(params)


def file_identifier(operation_name: str = "", now=None) -> str:
    """logging.jl:152-156: "FT:Partia:<version> (<operation | CommonOutput>): dd/mm/yyyy HH:MM:SS.s"
    (Julia's "s" is the millisecond count without padding)."""
    import datetime
    t = now or datetime.datetime.now()
    stamp = t.strftime("%d/%m/%Y %H:%M:%S") + f".{t.microsecond // 1000}"
    return f"FT:Partia:{PARTIA_VERSION} ({operation_name or 'CommonOutput'}): {stamp}"


def _float_name(dtype) -> str:
    dt = np.dtype(dtype)
    if dt == np.float32:
        return "Float32"
    if dt == np.float64:
        return "Float64"
    raise ArgumentError(_lib.PMX_ERR_ARGUMENT, f"unsupported grid element type {dt}")


def write_GridDataset(gd: GridDataset, filename: str = "PRGridDataset.h5") -> None:
    """write_grids.jl:29-74 (PointSamples) and 98-143 (StructuredGrid): every grid must share the
    first grid's coordinate / axis vectors (object identity, as the reference's ===)."""
    from . import griddataset as G
    grids = gd.data.grids
    g1 = grids[0]
    if isinstance(g1, StructuredGrid):
        kind, vecs, what = G.STRUCTURED, g1.axes, "StructuredGrid axes"
        for i, gi in enumerate(grids[1:], start=2):
            for d in range(len(vecs)):
                if gi.axes[d] is not vecs[d]:
                    raise ArgumentError(_lib.PMX_ERR_ARGUMENT, f"{what} not shared: grid={i} dim={d + 1}")
    elif isinstance(g1, PointSamples):
        kind, vecs, what = G.POINTSAMPLES, g1.coor, "PointSamples coor"
        for i, gi in enumerate(grids[1:], start=2):
            for d in range(len(vecs)):
                if gi.coor[d] is not vecs[d]:
                    raise ArgumentError(_lib.PMX_ERR_ARGUMENT, f"{what} not shared: grid={i} dim={d + 1}")
    else:
        raise ArgumentError(_lib.PMX_ERR_ARGUMENT, "GridDataset grids must be StructuredGrid or PointSamples")
    try:
        G.write(filename, kind, list(vecs), gd.data.names, [np.asarray(g.grid) for g in grids],
                list(gd.params.items()))
    except G.GridDatasetError as e:
        if e.code == G.GD_ERR_ARGUMENT:
            raise ArgumentError(_lib.PMX_ERR_ARGUMENT, str(e)) from None
        raise


def write_GridBundle(gb: GridBundle, filename: str = "PRGridDataset.h5", *, code_units: dict,
                     operation_name: str = "", params: dict | None = None) -> None:
    """write_grids.jl:181-215: schema, system, dtype, grid type, code units and file identifier
    added to the metadata, then write_GridDataset."""
    import platform
    p = {} if params is None else dict(params)
    required = ("umass", "udist", "utime", "umagfd")
    notfound = tuple(k for k in required if k not in code_units)
    if notfound:
        shown = "(" + ", ".join(":" + k for k in notfound) + ("," if len(notfound) == 1 else "") + ")"
        raise ArgumentError(_lib.PMX_ERR_ARGUMENT, f"code_units missing required keys: {shown}")
    p["schema_name"] = "GridDataset"
    p["schema_id"] = "GridDataset.v1"
    p["system_kernel"] = platform.system()              # Julia's string(Sys.KERNEL), e.g. "Linux"
    p["dtype"] = _float_name(np.asarray(gb.grids[0].grid).dtype)
    p["grid_type"] = type(gb.grids[0]).__name__
    for k, v in code_units.items():
        p[k] = float(v)
    p["file_identifier"] = file_identifier(operation_name)
    write_GridDataset(GridDataset(gb, p), filename)


def read_GridDataset(filename: str = "PRGridDataset.h5") -> GridDataset:
    """read_grids.jl:29-86: the grids share one set of vectors (read once); StructuredGrid values
    are reshaped column-major to the axes' lengths."""
    from . import griddataset as G
    try:
        kind, vecs, names, flat, params = G.read(filename)
    except G.GridDatasetError as e:
        if e.code == G.GD_ERR_ARGUMENT:
            raise ArgumentError(_lib.PMX_ERR_ARGUMENT, str(e)) from None
        raise
    if kind == G.STRUCTURED:
        axes = tuple(vecs)
        size = tuple(v.size for v in axes)
        grids = []
        for g in flat:
            sg = StructuredGrid(np.reshape(g, size, order="F"), axes)
            sg.axes = axes                     # one shared set of axis vectors
            grids.append(sg)
    else:
        coor = tuple(vecs)
        grids = [PointSamples(g, coor) for g in flat]
    return GridDataset(GridBundle(grids, names), params)


# --------------------------------------------------------------------------- drivers ----


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _build(backend, input: InterpolationInput):
    ctx = backend.ctx
    cols = [input.coord[0], input.coord[1], input.coord[2], input.m, input.h, input.ρ]
    for c in cols + list(input.quant):
        if isinstance(c, np.ndarray) and not c.flags.c_contiguous:
            raise ValueError("particle columns must be contiguous arrays")
    # LinearBVH! (InterpolationInput.jl:203-224): build and permute the caller's columns
    ctx.build(*cols, quant=list(input.quant), kernel=input.smoothed_kernel.id, writeback=True)
    return ctx


def PointSamples_interpolation(backend, grid_template: PointSamples, input: InterpolationInput,
                               catalog: InterpolationCatalog, itp_strategy=itpSymmetric) -> GridBundle:
    if not isinstance(backend, MI355XComputeBackend):
        raise ArgumentError(_lib.PMX_ERR_ARGUMENT, "this package provides the MI355XComputeBackend only")
    ctx = _build(backend, input)
    n = len(grid_template)
    dt = np.float32 if str(input.dtype).endswith("float32") else np.float64
    outs = [np.empty(n, dt) for _ in range(catalog.L)]
    qx, qy, qz = (_f64(c) for c in grid_template.coor)
    ctx.interpolate_points(itp_strategy.code, backend.precision, catalog.struct(), qx, qy, qz, outs)
    grids = [PointSamples(o, grid_template.coor) for o in outs]
    return GridBundle(grids, catalog.ordered_names)


def LineSamples_interpolation(backend, grid_template: LineSamples, input: InterpolationInput,
                              catalog: InterpolationCatalog, itp_strategy=itpScatter, *,
                              inv_h2: bool = False) -> GridBundle:
    """line_integrated_scalar_interpolation (itpScatter).  inv_h2 (an extension, off by default):
    weight each pair by Inorm(q) / h^2, the dimensionally correct column kernel, instead of the
    reference's Inorm(q) / h (line_integrated_kernel.jl:132-137)."""
    if itp_strategy is not itpScatter:
        raise ArgumentError(_lib.PMX_ERR_ARGUMENT,
                            "LineSamples_interpolation only supports itpScatter. Line-integrated samples do not have a "
                            "well-defined query smoothing length ha, so itpGather and itpSymmetric are not supported.")
    if catalog.grad_slots or catalog.div_slots or catalog.curl_slots:
        raise ArgumentError(_lib.PMX_ERR_ARGUMENT, "LineSamples_interpolation: catalog must hold scalars only")
    ctx = _build(backend, input)
    n = len(grid_template)
    dt = np.float32 if str(input.dtype).endswith("float32") else np.float64
    outs = [np.empty(n, dt) for _ in range(catalog.L)]
    o = [_f64(c) for c in grid_template.origin]
    d = [_f64(c) for c in grid_template.direction]
    prec = backend.precision | (_lib.PMX_LINE_INV_H2 if inv_h2 else 0)
    ctx.interpolate_lines(itpScatter.code, prec, catalog.struct(), o, d, outs)
    grids = [LineSamples(v, grid_template.origin, grid_template.direction) for v in outs]
    return GridBundle(grids, catalog.ordered_names)


def StructuredGrid_interpolation(backend, COORD, grid_template: StructuredGrid, input: InterpolationInput,
                                 catalog: InterpolationCatalog, itp_strategy=itpSymmetric) -> GridBundle:
    if COORD is Cartesian and len(grid_template.axes) == 3:
        # Cartesian samples are copies of the axis values: flattened on the device (bit-equal
        # to flatten + coordinate_grid, transform.jl:53-56, StructuredGrid.jl:126-141)
        if not isinstance(backend, MI355XComputeBackend):
            raise ArgumentError(_lib.PMX_ERR_ARGUMENT, "this package provides the MI355XComputeBackend only")
        ctx = _build(backend, input)
        axes = [_f64(a) for a in grid_template.axes]
        size = tuple(len(a) for a in axes)
        dt = np.float32 if str(input.dtype).endswith("float32") else np.float64
        outs = [np.empty(int(np.prod(size)), dt) for _ in range(catalog.L)]
        ctx.interpolate_cartesian_grid(itp_strategy.code, backend.precision, catalog.struct(), *axes, outs)
        grids = [StructuredGrid(np.reshape(o, size, order="F"), grid_template.axes) for o in outs]
        return GridBundle(grids, catalog.ordered_names)
    flat = flatten(COORD, grid_template)
    bundle = PointSamples_interpolation(backend, flat, input, catalog, itp_strategy)
    grids = [restore_struct(COORD, g, grid_template.axes) for g in bundle.grids]
    return GridBundle(grids, bundle.names)
