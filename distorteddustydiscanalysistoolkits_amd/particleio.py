"""Host mirror of ParticleIO.jl's input side for the interpolation path: read a Phantom dump,
add the density column, and build Partia's InterpolationInput + InterpolationCatalog — with the
particle columns read straight into device memory (HBM) when a device is given.

Reference:
  * read_phantom          ParticleIO.jl/src/julia/IO/phantomIO/read_phantom.jl:273-321
                          (format parsing :20-225 is native: pmx_phantom_* in libpartia_mi355x.so)
  * ParticleDataFrame     ParticleIO.jl/src/julia/Particles/ParticleDataFrame.jl:18-21, 156-223
  * add_rho!              ParticleIO.jl/src/julia/Particles/add_quantities.jl:294-347
  * MassFromColumn/Params ParticleIO.jl/src/julia/PartiaAdapter/MassSource.jl
  * build_input           ParticleIO.jl/src/julia/PartiaAdapter/build_input.jl:44-219 and
                          Partia.jl/src/julia/KernelInterpolation/interpolation_setup/constructor.jl:69-140

Differences, all forced by the host language: Julia's `missing` (columns absent from one part
of `separate_types="none"` with sinks) is NaN here, with integer columns widened to float64;
DataFrames' `groupby` order for `separate_types="all"` is taken as ascending itype.
"""
from __future__ import annotations

import math
import warnings

import numpy as np

from . import _lib
from .partia import ArgumentError, InterpolationCatalog, InterpolationInput, M5_spline

__all__ = ["ParticleDataFrame", "read_phantom", "get_dim", "get_time", "get_npart", "add_rho",
           "MassFromColumn", "MassFromParams", "build_input"]


class ParticleDataFrame:
    """Columns (`dfdata`, name -> host array or device tensor, in file order) and the dump's
    global parameters (`params`)."""

    def __init__(self, dfdata: dict, params: dict):
        self.dfdata = dict(dfdata)
        self.params = dict(params)

    def names(self):
        return list(self.dfdata)

    def __getitem__(self, key):
        if isinstance(key, tuple):
            rows, col = key
            c = self.dfdata[col]
            return c if (rows is Ellipsis or rows == slice(None)) else c[rows]
        return self.dfdata[key]

    def __setitem__(self, key, value):
        if self.dfdata and len(value) != get_npart(self):
            raise _lib.DimensionMismatch(_lib.PMX_ERR_DIMENSION,
                                         f"column {key} has {len(value)} rows, the data frame {get_npart(self)}")
        self.dfdata[key] = value

    def __contains__(self, key):
        return key in self.dfdata


def get_dim(data: ParticleDataFrame) -> int:
    return 3 if "z" in data.dfdata else 2


def get_time(data: ParticleDataFrame):
    if "time" in data.params:
        return data.params["time"]
    if "Time" in data.params:
        return data.params["Time"]
    return None


def get_npart(data: ParticleDataFrame) -> int:
    return len(next(iter(data.dfdata.values()))) if data.dfdata else 0


def _torch_device(device):
    if device is None:
        return None
    import torch
    return torch.device(device)


def _read(ph, cols, keep, device, have=None):
    """Columns (index, name) of the open dump, keeping rows `keep`, into host arrays or tensors.
    have: {column index: host array of the whole column} already read (uploaded, not re-read)."""
    out = {}
    for i, name in cols:
        _, blk, code = ph.columns[i]
        n = ph.block_sizes[blk] if keep is None else len(keep)
        if have and i in have:
            a = have[i] if keep is None else have[i][keep]
            if isinstance(a, np.ndarray):
                a = np.array(a, copy=True) if device is None else _to_device(a, device)
            out[name] = a
        elif device is None:
            out[name] = ph.read_column(i, keep=keep)
        else:
            import torch
            t = torch.empty(n, dtype=getattr(torch, np.dtype(_lib._NP_OF[code]).name), device=device)
            out[name] = ph.read_column(i, out=t, keep=keep)
    return out


def _to_device(a, device):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(device)


def read_phantom(filename: str, separate_types: str = "sinks", ignore_inactive: bool = True, device=None):
    """Read a Phantom dump (read_phantom.jl:273-321) into a list of ParticleDataFrame.

    separate_types: "sinks" (gas, then sinks when present), "all" (one frame per itype when the
    dump holds several, then sinks), anything else: one combined frame.  ignore_inactive drops
    particles with h <= 0 (or NaN).  device: None for host numpy columns, or a torch device
    (e.g. "cuda:0") to read every column straight into device memory."""
    dev = _torch_device(device)
    with _lib.Phantom(filename) as ph:
        header = dict(ph.header)
        header["default_Int"] = np.int32 if ph.int_bytes == 4 else np.int64
        header["default_Real"] = np.float32 if ph.real_bytes == 4 else np.float64
        header["file_identifier"] = ph.file_identifier
        header["COM_coordinate"] = np.zeros(6)
        header["Origin_sink_id"] = -1
        header["Origin_sink_mass"] = math.nan
        main = [(i, c[0]) for i, c in enumerate(ph.columns) if c[1] != 1]
        sinks = [(i, c[0]) for i, c in enumerate(ph.columns) if c[1] == 1]
        main_idx = {nm: i for i, nm in main}
        n_main = ph.block_sizes[ph.columns[main[0][0]][1]] if main else 0
        n_sink = ph.block_sizes[1] if len(ph.block_sizes) > 1 else 0
        sinks_nonempty = bool(sinks) and n_sink > 0

        keep = None
        have = {}
        ih = main_idx.get("h")
        if ignore_inactive:
            if ih is None:
                raise ArgumentError(_lib.PMX_ERR_ARGUMENT, "column name :h not found in the data frame")
            if dev is not None:
                # h straight into device memory, tested there; the host sees h only when some
                # particles are inactive and the kept rows must be listed
                import torch
                hd = ph.read_column(ih, out=torch.empty(n_main, dtype=getattr(
                    torch, np.dtype(_lib._NP_OF[ph.columns[ih][2]]).name), device=dev))
                if bool((hd > 0).all()):
                    have[ih] = hd
                else:
                    h = hd.cpu().numpy()
                    have[ih] = h
                    keep = np.flatnonzero(h > 0).astype(np.int64)
            else:
                h = ph.read_column(ih)
                have[ih] = h
                active = h > 0
                # every particle active (the usual case): whole columns, no row list to build
                keep = None if active.all() else np.flatnonzero(active).astype(np.int64)

        if separate_types == "all" and "itype" in main_idx:
            it = ph.read_column(main_idx["itype"])
            rows = np.arange(n_main, dtype=np.int64) if keep is None else keep
            if ih in have and not isinstance(have[ih], np.ndarray):
                have[ih] = have[ih].cpu().numpy()       # grouped rows are selected on the host
            types = np.unique(it[rows])
            if types.size > 1:
                frames = []
                for v in types:
                    itype = int(v)
                    mass_key = "massoftype" if itype == 1 else f"massoftype_{itype}"
                    if mass_key not in header:
                        raise KeyError(f"key :{mass_key} not found")
                    params = dict(header)
                    params["mass"] = header[mass_key]
                    params["itype"] = itype
                    frames.append(ParticleDataFrame(_read(ph, main, rows[it[rows] == v], dev, have), params))
                if sinks_nonempty:
                    frames.append(ParticleDataFrame(_read(ph, sinks, None, dev), header))
                return frames

        if separate_types in ("sinks", "all") and sinks_nonempty:
            params = dict(header)
            params["mass"] = None
            params["itype"] = math.nan
            return [ParticleDataFrame(_read(ph, main, keep, dev, have), params),
                    ParticleDataFrame(_read(ph, sinks, None, dev), header)]

        cols = _read(ph, main, keep, dev, have)
        if sinks_nonempty:
            cols = _vcat_union(cols, _read(ph, sinks, None, dev))
        params = dict(header)
        params["mass"] = None
        params["itype"] = None
        return [ParticleDataFrame(cols, params)]


def _vcat_union(a: dict, b: dict) -> dict:
    """vcat(a, b; cols = :union) with NaN for the rows of the part that lacks a column."""
    na = len(next(iter(a.values()))) if a else 0
    nb = len(next(iter(b.values()))) if b else 0
    out = {}
    for name in list(a) + [k for k in b if k not in a]:
        pa, pb = a.get(name), b.get(name)
        ref = pa if pa is not None else pb
        if isinstance(ref, np.ndarray):
            dt = ref.dtype if (pa is not None and pb is not None and pa.dtype == pb.dtype) else np.float64
            fill = lambda n: np.full(n, np.nan, dt)  # noqa: E731
            pa_ = pa.astype(dt) if pa is not None else fill(na)
            pb_ = pb.astype(dt) if pb is not None else fill(nb)
            out[name] = np.concatenate([pa_, pb_])
        else:
            import torch
            dt = ref.dtype if (pa is not None and pb is not None and pa.dtype == pb.dtype) else torch.float64
            fill = lambda n: torch.full((n,), math.nan, dtype=dt, device=ref.device)  # noqa: E731
            pa_ = pa.to(dt) if pa is not None else fill(na)
            pb_ = pb.to(dt) if pb is not None else fill(nb)
            out[name] = torch.cat([pa_, pb_])
    return out


def add_rho(data: ParticleDataFrame) -> None:
    """add_rho! (add_quantities.jl:331-347): rho = mass * hfact^d / h^d for equal-mass particles,
    in the mass parameter's type, computed by the library (on the device for device columns)."""
    if "m" in data.dfdata:
        warnings.warn("This function assumes all particles share the same mass. It is invalid if per-particle "
                      "masses are used.")
    if "mass" not in data.params:
        raise RuntimeError("KeyError: Missing required parameter :mass in data.params.")
    m, hfact = data.params["mass"], data.params["hfact"]
    tf = _float_type(m)
    if tf is None or _float_type(hfact) is not tf:
        raise TypeError(f"no method _density(h, m::{type(m).__name__}, hfact::{type(hfact).__name__}): the mass and "
                        "hfact parameters must share one floating-point type")
    h = data.dfdata["h"]
    if isinstance(h, np.ndarray):
        h = np.ascontiguousarray(h, dtype=tf)
    else:
        import torch
        h = h.to(getattr(torch, np.dtype(tf).name)).contiguous()
    data.dfdata["rho"] = _lib.density_from_h(h, m, hfact, dim=get_dim(data))


def _float_type(v):
    if isinstance(v, (np.float32,)):
        return np.float32
    if isinstance(v, (float, np.float64)):
        return np.float64
    return None


class MassFromColumn:
    """Particle masses from a data-frame column (MassSource.jl)."""

    def __init__(self, name: str):
        self.name = name


class MassFromParams:
    """One particle mass for all particles, from data.params (MassSource.jl)."""

    def __init__(self, name: str):
        self.name = name


def _vector_components(name):
    return tuple(f"{name}{s}" for s in ("x", "y", "z"))


def _quantity_column_names(scalars, gradients, divergences, curls):
    names = []
    for nm in list(scalars) + list(gradients):
        if nm not in names:
            names.append(nm)
    for base in list(divergences) + list(curls):
        for comp in _vector_components(base):
            if comp not in names:
                names.append(comp)
    return tuple(names)


def _missing_particle_columns(data, column_names, mass_column=None):
    have = set(data.names())
    missing = []
    for base in ("x", "y", "z", "h", "rho"):
        if base not in have and base not in missing:
            missing.append(base)
    if mass_column is not None and mass_column not in have and mass_column not in missing:
        missing.append(mass_column)
    for nm in column_names:
        if nm not in have and nm not in missing:
            missing.append(nm)
    return missing


def _promoted(cols):
    """Julia promote_type over the columns' element types: Float64 if any is Float64, else
    Float32 if any is Float32 (integers promote to the float type)."""
    names = {str(c.dtype).replace("torch.", "") for c in cols}
    if "float64" in names:
        return "float64"
    if "float32" in names:
        return "float32"
    raise TypeError("InterpolationInput needs floating-point columns")


def _as(c, dt):
    if isinstance(c, np.ndarray):
        return np.array(c, dtype=dt, copy=True)
    import torch
    return c.to(getattr(torch, dt), copy=True).contiguous()


def build_input(data: ParticleDataFrame, mass_source, scalars=(), gradients=(), divergences=(), curls=(),
                smoothed_kernel=M5_spline):
    """InterpolationInput + InterpolationCatalog from a ParticleDataFrame (build_input.jl:44-219).
    Device-tensor columns stay on the device."""
    column_names = _quantity_column_names(scalars, gradients, divergences, curls)
    mass_col = mass_source.name if isinstance(mass_source, MassFromColumn) else None
    if not isinstance(mass_source, (MassFromColumn, MassFromParams)):
        raise TypeError("mass_source must be MassFromColumn or MassFromParams")
    missing = _missing_particle_columns(data, column_names, mass_col)
    if missing:
        raise ArgumentError(_lib.PMX_ERR_ARGUMENT, "Missing columns in ParticleDataFrame: " + ", ".join(missing))
    x, y, z, h, rho = (data.dfdata[k] for k in ("x", "y", "z", "h", "rho"))
    if mass_col is not None:
        m = data.dfdata[mass_col]
    else:
        if mass_source.name not in data.params:
            raise ArgumentError(_lib.PMX_ERR_ARGUMENT,
                                f"Missing parameter in ParticleDataFrame.params: {mass_source.name}")
        mv = data.params[mass_source.name]
        n = get_npart(data)
        mdt = _float_type(mv) or np.float64
        if isinstance(x, np.ndarray):
            m = np.full(n, mv, dtype=mdt)
        else:
            import torch
            m = torch.full((n,), float(mv), dtype=getattr(torch, np.dtype(mdt).name), device=x.device)
    quant = [data.dfdata[nm] for nm in column_names]
    dt = _promoted([x, y, z, m, h, rho] + quant)
    inp = InterpolationInput(*(_as(c, dt) for c in (x, y, z, m, h, rho)), quant=tuple(_as(q, dt) for q in quant),
                             smoothed_kernel=smoothed_kernel)
    cat = InterpolationCatalog(column_names, scalars=tuple(scalars), gradients=tuple(gradients),
                               divergences=tuple(divergences), curls=tuple(curls))
    return inp, cat
