"""Shared builders for the parity tests: synthetic discs, catalogs, oracle runs."""
import numpy as np

from distorteddustydiscanalysistoolkits_amd import disc
from oracle import oracle as O

# The 8-output field set: columns (rho, vx, vy, vz); scalars (rho,), gradients (rho,), divergences (v,),
# curls (v,) -> L = 8 (rho is not Shepard-normalised, InterpolationCatalog.jl:222-225).
CAT_RHO_V = dict(scalar_slots=[0], scalar_shepard=[0], grad_slots=[0], div_slots=[1, 2, 3],
              curl_slots=[1, 2, 3])


def disc_columns(n, seed=0, warped=False, dtype=np.float64):
    d = disc.make_disc(n, seed, warped)
    cols = [np.ascontiguousarray(d[k], dtype) for k in ("x", "y", "z", "m", "h", "rho")]
    quant = [np.ascontiguousarray(d[k], dtype) for k in ("rho", "vx", "vy", "vz")]
    return cols, quant


def oracle_tree(cols, quant):
    """Oracle LinearBVH! on FP64 copies (promoting FP32 exactly)."""
    c = [np.array(a, np.float64) for a in cols]
    q = [np.array(a, np.float64) for a in quant]
    return O.LBVH(*c, quant=q), c, q


def cyl_points(ns, nphi, nz, zr=5.0):
    sa, pa, za = disc.cylindrical_axes((1.0, 100.0, ns), (0.0, 2 * np.pi, nphi), (-zr, zr, nz))
    return disc.flatten_cylindrical(sa, pa, za)


def csr_sets(off, idx):
    return [idx[off[i]:off[i + 1]] for i in range(len(off) - 1)]


def rel_norms(got, ref):
    """Normwise relative L-inf and L2 error over finite reference entries."""
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    m = np.isfinite(ref)
    if not m.any():
        return 0.0, 0.0
    e = got[m] - ref[m]
    linf = np.abs(e).max() / max(np.abs(ref[m]).max(), 1e-300)
    l2 = np.linalg.norm(e) / max(np.linalg.norm(ref[m]), 1e-300)
    return float(linf), float(l2)
