"""Deterministic synthetic protoplanetary-disc particle sets and sample geometries.

Counter-based SplitMix64 (one independent stream per variable) plus Box-Muller normals,
so every particle is a pure function of (seed, stream, index) and any size can be drawn
reproducibly in numpy.

Disc model (flat):  r ~ U[1, 100] (surface density Sigma ~ 1/r), phi ~ U[0, 2 pi),
z ~ N(0, H) with H = 0.05 r^1.25; equal masses summing to 1e-3; analytic density
rho = Sigma / (sqrt(2 pi) H) exp(-z^2 / 2H^2) with Sigma = Sigma0 / r,
Sigma0 = 1e-3 / (198 pi); h = 1.2 (m / rho)^(1/3); Keplerian v_phi = r^-1/2 plus 1 %
Gaussian noise per component.
Warped disc: positions and velocities rotated by the tilt beta(r) = 0.2 r / 100 about
the in-plane axis (cos gamma, sin gamma, 0), gamma = 0.5 ln r (Rodrigues, pre-warp r).
Dust: H_d = 0.3 H, Sigma_d = 0.01 Sigma_g, v_phi lowered by 0.5 %.
"""
from __future__ import annotations

from fractions import Fraction

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def _mix(z: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def splitmix64(seed: int, stream: int, n: int, start: int = 0) -> np.ndarray:
    """n outputs of the SplitMix64 sequence keyed by (seed, stream)."""
    with np.errstate(over="ignore"):
        key = _mix(np.array([np.uint64(seed) * _GOLDEN + np.uint64(stream + 1) * _M2], np.uint64))[0]
        ctr = np.arange(start + 1, start + n + 1, dtype=np.uint64) * _GOLDEN + key
    return _mix(ctr)


def uniform(seed: int, stream: int, n: int) -> np.ndarray:
    """Uniform doubles in [0, 1) from the top 53 bits."""
    return (splitmix64(seed, stream, n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def normal(seed: int, stream: int, n: int) -> np.ndarray:
    """Standard normals by Box-Muller from two uniform streams (stream, stream + 1000)."""
    u1 = 1.0 - uniform(seed, stream, n)          # (0, 1]
    u2 = uniform(seed, stream + 1000, n)
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)


SIGMA0 = 1e-3 / (198.0 * np.pi)


def _rotate(vx, vy, vz, ax, ay, beta):
    """Rodrigues rotation by angle beta about the unit in-plane axis (ax, ay, 0)."""
    c, s = np.cos(beta), np.sin(beta)
    dot = ax * vx + ay * vy
    cx = ay * vz            # axis x v = (ay*vz, -ax*vz, ax*vy - ay*vx)
    cy = -ax * vz
    cz = ax * vy - ay * vx
    rx = vx * c + cx * s + ax * dot * (1 - c)
    ry = vy * c + cy * s + ay * dot * (1 - c)
    rz = vz * c + cz * s
    return rx, ry, rz


def make_disc(n: int, seed: int = 0, warped: bool = False, *, total_mass: float = 1e-3,
              h_scale: float = 1.0, vphi_factor: float = 1.0, stream_base: int = 0) -> dict:
    """Draw one disc species.  Returns float64 columns x, y, z, m, h, rho, vx, vy, vz."""
    b = stream_base
    r = 1.0 + 99.0 * uniform(seed, b + 1, n)
    phi = 2.0 * np.pi * uniform(seed, b + 2, n)
    H = h_scale * 0.05 * r ** 1.25
    z = H * normal(seed, b + 3, n)
    m = np.full(n, total_mass / n)
    sigma = (total_mass / 1e-3) * SIGMA0 / r
    rho = sigma / (np.sqrt(2.0 * np.pi) * H) * np.exp(-z * z / (2.0 * H * H))
    h = 1.2 * np.cbrt(m / rho)
    vk = vphi_factor * r ** -0.5
    noise = 0.01 * r ** -0.5
    vx = -vk * np.sin(phi) + noise * normal(seed, b + 4, n)
    vy = vk * np.cos(phi) + noise * normal(seed, b + 5, n)
    vz = noise * normal(seed, b + 6, n)
    x = r * np.cos(phi)
    y = r * np.sin(phi)
    if warped:
        beta = 0.2 * (r / 100.0)
        gamma = 0.5 * np.log(r)
        ax, ay = np.cos(gamma), np.sin(gamma)
        x, y, z = _rotate(x, y, z, ax, ay, beta)
        vx, vy, vz = _rotate(vx, vy, vz, ax, ay, beta)
    return {"x": x, "y": y, "z": z, "m": m, "h": h, "rho": rho, "vx": vx, "vy": vy, "vz": vz}


def make_dust(n: int, seed: int = 0, warped: bool = False, gas_mass: float = 1e-3) -> dict:
    """Dust species: H_d = 0.3 H, dust-to-gas mass 0.01, v_phi x 0.995."""
    return make_disc(n, seed, warped, total_mass=0.01 * gas_mass, h_scale=0.3,
                     vphi_factor=0.995, stream_base=100)


# ------------------------------------------------------------------------------------
# Sample geometries (host-side flatten, as Partia's StructuredGrid flatten runs on the
# host: P/src/julia/Grids/grids/transform.jl:53-72, StructuredGrid.jl:126-204).
# ------------------------------------------------------------------------------------
def _fma(x: float, y: float, z: float) -> float:
    """Correctly rounded x*y + z (exact rational arithmetic)."""
    return float(Fraction(x) * Fraction(y) + Fraction(z))


def linrange(a: float, b: float, n: int) -> np.ndarray:
    """LinRange{Float64}(a, b, n) element values: t = j/(n-1), fma(t, b, fma(-t, a, a))
    (Julia Base.lerpi)."""
    if n == 1:
        return np.array([a], np.float64)
    a, b = float(a), float(b)
    out = np.empty(n, np.float64)
    for j in range(n):
        t = j / (n - 1)
        out[j] = _fma(t, b, _fma(-t, a, a))
    return out


def cylindrical_axes(s=(1.0, 100.0, 128), phi=(0.0, 2 * np.pi, 128), z=(-5.0, 5.0, 32)):
    """Axes of StructuredGrid(Cylindrical, s, phi, z) (StructuredGrid.jl:366-390):
    LinRange on s and z, half-open [phimin, phimax) with step (phimax - phimin) / nphi."""
    sa = linrange(*s)
    dphi = (phi[1] - phi[0]) / phi[2]
    pa = phi[0] + dphi * np.arange(phi[2], dtype=np.float64)
    za = linrange(*z)
    return sa, pa, za


def flatten_cylindrical(sa, pa, za):
    """Cartesian sample points of a cylindrical grid in column-major (s fastest) order,
    matching vec(grid.grid) (StructuredGrid.jl:162-181, _cylin2cart)."""
    S, P, Z = np.meshgrid(sa, pa, za, indexing="ij")
    S, P, Z = (a.ravel(order="F") for a in (S, P, Z))
    return S * np.cos(P), S * np.sin(P), Z.copy()


def spherical_axes(r=(1.0, 100.0, 64), phi=(0.0, 2 * np.pi, 64), theta=(0.0, np.pi, 32)):
    """Axes of StructuredGrid(Spherical, r, phi, theta) (StructuredGrid.jl:492-518): LinRange on r,
    half-open [phimin, phimax) with step (phimax - phimin) / nphi, and theta = acos of a LinRange
    from cos(thetamin) to cos(thetamax) (equal-area spacing).  A number for r gives the shell of
    StructuredGrid(Spherical, rconst, phi, theta) (StructuredGrid.jl:427-454).  cos / acos are the
    platform's (Julia's own may differ in the last bit)."""
    ra = np.array([float(r)], np.float64) if np.isscalar(r) else linrange(*r)
    dphi = (phi[1] - phi[0]) / phi[2]
    pa = phi[0] + dphi * np.arange(phi[2], dtype=np.float64)
    ta = np.arccos(linrange(np.cos(theta[0]), np.cos(theta[1]), theta[2]))
    return ra, pa, ta


def flatten_spherical(ra, pa, ta):
    """Cartesian sample points of a spherical grid in column-major (r fastest) order (coordinate_grid
    for Spherical, StructuredGrid.jl:184-204, with _sph2cart: x = r sin(theta) cos(phi),
    y = r sin(theta) sin(phi), z = r cos(theta), coordinate_transformations.jl:197-203)."""
    R, P, T = np.meshgrid(ra, pa, ta, indexing="ij")
    R, P, T = (a.ravel(order="F") for a in (R, P, T))
    rs = R * np.sin(T)
    return rs * np.cos(P), rs * np.sin(P), R * np.cos(T)


def flatten_cartesian(xa, ya, za):
    X, Y, Z = np.meshgrid(xa, ya, za, indexing="ij")
    return X.ravel(order="F"), Y.ravel(order="F"), Z.ravel(order="F")


def perspective_lines(n: int, lo: float = -100.0, hi: float = 100.0, height: float = 500.0):
    """n x n column-map lines of a pinhole camera at (0, 0, height) looking down on the disc: the
    line of image point (x_i, y_j, 0) passes through it and the camera; origins (x_i, y_j, 0),
    unit directions towards the camera.  Slopes reach (hi - lo) / (2 height) at the image edges
    (0.2 with the defaults)."""
    ax = linrange(lo, hi, n)
    X, Y = np.meshgrid(ax, ax, indexing="ij")
    ox, oy = X.ravel(order="F"), Y.ravel(order="F")
    oz = np.zeros_like(ox)
    dx, dy, dz = -ox, -oy, np.full_like(ox, height)
    nrm = np.sqrt(dx * dx + dy * dy + dz * dz)
    return (ox, oy, oz), (dx / nrm, dy / nrm, dz / nrm)


def faceon_lines(n: int, lo: float = -100.0, hi: float = 100.0):
    """n x n face-on column-map lines: origins (x_i, y_j, 0), direction (0, 0, 1)."""
    ax = linrange(lo, hi, n)
    X, Y = np.meshgrid(ax, ax, indexing="ij")
    ox, oy = X.ravel(order="F"), Y.ravel(order="F")
    oz = np.zeros_like(ox)
    return (ox, oy, oz), (np.zeros_like(ox), np.zeros_like(ox), np.ones_like(ox))
