"""Pins the oracle's interpolated fields (general_interpolation.jl, line_integrated_*.jl).

* O(N) brute-force evaluators over all particles in sorted order, restating the
  reference's test evaluators (interpolation_test_common.jl:72-335): density, quantity,
  gradient, divergence, curl, line-integrated quantity; tolerances as the reference's
  (interpolation_tests.jl:242-297);
* analytic linear fields q = x+y+z, v = (x,y,z) on a 12^3 uniform cloud with eta = 1.2
  and variable h (interpolation_analytic_test_common.jl:12-52) with the reference's
  mean-absolute-error bounds (scalar 2e-2, gradient 6e-2, divergence 1e-1, curl 5e-2);
* uniform field: div = curl = 0 (interpolation_tests.jl:304-335);
* empty support: every output NaN (interpolation_tests.jl:51-116);
* the 4-particle hand fixture (grid_interpolation_test_common.jl:17-41).
"""
import numpy as np
import pytest

from oracle import oracle as O

CAT = dict(scalar_slots=[0], scalar_shepard=[1], grad_slots=[0], div_slots=[1, 2, 3], curl_slots=[1, 2, 3])


def brute(t, cols, kernel, strategy, p, ha):
    """Brute-force scalar / gradient / divergence / curl of columns (q, vx, vy, vz)."""
    lx, ly, lz, lh = t.leaf
    m, rho = cols[3], cols[5]
    q, vx, vy, vz = cols[6:10]
    K = O.kernel_support(kernel)
    S1 = A = 0.0
    gf = np.zeros(3); gb = np.zeros(3)
    df = 0.0; db = np.zeros(3); V = np.zeros(3)
    cf = np.zeros(3)
    for b in range(lx.size):
        dx, dy, dz = p[0] - lx[b], p[1] - ly[b], p[2] - lz[b]
        d2 = ((0.0 + dx * dx) + dy * dy) + dz * dz
        hb = lh[b]
        if strategy == "gather":
            hsel = [ha]
            rad = K * ha
        elif strategy == "scatter":
            hsel = [hb]
            rad = K * hb
        else:
            hsel = [ha, hb]
            rad = K * max(ha, hb)
        if not d2 <= rad * rad:
            continue
        r = np.sqrt(d2)
        W = np.mean([O.W(kernel, r, hh) for hh in hsel])
        g = np.mean([O.gradW(kernel, dx, dy, dz, hh) for hh in hsel], axis=0)
        w = m[b] * W / rho[b]
        gw = m[b] / rho[b] * g
        S1 += w
        A += q[b] * w
        gf += q[b] * gw
        gb += gw
        vb = np.array([vx[b], vy[b], vz[b]])
        V += vb * w
        df += vb @ gw
        cf += np.cross(vb, gw)
    if S1 == 0.0:
        return np.full(8, np.nan)
    At = A / S1
    Vt = V / S1
    return np.concatenate([[At], gf - At * gb, [df - Vt @ gb], -(cf - np.cross(Vt, gb))])


def random_cloud(n, seed):
    rng = np.random.default_rng(seed)
    x, y, z = rng.random(n), rng.random(n), rng.random(n)
    m = rng.uniform(0.5, 1.5, n) / n
    h = rng.uniform(0.08, 0.2, n)
    rho = rng.uniform(0.8, 1.2, n)
    quant = [rng.normal(size=n), rng.normal(size=n), rng.normal(size=n), rng.normal(size=n)]
    return [x, y, z, m, h, rho], quant


@pytest.mark.parametrize("strategy", ["gather", "scatter", "symmetric"])
@pytest.mark.parametrize("kernel", ["M4", "C4"])
def test_fields_match_bruteforce(strategy, kernel):
    cols, quant = random_cloud(400, 1)
    t = O.LBVH(*cols, quant=quant)
    allc = cols + quant
    rng = np.random.default_rng(8)
    pts = rng.uniform(0.1, 0.9, (40, 3))
    out = t.point_samples(kernel, strategy, CAT, pts[:, 0], pts[:, 1], pts[:, 2])
    _, ha = t.nearest(pts[:, 0], pts[:, 1], pts[:, 2])
    for i in range(len(pts)):
        ref = brute(t, allc, kernel, strategy, pts[i], ha[i])
        got = np.array([o[i] for o in out])
        assert np.allclose(got, ref, rtol=1e-8, atol=5e-10 * np.abs(ref).max()), (i, got, ref)


def uniform_cloud(nx, eta=1.2, variable_h=True):
    dx = 1.0 / nx
    c = np.arange(dx / 2, 1.0, dx)
    X, Y, Z = np.meshgrid(c, c, c, indexing="ij")
    x, y, z = X.ravel(), Y.ravel(), Z.ravel()
    n = x.size
    h = eta * dx * ((1.0 + 0.18 * (x - 0.5) - 0.10 * (y - 0.5) + 0.08 * (z - 0.5)) if variable_h else np.ones(n))
    return [x.copy(), y.copy(), z.copy(), np.full(n, dx**3), h, np.ones(n)], [x + y + z, x.copy(), y.copy(), z.copy()]


@pytest.mark.parametrize("strategy", ["gather", "scatter", "symmetric"])
def test_analytic_linear_field(strategy):
    cols, quant = uniform_cloud(12)
    t = O.LBVH(*cols, quant=quant)
    rng = np.random.default_rng(0xC0FFEE)
    margin = 1.5 * 2.0 * 1.2 / 12
    p = rng.uniform(margin, 1 - margin, (64, 3))
    out = t.point_samples("M4", strategy, CAT, p[:, 0], p[:, 1], p[:, 2])
    s, gx, gy, gz, div, cx, cy, cz = out
    assert np.mean(np.abs(s - p.sum(axis=1))) < 2e-2
    for g in (gx, gy, gz):
        assert np.mean(np.abs(g - 1.0)) < 6e-2
    assert np.mean(np.abs(div - 3.0)) < 1e-1
    for cc in (cx, cy, cz):
        assert np.mean(np.abs(cc)) < 5e-2


def test_uniform_field_div_curl_zero():
    cols, quant = random_cloud(300, 3)
    quant = [np.full(300, 2.0), np.full(300, 0.7), np.full(300, -1.3), np.full(300, 0.4)]
    t = O.LBVH(*cols, quant=quant)
    rng = np.random.default_rng(5)
    p = rng.uniform(0.2, 0.8, (30, 3))
    s, gx, gy, gz, div, cx, cy, cz = t.point_samples("M4", "symmetric", CAT, p[:, 0], p[:, 1], p[:, 2])
    assert np.allclose(s, 2.0, rtol=1e-12)
    for v in (gx, gy, gz, div, cx, cy, cz):
        assert np.all(np.abs(v) < 1e-12 * 200)


def test_empty_support_is_nan():
    cols, quant = random_cloud(50, 4)
    t = O.LBVH(*cols, quant=quant)
    p = np.array([[10.0, 10.0, 10.0]])
    out = t.point_samples("M4", "scatter", CAT, p[:, 0], p[:, 1], p[:, 2])
    assert all(np.isnan(o[0]) for o in out)
    o = [np.array([10.0]), np.array([10.0]), np.array([0.0])]
    d = [np.array([0.0]), np.array([0.0]), np.array([1.0])]
    lo = t.line_samples("M4", "scatter", [0], [0], o, d)
    assert np.isnan(lo[0][0])


def test_hand_fixture_divergence():
    # make_grid_interpolation_fixture (grid_interpolation_test_common.jl:17-41)
    x = np.array([0.15, 0.35, 0.55, 0.75]); y = np.array([0.20, 0.60, 0.25, 0.70])
    z = np.array([0.25, 0.45, 0.80, 0.30]); h = np.array([0.28, 0.24, 0.26, 0.22])
    rho = np.array([1.0, 0.95, 1.05, 1.1]); m = np.array([0.4, 0.45, 0.42, 0.38])
    temp = np.array([10.0, 11.0, 13.0, 12.0]); vx = np.array([0.1, -0.2, 0.3, -0.1])
    vy = np.array([0.0, 0.25, -0.15, 0.2]); vz = np.array([-0.05, 0.1, 0.2, -0.1])
    cols = [x, y, z, m, h, rho]
    t = O.LBVH(*cols, quant=[temp, vx, vy, vz])
    p = np.array([[0.20, 0.25, 0.30], [0.40, 0.55, 0.50], [0.65, 0.50, 0.35]])
    cat = dict(scalar_slots=[0], scalar_shepard=[1], div_slots=[1, 2, 3])
    for strategy in ("gather", "scatter", "symmetric"):
        out = t.point_samples("M4", strategy, cat, p[:, 0], p[:, 1], p[:, 2])
        _, ha = t.nearest(p[:, 0], p[:, 1], p[:, 2])
        allc = cols + [temp, vx, vy, vz]
        for i in range(3):
            ref = brute(t, allc, "M4", strategy, p[i], ha[i])
            got = np.array([out[0][i], out[1][i]])
            expect = np.array([ref[0], ref[4]])
            assert np.allclose(got, expect, rtol=1e-10, atol=1e-12, equal_nan=True)


def test_line_fields_match_bruteforce():
    cols, quant = random_cloud(300, 6)
    t = O.LBVH(*cols, quant=quant)
    lx, ly, lz, lh = t.leaf
    m, rho = cols[3], cols[5]
    q = quant[0]
    inv = 1 / np.sqrt(2.0)
    o = [np.array([0.2, 0.4, 0.55]), np.array([0.25, 0.35, 0.45]), np.array([0.0, 0.1, 0.2])]
    d = [np.array([0.0, inv, 1.0]), np.array([0.0, 0.0, 0.0]), np.array([1.0, inv, 0.0])]
    out = t.line_samples("C4", "scatter", [0, 0], [1, 0], o, d)
    for i in range(3):
        num = den = 0.0
        for b in range(lx.size):
            D = np.array([lx[b] - o[0][i], ly[b] - o[1][i], lz[b] - o[2][i]])
            dd = np.array([d[0][i], d[1][i], d[2][i]])
            d2 = D @ D - (D @ dd) ** 2
            if d2 <= (2.0 * lh[b]) ** 2:
                W = O.line_W("C4", np.sqrt(max(d2, 0.0)), lh[b])
                w = m[b] * W / rho[b]
                num += q[b] * w
                den += w
        assert np.isclose(out[0][i], num / den, rtol=1e-10)
        assert np.isclose(out[1][i], num, rtol=1e-10)
