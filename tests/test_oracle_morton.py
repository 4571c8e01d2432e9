"""Pins the oracle's Morton encoding, stable sort and de-duplication (MortonCode.jl).

Checks are independent of the oracle's bit-spreading masks: codes are rebuilt by an
explicit bit-by-bit interleave, the sort by numpy's stable argsort, the de-dup by the
reference's serial loop.
"""
import numpy as np
import pytest

from oracle import oracle as O


def interleave(ix, iy, iz):
    """x bit k -> code bit 3k+2, y -> 3k+1, z -> 3k (MortonCode.jl:199 layout)."""
    c = 0
    for k in range(21):
        c |= ((ix >> k) & 1) << (3 * k + 2)
        c |= ((iy >> k) & 1) << (3 * k + 1)
        c |= ((iz >> k) & 1) << (3 * k)
    return c


def test_square_fixture_codes():
    # LBVH query fixture of neighbor_search_tests.jl:319-325: x, y in {0, 1}, z = 0 (flat)
    x = np.array([0.0, 1.0, 0.0, 1.0])
    y = np.array([0.0, 0.0, 1.0, 1.0])
    z = np.zeros(4)
    assert list(O.quantize_axis(x)) == [0, 2097151, 0, 2097151]
    assert list(O.quantize_axis(z)) == [1048575] * 4        # flat axis: floor(2097151 * 0.5)
    codes = O.morton_codes(x, y, z)
    expect = [interleave(ix, iy, 1048575) for ix, iy in [(0, 0), (2097151, 0), (0, 2097151), (2097151, 2097151)]]
    assert [int(c) for c in codes] == expect
    # hand values: all-ones x and y bits with z = 0b0111...1 (bit 20 clear)
    assert expect[3] == (0x7FFFFFFFFFFFFFFF & ~(1 << 60))


def test_expand_compact_roundtrip():
    rng = np.random.default_rng(1)
    for v in list(rng.integers(0, 1 << 21, 200)) + [0, 1, (1 << 21) - 1]:
        v = int(v)
        e = O.expand_bits3(v)
        assert e == interleave(0, 0, v)
        assert O.compact_bits3(e) == v
    for _ in range(100):
        ix, iy, iz = (int(a) for a in rng.integers(0, 1 << 21, 3))
        assert O.encode3(ix, iy, iz) == interleave(ix, iy, iz)


def test_quantize_spot_values():
    v = np.array([0.0, 0.25, 0.5, 1.0 - 2.0 ** -53, 1.0])
    q = O.quantize_axis(v)
    assert list(q) == [0, 524287, 1048575, 2097150, 2097151]
    # scaled, shifted axis: (v - vmin) * inv(vmax - vmin) with two roundings
    w = 3.0 + 7.0 * v
    assert list(O.quantize_axis(w)) == [int(np.floor(2097151.0 * ((a - 3.0) * (1.0 / 7.0)))) for a in w]


@pytest.mark.parametrize("n,dups", [(1, 0), (2, 1), (50, 0), (500, 200), (4096, 3000)])
def test_sort_and_dedup(n, dups):
    rng = np.random.default_rng(n)
    x, y, z = rng.random(n), rng.random(n), rng.random(n)
    if dups:
        src = rng.integers(0, max(1, n - dups), dups)
        x[:dups], y[:dups], z[:dups] = x[src], y[src], z[src]
    raw = O.morton_codes(x, y, z)
    perm = np.argsort(raw, kind="stable")
    ded = [int(c) for c in raw[perm]]
    for i in range(1, n):                      # MortonCode.jl:189-193
        if ded[i] <= ded[i - 1]:
            ded[i] = ded[i - 1] + 1
    cols = [x.copy(), y.copy(), z.copy(), np.ones(n), np.full(n, 0.1), np.ones(n)]
    t = O.LBVH(*cols)
    assert np.array_equal(t.order - 1, perm)
    assert [int(c) for c in t.codes] == ded
    assert bool(np.all(t.codes[1:] > t.codes[:-1]))
    # caller columns are permuted in place (InterpolationInput.jl:211-219)
    assert np.array_equal(cols[0], x[perm]) and np.array_equal(cols[2], z[perm])


def test_identical_points_codes_strictly_increase():
    n = 16
    t = O.LBVH(np.full(n, 0.5), np.full(n, 0.5), np.full(n, 0.5), np.ones(n), np.full(n, 0.5), np.ones(n))
    c0 = interleave(1048575, 1048575, 1048575)
    assert [int(c) for c in t.codes] == [c0 + i for i in range(n)]
    assert np.array_equal(t.order, np.arange(1, n + 1))
