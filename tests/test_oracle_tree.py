"""Pins the oracle's radix tree, escape links and refit (BinaryRadixTree.jl, LinearBVH.jl).

The checks use definitions, not the oracle's construction steps:
  * every internal node covers a contiguous leaf range, its two children partition that
    range, and the split sits where the first bit differing between the range's first and
    last code flips (the defining property of a binary radix tree, Karras 2012);
  * internal node i's range has leaf i as one end (Karras' node numbering, which fixes the ids:
    a consistent relabelling of the internal nodes fails it);
  * escape[id] is the DFS-preorder successor outside id's subtree, computed by explicit
    recursion (as the reference's _build_escape_stackless!, BinaryRadixTree.jl:283-312);
  * node boxes / hmax equal brute-force min/max over the covered leaves.
Coordinates follow the reference's prime-hash generator (neighbor_search_tests.jl:49-61).
"""
import numpy as np
import pytest

from oracle import oracle as O

PRIMES = (37, 61, 97, 131, 197, 263)


def prime_coords(n, offset, dim=3):
    return [np.array([float((PRIMES[d] * i + PRIMES[d + dim] * offset) % 1021) / 1021 for i in range(1, n + 1)])
            for d in range(dim)]


def leaf_range(t, node):
    n = t.n
    if node >= n:
        k = node - (n - 1)
        return k, k
    lo, _ = leaf_range(t, int(t.left[node - 1]))
    _, hi = leaf_range(t, int(t.right[node - 1]))
    return lo, hi


def check_tree(t):
    n = t.n
    codes = [int(c) for c in t.codes]
    assert t.root == (1 if n >= 2 else 0)
    assert len(t.left) == 2 * n - 1
    if n == 1:
        assert t.escape[0] == 0 and t.parent[0] == 0
        return
    assert t.parent[0] == 0
    assert np.count_nonzero(t.parent == 0) == 1
    for i in range(1, n):
        l, r = int(t.left[i - 1]), int(t.right[i - 1])
        assert t.parent[l - 1] == i and t.parent[r - 1] == i
        lo, mid = leaf_range(t, l)
        mid2, hi = leaf_range(t, r)
        assert mid2 == mid + 1
        # Karras id assignment: internal node i's leaf range has leaf i as one end
        assert lo == i or hi == i, (i, lo, hi)
        # and the right child is the escape (DFS successor) of the left child
        assert int(t.escape[l - 1]) == r
        a, b = codes[lo - 1], codes[hi - 1]
        bit = 63 - (a ^ b).bit_length() + 1          # leading zeros of a ^ b
        pos = 63 - bit                               # index of the first differing bit
        assert all(((codes[k - 1] >> pos) & 1) == 0 for k in range(lo, mid + 1))
        assert all(((codes[k - 1] >> pos) & 1) == 1 for k in range(mid + 1, hi + 1))
    # escape = DFS-preorder successor outside the subtree
    esc = {}

    def visit(node, after):
        esc[node] = after
        if node < n:
            l, r = int(t.left[node - 1]), int(t.right[node - 1])
            visit(l, r)
            visit(r, after)

    visit(1, 0)
    assert all(int(t.escape[i - 1]) == esc[i] for i in range(1, 2 * n))


@pytest.mark.parametrize("offset", range(5))
def test_brt_invariants_prime_hash(offset):
    for n in range(1, 21):
        x, y, z = prime_coords(n, offset)
        t = O.LBVH(x, y, z, np.ones(n), np.full(n, 0.1), np.ones(n))
        check_tree(t)


@pytest.mark.parametrize("n", [1, 2, 8, 16, 257])
def test_brt_identical_codes(n):
    t = O.LBVH(np.full(n, 0.5), np.full(n, 0.5), np.full(n, 0.5), np.ones(n), np.full(n, 0.5), np.ones(n))
    check_tree(t)


def test_brt_random_large():
    rng = np.random.default_rng(7)
    n = 3000
    t = O.LBVH(rng.random(n), rng.random(n) ** 3, rng.normal(size=n), np.ones(n), rng.random(n), np.ones(n))
    check_tree(t)


@pytest.mark.parametrize("n", [2, 8, 32, 500])
def test_refit_matches_bruteforce(n):
    rng = np.random.default_rng(n)
    x, y, z = rng.random(n), rng.random(n), rng.normal(size=n)
    h = rng.random(n)
    t = O.LBVH(x.copy(), y.copy(), z.copy(), np.ones(n), h.copy(), np.ones(n))
    lx, ly, lz, lh = t.leaf
    for i in range(1, n):
        lo, hi = leaf_range(t, i)
        s = slice(lo - 1, hi)
        for k, arr in enumerate((lx, ly, lz)):
            assert t.node_min[k][i - 1] == arr[s].min()
            assert t.node_max[k][i - 1] == arr[s].max()
        assert t.node_hmax[i - 1] == lh[s].max()


def test_leaf_arrays_are_sorted_columns():
    n = 64
    rng = np.random.default_rng(0)
    x, y, z = rng.random(n), rng.random(n), rng.random(n)
    t = O.LBVH(x.copy(), y.copy(), z.copy(), np.ones(n), np.full(n, 0.2), np.ones(n))
    order = t.order - 1
    assert np.array_equal(t.leaf[0], x[order]) and np.array_equal(t.leaf[2], z[order])
