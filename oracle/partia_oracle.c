/*
 * partia_oracle.c — CPU FP64 restatement of Partia.jl's LBVH + SPH interpolation path.
 *
 * TEST INFRASTRUCTURE: the checker for the HIP library and the CPU baseline timed by
 * bench.py.  It is never linked into, called by, or used as a fallback of the product.
 *
 * Parallel structure mirrors the reference: OpenMP where Julia uses `Threads.@threads`
 * (quantise, encode, BRT child/adjacency/escape, refit, per-sample loops) and serial
 * where the reference is serial (bbox normalisation, sortperm, permute!, de-dup scan,
 * leaf escapes).
 *
 * How each part is pinned (tests/test_oracle_*.py):
 *   Morton quantise/encode ... hand-computed codes of the reference's 4-point square
 *                              fixture, decode∘encode round trip, spot values.
 *   sort + de-dup ............ brute-force stable sort + closed-form max-scan.
 *   BRT + escape ............. the reference's structural invariants
 *                              (neighbor_search_tests.jl:186-256) and an independent
 *                              DFS-preorder escape (BinaryRadixTree.jl:283-312).
 *   refit .................... brute-force subtree min/max/hmax.
 *   traversal / nearest ...... brute-force predicate sets (O(N*P)) with the reference
 *                              operation order.
 *   kernels .................. closed forms, sigma values, finite differences
 *                              (kernel_function_tests.jl:42-195).
 *   line LUT ................. sampled entries of the reference's own table
 *                              (tests/golden/line_lut_sample.npz) and I_M4(0)=3/2.
 *   fields ................... brute-force evaluators (interpolation_test_common.jl),
 *                              hand fixtures, analytic linear fields, uniform-field
 *                              div = curl = 0.
 *   orc_jl_pow ............... bit-level parity with Julia's `^(Float64,Integer)` is
 *                              unpinned (no Julia here); pinned only to 1e-15 relative
 *                              against exact integer powers.
 */
#include "partia_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define PRAGMA_OMP(x) _Pragma(#x)

static inline int clz64(uint64_t x) { return x ? __builtin_clzll(x) : 64; }

/* ===================================================================================
 * Morton codes — NS/BinaryRadixTree/MortonCode.jl
 * =================================================================================== */

/* _expand_bits3D(::UInt64), MortonCode.jl:209-217 (21 -> 63 bit spread). */
uint64_t orc_expand_bits3(uint64_t x) {
    x = (x | (x << 32)) & 0x1f00000000ffffULL;
    x = (x | (x << 16)) & 0x1f0000ff0000ffULL;
    x = (x | (x << 8)) & 0x100f00f00f00f00fULL;
    x = (x | (x << 4)) & 0x10c30c30c30c30c3ULL;
    x = (x | (x << 2)) & 0x1249249249249249ULL;
    return x;
}

/* _compact_bits3D(::UInt64), MortonCode.jl:228-236. */
uint64_t orc_compact_bits3(uint64_t x) {
    x &= 0x1249249249249249ULL;
    x = (x ^ (x >> 2)) & 0x10c30c30c30c30c3ULL;
    x = (x ^ (x >> 4)) & 0x100f00f00f00f00fULL;
    x = (x ^ (x >> 8)) & 0x1f0000ff0000ffULL;
    x = (x ^ (x >> 16)) & 0x1f00000000ffffULL;
    x = (x ^ (x >> 32)) & 0x1fffffULL;
    return x;
}

/* _encode_morton_code3D, MortonCode.jl:275-280: x is the most significant of each triple. */
uint64_t orc_encode3(uint64_t ix, uint64_t iy, uint64_t iz) {
    return (orc_expand_bits3(ix) << 2) | (orc_expand_bits3(iy) << 1) | orc_expand_bits3(iz);
}

/* _normalize_vector (MortonCode.jl:443-456, serial) followed by the quantisation of
 * _quantize_coords (MortonCode.jl:473-498): f = (v - vmin) * inv(vmax - vmin), or 0.5
 * when the axis is flat; i = UInt64(floor(scale * f)) with scale = exp2(21) - 1
 * (_axis_scale, MortonCode.jl:431-434). */
void orc_quantize_axis(int64_t n, const double *v, uint64_t *out) {
    double vmin = v[0], vmax = v[0];
    for (int64_t i = 1; i < n; ++i) {   /* Base.minimum / maximum */
        if (v[i] < vmin) vmin = v[i];
        if (v[i] > vmax) vmax = v[i];
    }
    const double scale = 2097151.0;     /* exp2(21.0) - 1.0, exact */
    const double dv = vmax - vmin;
    if (dv == 0.0) {
        uint64_t c = (uint64_t)floor(scale * 0.5);
        for (int64_t i = 0; i < n; ++i) out[i] = c;
        return;
    }
    const double inv = 1.0 / dv;
    double *f = (double *)malloc(sizeof(double) * (size_t)n);
    for (int64_t i = 0; i < n; ++i) f[i] = (v[i] - vmin) * inv;     /* serial normalise */
    PRAGMA_OMP(omp parallel for schedule(static))
    for (int64_t i = 0; i < n; ++i) out[i] = (uint64_t)floor(scale * f[i]);
    free(f);
}

void orc_morton_codes(int64_t n, const double *x, const double *y, const double *z,
                      uint64_t *codes) {
    uint64_t *ix = (uint64_t *)malloc(8 * (size_t)n), *iy = (uint64_t *)malloc(8 * (size_t)n),
             *iz = (uint64_t *)malloc(8 * (size_t)n);
    orc_quantize_axis(n, x, ix);
    orc_quantize_axis(n, y, iy);
    orc_quantize_axis(n, z, iz);
    PRAGMA_OMP(omp parallel for schedule(static))          /* MortonCode.jl:301-303 */
    for (int64_t i = 0; i < n; ++i) codes[i] = orc_encode3(ix[i], iy[i], iz[i]);
    free(ix); free(iy); free(iz);
}

/* sortperm(codes; alg=QuickSort) with Base.Order.Perm (MortonCode.jl:182): Perm breaks
 * ties by index, so the permutation is the unique order of (code, index) pairs, i.e. the
 * stable order.  A serial quicksort on (code, index) keys reproduces it. */
static inline int key_less(const uint64_t *c, int64_t a, int64_t b) {
    return c[a] < c[b] || (c[a] == c[b] && a < b);
}
static void insertion_sort(const uint64_t *c, int64_t *p, int64_t lo, int64_t hi) {
    for (int64_t i = lo + 1; i <= hi; ++i) {
        int64_t v = p[i], j = i - 1;
        while (j >= lo && key_less(c, v, p[j])) { p[j + 1] = p[j]; --j; }
        p[j + 1] = v;
    }
}
static void quicksort_perm(const uint64_t *c, int64_t *p, int64_t lo, int64_t hi) {
    while (hi - lo > 16) {
        int64_t mid = lo + (hi - lo) / 2;
        /* median of three */
        if (key_less(c, p[mid], p[lo])) { int64_t t = p[mid]; p[mid] = p[lo]; p[lo] = t; }
        if (key_less(c, p[hi], p[lo])) { int64_t t = p[hi]; p[hi] = p[lo]; p[lo] = t; }
        if (key_less(c, p[hi], p[mid])) { int64_t t = p[hi]; p[hi] = p[mid]; p[mid] = t; }
        int64_t pivot = p[mid];
        int64_t i = lo, j = hi;
        while (i <= j) {
            while (key_less(c, p[i], pivot)) ++i;
            while (key_less(c, pivot, p[j])) --j;
            if (i <= j) { int64_t t = p[i]; p[i] = p[j]; p[j] = t; ++i; --j; }
        }
        if (j - lo < hi - i) { quicksort_perm(c, p, lo, j); lo = i; }
        else { quicksort_perm(c, p, i, hi); hi = j; }
    }
    insertion_sort(c, p, lo, hi);
}

static void permute_f64(double *a, const int64_t *p, int64_t n, double *tmp) {
    for (int64_t i = 0; i < n; ++i) tmp[i] = a[p[i]];               /* Base.permute! (serial) */
    memcpy(a, tmp, sizeof(double) * (size_t)n);
}

/* ===================================================================================
 * Binary radix tree — NS/BinaryRadixTree/BinaryRadixTree.jl (1-based indices kept)
 * =================================================================================== */

/* _longest_common_prefix_length(codes, i, j), MortonCode.jl:380-387 (1-based i, j). */
static inline int64_t delta(const uint64_t *codes1, int64_t i, int64_t j) {
    uint64_t a = codes1[i], b = codes1[j];
    if (a == b) return clz64((uint64_t)i ^ (uint64_t)j);
    return clz64(a ^ b);
}

/* _range_direction / _find_range, BinaryRadixTree.jl:140-179; codes1 is 1-based. */
void orc_find_range(const uint64_t *codes1, int64_t n, int64_t i, int64_t *lo, int64_t *hi) {
    int64_t dl = (i > 1) ? delta(codes1, i, i - 1) : -1;
    int64_t dr = (i < n) ? delta(codes1, i, i + 1) : -1;
    int64_t d = (dr - dl > 0) ? 1 : ((dr - dl < 0) ? -1 : 1);
    int64_t dmin = (1 <= i - d && i - d <= n) ? delta(codes1, i, i - d) : -1;
    int64_t lmax = 2;
    for (;;) {
        int64_t j = i + lmax * d;
        if (j < 1 || j > n) break;
        if (delta(codes1, i, j) <= dmin) break;
        lmax *= 2;
    }
    int64_t l = 0;
    for (int64_t t = lmax >> 1; t > 0; t >>= 1) {
        int64_t j = i + (l + t) * d;
        if (j >= 1 && j <= n && delta(codes1, i, j) > dmin) l += t;
    }
    int64_t j = i + l * d;
    *lo = i < j ? i : j;
    *hi = i < j ? j : i;
}

/* _split_position, BinaryRadixTree.jl:181-202. */
int orc_split_position(const uint64_t *codes1, int64_t first, int64_t last) {
    if (codes1[first] == codes1[last]) {
        int64_t s = (first + last) >> 1;
        if (s < first) s = first;
        if (s > last - 1) s = last - 1;
        return (int)s;
    }
    int64_t prefix = delta(codes1, first, last);
    int64_t split = first, step = last - first;
    while (step > 1) {
        step = (step + 1) >> 1;
        int64_t ns = split + step;
        if (ns < last && delta(codes1, first, ns) > prefix) split = ns;
    }
    return (int)split;
}

/* BinaryRadixTree(enc), BinaryRadixTree.jl:81-132 with _build_child! (:209-243),
 * _build_adjacent! (:204-207), internal _build_escape! (:245-260) and the serial leaf
 * _build_escape! loop (:262-281). */
static void build_brt(orc_lbvh *t) {
    const int64_t n = t->n;
    const uint64_t *codes1 = t->codes - 1;      /* 1-based view */
    const int64_t ntot = 2 * n - 1, off = n - 1;
    t->left = (int32_t *)calloc((size_t)(ntot + 1), 4);
    t->right = (int32_t *)calloc((size_t)(ntot + 1), 4);
    t->escape = (int32_t *)calloc((size_t)(ntot + 1), 4);
    t->parent = (int32_t *)calloc((size_t)(ntot + 1), 4);
    t->root = n >= 2 ? 1 : 0;
    if (n < 2) return;
    int32_t *range_hi = (int32_t *)malloc(4 * (size_t)(n + 1));
    int32_t *adj = (int32_t *)malloc(4 * (size_t)(n + 1));
    PRAGMA_OMP(omp parallel for schedule(static))
    for (int64_t i = 1; i <= n - 1; ++i) {
        int64_t lo, hi;
        orc_find_range(codes1, n, i, &lo, &hi);
        range_hi[i] = (int32_t)hi;
        int64_t s = orc_split_position(codes1, lo, hi);
        int64_t l = (s == lo) ? lo + off : s;
        int64_t r = (s + 1 == hi) ? hi + off : s + 1;
        t->left[i] = (int32_t)l; t->parent[l] = (int32_t)i;
        t->right[i] = (int32_t)r; t->parent[r] = (int32_t)i;
    }
    PRAGMA_OMP(omp parallel for schedule(static))
    for (int64_t i = 1; i <= n - 1; ++i) adj[i] = (int32_t)delta(codes1, i, i + 1);
    adj[n] = 2147483647;
    PRAGMA_OMP(omp parallel for schedule(static))
    for (int64_t i = 1; i <= n - 1; ++i) {
        int64_t hi = range_hi[i];
        if (hi == n) t->escape[i] = 0;
        else if (hi == n - 1) t->escape[i] = (int32_t)(off + n);
        else {
            int64_t next = hi + 1;
            t->escape[i] = (adj[next] < adj[next - 1]) ? (int32_t)(off + next) : (int32_t)next;
        }
    }
    for (int64_t leaf = n; leaf <= ntot; ++leaf) {     /* serial, as in the reference */
        int32_t p = t->parent[leaf];
        if (p == 0) { t->escape[leaf] = 0; continue; }
        t->escape[leaf] = (t->left[p] == (int32_t)leaf) ? t->right[p] : t->escape[p];
    }
    free(range_hi); free(adj);
}

/* LinearBVH(enc, brt), LinearBVH.jl:55-87, with _build_internal_aabb! (:98-156): one
 * climber per leaf; the second arrival at an internal node combines its children. */
static void refit(orc_lbvh *t) {
    const int64_t n = t->n;
    t->nminx = (double *)calloc((size_t)n, 8); t->nminy = (double *)calloc((size_t)n, 8);
    t->nminz = (double *)calloc((size_t)n, 8); t->nmaxx = (double *)calloc((size_t)n, 8);
    t->nmaxy = (double *)calloc((size_t)n, 8); t->nmaxz = (double *)calloc((size_t)n, 8);
    t->nhmax = (double *)calloc((size_t)n, 8);
    if (n < 2) return;
    uint32_t *visited = (uint32_t *)calloc((size_t)n, 4);
    const int64_t off = n - 1;
    PRAGMA_OMP(omp parallel for schedule(static))
    for (int64_t start = n; start <= 2 * n - 1; ++start) {
        int32_t p = t->parent[start];
        while (p != 0) {
            uint32_t nv = __atomic_add_fetch(&visited[p], 1u, __ATOMIC_SEQ_CST);
            if (nv != 2u) break;
            int32_t l = t->left[p], r = t->right[p];
            int ll = l > off, rl = r > off;
#define CH(arr_leaf, arr_node, c, isleaf) ((isleaf) ? (arr_leaf)[(c) - off - 1] : (arr_node)[c])
            double hl = CH(t->lh, t->nhmax, l, ll), hr = CH(t->lh, t->nhmax, r, rl);
            t->nhmax[p] = hl > hr ? hl : hr;
            double a, b;
            a = CH(t->lx, t->nminx, l, ll); b = CH(t->lx, t->nminx, r, rl); t->nminx[p] = a < b ? a : b;
            a = CH(t->lx, t->nmaxx, l, ll); b = CH(t->lx, t->nmaxx, r, rl); t->nmaxx[p] = a > b ? a : b;
            a = CH(t->ly, t->nminy, l, ll); b = CH(t->ly, t->nminy, r, rl); t->nminy[p] = a < b ? a : b;
            a = CH(t->ly, t->nmaxy, l, ll); b = CH(t->ly, t->nmaxy, r, rl); t->nmaxy[p] = a > b ? a : b;
            a = CH(t->lz, t->nminz, l, ll); b = CH(t->lz, t->nminz, r, rl); t->nminz[p] = a < b ? a : b;
            a = CH(t->lz, t->nmaxz, l, ll); b = CH(t->lz, t->nmaxz, r, rl); t->nmaxz[p] = a > b ? a : b;
#undef CH
            __atomic_thread_fence(__ATOMIC_SEQ_CST);
            p = t->parent[p];
        }
    }
    free(visited);
}

/* LinearBVH!(input, Val(3)), InterpolationInput.jl:203-224: MortonEncoding (copies,
 * quantise, encode, sort_by_morton! with de-dup), permute! of every caller column in
 * place, BinaryRadixTree, LinearBVH.  nthreads <= 0 keeps the OpenMP default. */
orc_lbvh *orc_linear_bvh(int64_t n, double *x, double *y, double *z, double *m, double *h,
                         double *rho, int ncol, double **quant, int nthreads) {
    if (n < 1 || n > 1073741824LL) return NULL;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
    orc_lbvh *t = (orc_lbvh *)calloc(1, sizeof(orc_lbvh));
    t->n = n;
    t->codes = (uint64_t *)malloc(8 * (size_t)n);
    t->order = (int64_t *)malloc(8 * (size_t)n);
    orc_morton_codes(n, x, y, z, t->codes);
    int64_t *p = (int64_t *)malloc(8 * (size_t)n);
    for (int64_t i = 0; i < n; ++i) p[i] = i;
    quicksort_perm(t->codes, p, 0, n - 1);
    {
        uint64_t *tmp = (uint64_t *)malloc(8 * (size_t)n);
        for (int64_t i = 0; i < n; ++i) tmp[i] = t->codes[p[i]];
        memcpy(t->codes, tmp, 8 * (size_t)n);
        free(tmp);
    }
    for (int64_t i = 0; i < n; ++i) t->order[i] = p[i] + 1;
    for (int64_t i = 1; i < n; ++i)        /* de-dup scan, MortonCode.jl:189-193 */
        if (t->codes[i] <= t->codes[i - 1]) t->codes[i] = t->codes[i - 1] + 1;
    double *tmp = (double *)malloc(8 * (size_t)n);
    permute_f64(x, p, n, tmp); permute_f64(y, p, n, tmp); permute_f64(z, p, n, tmp);
    permute_f64(m, p, n, tmp); permute_f64(h, p, n, tmp); permute_f64(rho, p, n, tmp);
    for (int c = 0; c < ncol; ++c) permute_f64(quant[c], p, n, tmp);
    free(tmp); free(p);
    t->lx = (double *)malloc(8 * (size_t)n); memcpy(t->lx, x, 8 * (size_t)n);
    t->ly = (double *)malloc(8 * (size_t)n); memcpy(t->ly, y, 8 * (size_t)n);
    t->lz = (double *)malloc(8 * (size_t)n); memcpy(t->lz, z, 8 * (size_t)n);
    t->lh = (double *)malloc(8 * (size_t)n); memcpy(t->lh, h, 8 * (size_t)n);
    build_brt(t);
    refit(t);
    return t;
}

void orc_lbvh_free(orc_lbvh *t) {
    if (!t) return;
    free(t->codes); free(t->order); free(t->left); free(t->right); free(t->escape);
    free(t->parent); free(t->lx); free(t->ly); free(t->lz); free(t->lh);
    free(t->nminx); free(t->nminy); free(t->nminz); free(t->nmaxx); free(t->nmaxy);
    free(t->nmaxz); free(t->nhmax); free(t);
}

/* ===================================================================================
 * SPH kernels — KI/kernel_function/kernels/{M4..C6}.jl, kernel.jl
 * =================================================================================== */

/* Julia Base `^(x::Float64, n::Integer)` for the exponents used by the kernels (4..8):
 * compensated power-by-squaring (Base.Math.pow_body) with two_mul via fma and muladd
 * taken as fused.  n == 3 is x*x*x, as Base.literal_pow for Val(3). */
double orc_jl_pow(double x, int64_t n) {
    if (n == 0) return 1.0;
    if (n == 2) return x * x;
    if (n == 3) return x * x * x;
    double y = 1.0, xnlo = 0.0, ynlo = 0.0;
    while (n > 1) {
        if (n & 1) {
            double err = fma(y, xnlo, x * ynlo);
            double yh = x * y;
            double yl = fma(x, y, -yh);
            y = yh; ynlo = yl + err;
        }
        double err = x * 2.0 * xnlo;
        double xh = x * x;
        double xl = fma(x, x, -xh);
        x = xh; xnlo = xl + err;
        n >>= 1;
    }
    double err = fma(y, xnlo, x * ynlo);
    return (isfinite(x) && isfinite(err)) ? fma(x, y, err) : x * y;
}
#define P(x, n) orc_jl_pow((x), (n))

/* Kernel shapes f(q) (M4_spline.jl:9-20, M5_spline.jl:9-22, M6_spline.jl:9-22,
 * C2_Wendland.jl:9-18, C4_Wendland.jl:9-18, C6_Wendland.jl:9-18). */
double orc_kernel_f(int k, double q) {
    if (q < 0.0) return NAN;
    switch (k) {
    case ORC_M4:
        if (q < 1.0) { double a = 2.0 - q, b = 1.0 - q; return 0.25 * (a * a * a) - b * b * b; }
        if (q < 2.0) { double a = 2.0 - q; return 0.25 * (a * a * a); }
        return 0.0;
    case ORC_M5:
        if (q < 0.5) return (P(2.5 - q, 4) - 5.0 * P(1.5 - q, 4)) + 10.0 * P(0.5 - q, 4);
        if (q < 1.5) return P(2.5 - q, 4) - 5.0 * P(1.5 - q, 4);
        if (q < 2.5) return P(2.5 - q, 4);
        return 0.0;
    case ORC_M6:
        if (q < 1.0) return (P(3.0 - q, 5) - 6.0 * P(2.0 - q, 5)) + 15.0 * P(1.0 - q, 5);
        if (q < 2.0) return P(3.0 - q, 5) - 6.0 * P(2.0 - q, 5);
        if (q < 3.0) return P(3.0 - q, 5);
        return 0.0;
    case ORC_C2:
        if (q < 2.0) return P(1.0 - 0.5 * q, 4) * (2.0 * q + 1.0);
        return 0.0;
    case ORC_C4:
        if (q < 2.0) return P(1.0 - 0.5 * q, 6) * (((35.0 / 12.0) * (q * q) + 3.0 * q) + 1.0);
        return 0.0;
    case ORC_C6:
        if (q < 2.0)
            return P(1.0 - 0.5 * q, 8) *
                   (((4.0 * (q * q * q) + 6.25 * (q * q)) + 4.0 * q) + 1.0);
        return 0.0;
    }
    return NAN;
}

/* Kernel derivatives f'(q) (M4_spline.jl:22-33, C4_Wendland.jl:20-32 and siblings). */
double orc_kernel_df(int k, double q) {
    if (q < 0.0) return NAN;
    switch (k) {
    case ORC_M4:
        if (q < 1.0) { double a = 2.0 - q, b = 1.0 - q; return -0.75 * (a * a) + 3.0 * (b * b); }
        if (q < 2.0) { double a = 2.0 - q; return -0.75 * (a * a); }
        return 0.0;
    case ORC_M5:
        if (q < 0.5)
            return (-4.0 * P(2.5 - q, 3) + 20.0 * P(1.5 - q, 3)) - 40.0 * P(0.5 - q, 3);
        if (q < 1.5) return -4.0 * P(2.5 - q, 3) + 20.0 * P(1.5 - q, 3);
        if (q < 2.5) return -4.0 * P(2.5 - q, 3);
        return 0.0;
    case ORC_M6:
        if (q < 1.0)
            return (-5.0 * P(3.0 - q, 4) + 30.0 * P(2.0 - q, 4)) - 75.0 * P(1.0 - q, 4);
        if (q < 2.0) return -5.0 * P(3.0 - q, 4) + 30.0 * P(2.0 - q, 4);
        if (q < 3.0) return -5.0 * P(3.0 - q, 4);
        return 0.0;
    case ORC_C2:
        if (q < 2.0) return P(1.0 - 0.5 * q, 4) * 2.0 - P(1.0 - 0.5 * q, 3) * (4.0 * q + 2.0);
        return 0.0;
    case ORC_C4:
        if (q < 2.0)
            return P(1.0 - 0.5 * q, 6) * ((35.0 / 6.0) * q + 3.0) -
                   P(1.0 - 0.5 * q, 5) * (((35.0 / 4.0) * (q * q) + 9.0 * q) + 3.0);
        return 0.0;
    case ORC_C6:
        if (q < 2.0)
            return P(1.0 - 0.5 * q, 8) * ((12.0 * (q * q) + 12.5 * q) + 4.0) -
                   P(1.0 - 0.5 * q, 7) *
                       (((16.0 * (q * q * q) + 25.0 * (q * q)) + 16.0 * q) + 4.0);
        return 0.0;
    }
    return NAN;
}
#undef P

/* KernelFunctionnorm(K, Val(3), Float64), e.g. M4_spline.jl:97, C4_Wendland.jl:96. */
double orc_kernel_norm3(int k) {
    const double pi = 3.141592653589793;
    switch (k) {
    case ORC_M4: return 1.0 / pi;
    case ORC_M5: return 0.05 / pi;
    case ORC_M6: return 1.0 / (120.0 * pi);
    case ORC_C2: return 21.0 / (16.0 * pi);
    case ORC_C4: return 495.0 / (256.0 * pi);
    case ORC_C6: return 1365.0 / (512.0 * pi);
    }
    return NAN;
}

/* KernelFunctionValid(K, Float64), e.g. M4_spline.jl:62. */
double orc_kernel_support(int k) {
    switch (k) {
    case ORC_M4: return 2.0;
    case ORC_M5: return 2.5;
    case ORC_M6: return 3.0;
    default: return 2.0;
    }
}

/* Smoothed_kernel_function(K, r, h, Val(3)), kernel.jl:152-156. */
double orc_W(int k, double r, double h) {
    double invh = 1.0 / h;
    double inv_hD = invh * invh * invh;
    double q = r * invh;
    return inv_hD * (orc_kernel_norm3(k) * orc_kernel_f(k, q));
}

/* Smoothed_gradient_kernel_function(K, dx, dy, dz, h), kernel.jl:371-376, with the
 * dimensionless part at kernel.jl:284-296. */
void orc_gradW(int k, double dx, double dy, double dz, double h, double *g) {
    double invh = 1.0 / h;
    double inv_hDp1 = invh * invh * invh * invh;
    double r2 = dx * dx + dy * dy + dz * dz;
    double gx = 0.0, gy = 0.0, gz = 0.0;
    if (r2 != 0.0) {
        double r = sqrt(r2);
        double invr = 1.0 / r;
        double q = r / h;
        double coeff = orc_kernel_df(k, q) * orc_kernel_norm3(k) * invr;
        gx = dx * coeff; gy = dy * coeff; gz = dz * coeff;
    }
    g[0] = inv_hDp1 * gx; g[1] = inv_hDp1 * gy; g[2] = inv_hDp1 * gz;
}

/* ---- line-integrated kernel table: line_integrated_kernel_tables.jl holds, per kernel,
 * Q (4096 points on [0, support]) and Inorm = sigma_3 * I(q) with
 * I(q) = integral over the whole line of f(sqrt(q^2 + s^2)) ds.  The table is regenerated
 * here from that definition: s = q sinh(t) turns each polynomial piece into an entire
 * function of t, integrated by composite 12-point Gauss-Legendre in long double. */
static const double kBreaks[6][3] = {{1.0, 0, 0}, {0.5, 1.5, 0}, {1.0, 2.0, 0},
                                     {0, 0, 0},   {0, 0, 0},     {0, 0, 0}};
static const int kNBreaks[6] = {1, 2, 2, 0, 0, 0};

static void gauss_legendre(int npt, long double *x, long double *w) {
    for (int i = 0; i < npt; ++i) {
        long double z = cosl(3.14159265358979323846264338327950288L * (i + 0.75L) / (npt + 0.5L));
        for (int it = 0; it < 100; ++it) {
            long double p1 = 1.0L, p2 = 0.0L;
            for (int j = 1; j <= npt; ++j) {
                long double p3 = p2; p2 = p1;
                p1 = ((2.0L * j - 1.0L) * z * p2 - (j - 1.0L) * p3) / j;
            }
            long double pp = npt * (z * p1 - p2) / (z * z - 1.0L);
            long double z1 = z;
            z = z1 - p1 / pp;
            if (fabsl(z - z1) < 1e-19L) {
                x[i] = z; w[i] = 2.0L / ((1.0L - z * z) * pp * pp);
                break;
            }
            x[i] = z; w[i] = 2.0L / ((1.0L - z * z) * pp * pp);
        }
    }
}

static long double shape_ld(int k, long double q) {
    /* the kernel shapes in long double (same piecewise polynomials as orc_kernel_f) */
    switch (k) {
    case ORC_M4:
        if (q < 1) return 0.25L * powl(2 - q, 3) - powl(1 - q, 3);
        if (q < 2) return 0.25L * powl(2 - q, 3);
        return 0;
    case ORC_M5:
        if (q < 0.5L) return powl(2.5L - q, 4) - 5 * powl(1.5L - q, 4) + 10 * powl(0.5L - q, 4);
        if (q < 1.5L) return powl(2.5L - q, 4) - 5 * powl(1.5L - q, 4);
        if (q < 2.5L) return powl(2.5L - q, 4);
        return 0;
    case ORC_M6:
        if (q < 1) return powl(3 - q, 5) - 6 * powl(2 - q, 5) + 15 * powl(1 - q, 5);
        if (q < 2) return powl(3 - q, 5) - 6 * powl(2 - q, 5);
        if (q < 3) return powl(3 - q, 5);
        return 0;
    case ORC_C2: return q < 2 ? powl(1 - q / 2, 4) * (2 * q + 1) : 0;
    case ORC_C4: return q < 2 ? powl(1 - q / 2, 6) * (35.0L / 12 * q * q + 3 * q + 1) : 0;
    case ORC_C6:
        return q < 2 ? powl(1 - q / 2, 8) * (4 * q * q * q + 6.25L * q * q + 4 * q + 1) : 0;
    }
    return 0;
}

static long double line_integral_ld(int k, long double q, const long double *gx,
                                    const long double *gw, int npt) {
    const long double qmax = orc_kernel_support(k);
    if (q >= qmax) return 0;
    long double edges[8]; int ne = 0;
    long double total = 0;
    if (q == 0) {
        edges[ne++] = 0;
        for (int b = 0; b < kNBreaks[k]; ++b) edges[ne++] = kBreaks[k][b];
        edges[ne++] = qmax;
        for (int e = 0; e + 1 < ne; ++e) {
            long double a = edges[e], c = edges[e + 1];
            int nsub = 16;
            for (int s = 0; s < nsub; ++s) {
                long double lo = a + (c - a) * s / nsub, hi = a + (c - a) * (s + 1) / nsub;
                long double hm = (hi - lo) / 2, mid = (hi + lo) / 2;
                for (int i = 0; i < npt; ++i) total += gw[i] * hm * shape_ld(k, mid + hm * gx[i]);
            }
        }
        return 2 * total;
    }
    edges[ne++] = 0;
    for (int b = 0; b < kNBreaks[k]; ++b)
        if (kBreaks[k][b] > q) edges[ne++] = acoshl((long double)kBreaks[k][b] / q);
    edges[ne++] = acoshl(qmax / q);
    for (int e = 0; e + 1 < ne; ++e) {
        long double a = edges[e], c = edges[e + 1];
        int nsub = (int)ceill((c - a) / 0.2L);
        if (nsub < 1) nsub = 1;
        for (int s = 0; s < nsub; ++s) {
            long double lo = a + (c - a) * s / nsub, hi = a + (c - a) * (s + 1) / nsub;
            long double hm = (hi - lo) / 2, mid = (hi + lo) / 2;
            for (int i = 0; i < npt; ++i) {
                long double tt = mid + hm * gx[i];
                long double ch = coshl(tt);
                total += gw[i] * hm * shape_ld(k, q * ch) * q * ch;
            }
        }
    }
    return 2 * total;
}

void orc_line_table(int k, double *Q, double *Inorm) {
    long double gx[12], gw[12];
    gauss_legendre(12, gx, gw);
    const double qmax = orc_kernel_support(k);
    const long double sigma = (long double)orc_kernel_norm3(k);
    for (int i = 0; i < 4096; ++i) {
        Q[i] = (double)i * qmax / 4095.0;          /* correctly rounded (i*qmax)/4095 */
        Inorm[i] = (double)(sigma * line_integral_ld(k, (long double)Q[i], gx, gw, 12));
    }
}

typedef struct { double Q[4096]; double I[4096]; } line_table_t;
static line_table_t g_tables[6];
static int g_tables_ready[6];

static const line_table_t *line_table(int k) {
    if (!g_tables_ready[k]) {
        PRAGMA_OMP(omp critical(orc_lut))
        {
            if (!g_tables_ready[k]) {
                orc_line_table(k, g_tables[k].Q, g_tables[k].I);
                __atomic_store_n(&g_tables_ready[k], 1, __ATOMIC_RELEASE);
            }
        }
    }
    return &g_tables[k];
}

/* _lin_lut, line_integrated_kernel.jl:13-19 (1-based i clamped to 1..N-1). */
static double lin_lut(double q, const double *Q, const double *I) {
    double dq = Q[1] - Q[0];
    double idxf = q / dq + 1.0;
    int64_t i = (int64_t)(int32_t)idxf;           /* unsafe_trunc(Int32, idxf) */
    if (i < 1) i = 1;
    if (i > 4095) i = 4095;
    double t = idxf - (double)i;
    return I[i - 1] * (1.0 - t) + I[i] * t;
}

/* line_integrated_kernel_function(K, r, h), line_integrated_kernel.jl:90-94, 132-137:
 * invh * Inorm(q), i.e. a 1/h prefactor as in the reference. */
double orc_line_W(int k, double r, double h) {
    double invh = 1.0 / h;
    double q = r * invh;
    double I = (q >= orc_kernel_support(k)) ? 0.0 : lin_lut(q, line_table(k)->Q, line_table(k)->I);
    return invh * I;
}

/* ===================================================================================
 * Distances — NS/LinearBVH/LinearBVHTraversalMacros.jl:800-953
 * =================================================================================== */
static inline double d2_point_leaf(const orc_lbvh *t, const double *p, int64_t li) {
    double s = 0.0, dd;
    dd = p[0] - t->lx[li]; s += dd * dd;
    dd = p[1] - t->ly[li]; s += dd * dd;
    dd = p[2] - t->lz[li]; s += dd * dd;
    return s;
}
static inline double clampd(double p, double a, double b) {
    return p < a ? (a - p) : (p > b ? (p - b) : 0.0);
}
static inline double d2_point_aabb(const orc_lbvh *t, const double *p, int32_t id) {
    double s = 0.0, tt;
    tt = clampd(p[0], t->nminx[id], t->nmaxx[id]); s += tt * tt;
    tt = clampd(p[1], t->nminy[id], t->nmaxy[id]); s += tt * tt;
    tt = clampd(p[2], t->nminz[id], t->nmaxz[id]); s += tt * tt;
    return s;
}
static inline double d2_line_leaf(const orc_lbvh *t, const double *o, const double *d, int64_t li) {
    double D2 = 0.0, Dm = 0.0, dd;
    dd = t->lx[li] - o[0]; D2 += dd * dd; Dm += dd * d[0];
    dd = t->ly[li] - o[1]; D2 += dd * dd; Dm += dd * d[1];
    dd = t->lz[li] - o[2]; D2 += dd * dd; Dm += dd * d[2];
    return D2 - Dm * Dm;
}
static inline double jl_max(double a, double b) {
    if (a != a || b != b) return a + b;
    if (a > b) return a;
    if (b > a) return b;
    return signbit(a) ? b : a;
}
static inline double jl_min(double a, double b) {
    if (a != a || b != b) return a + b;
    if (a < b) return a;
    if (b < a) return b;
    return signbit(a) ? a : b;
}
static int line_intersects_aabb(const double *o, const double *d, const double *mn,
                                const double *mx) {
    double tmin = -INFINITY, tmax = INFINITY;
    for (int k = 0; k < 3; ++k) {
        double a = mn[k], b = mx[k];
        if (d[k] == 0.0) {
            if (o[k] < a || o[k] > b) return 0;
        } else {
            double t0 = (a - o[k]) / d[k], t1 = (b - o[k]) / d[k];
            if (t0 > t1) { double s = t0; t0 = t1; t1 = s; }
            tmin = jl_max(tmin, t0);
            tmax = jl_min(tmax, t1);
            if (tmin > tmax) return 0;
        }
    }
    return 1;
}
static double d2_line_aabb_lb(const orc_lbvh *t, const double *o, const double *d, int32_t id) {
    double mn[3] = {t->nminx[id], t->nminy[id], t->nminz[id]};
    double mx[3] = {t->nmaxx[id], t->nmaxy[id], t->nmaxz[id]};
    if (line_intersects_aabb(o, d, mn, mx)) return 0.0;
    double c[3], r2 = 0.0;
    for (int k = 0; k < 3; ++k) c[k] = (mn[k] + mx[k]) * 0.5;
    for (int k = 0; k < 3; ++k) { double hh = (mx[k] - mn[k]) * 0.5; r2 += hh * hh; }
    double D2 = 0.0, Dm = 0.0;
    for (int k = 0; k < 3; ++k) { double dd = c[k] - o[k]; D2 += dd * dd; Dm += dd * d[k]; }
    double dc2 = D2 - Dm * Dm;
    if (dc2 <= r2) return 0.0;
    double dl = sqrt(dc2) - sqrt(r2);
    return dl * dl;
}

/* ===================================================================================
 * Traversals — stackless left/escape DFS (LinearBVHTraversalMacros.jl).  Leaves are
 * visited in increasing sorted index, which is also the accumulation order.
 * The visitor returns nothing; `mode` selects the radius rule:
 *   gather:    r2 = radius2                                 (:26-108, :408-494)
 *   scatter:   r2 = (K h)^2, h = leaf_h / node_hmax          (:135-236, :539-644)
 *   symmetric: r2 = max(radius2, (K h)^2)                   (:268-371, :691-796)
 * =================================================================================== */
typedef void (*leaf_fn)(void *ctx, int64_t leaf, double d2, double hb);

static inline double radius2_rule(int strategy, double Kv, double radius2, double h) {
    if (strategy == ORC_GATHER) return radius2;
    double r = Kv * h;
    if (strategy == ORC_SCATTER) return r * r;
    return jl_max(radius2, r * r);
}

static void traverse_point(const orc_lbvh *t, const double *p, int strategy, double Kv,
                           double radius2, leaf_fn fn, void *ctx, int64_t *ni, int64_t *nl) {
    const int64_t n = t->n, off = n - 1;
    if (t->root == 0) {
        for (int64_t li = 0; li < n; ++li) {
            double hb = t->lh[li];
            double r2 = radius2_rule(strategy, Kv, radius2, hb);
            double d2 = d2_point_leaf(t, p, li);
            if (nl) ++*nl;
            if (d2 <= r2) fn(ctx, li, d2, hb);
        }
        return;
    }
    int32_t node = t->root;
    while (node != 0) {
        if (node > off) {
            int64_t li = node - off - 1;
            double hb = t->lh[li];
            double r2 = radius2_rule(strategy, Kv, radius2, hb);
            double d2 = d2_point_leaf(t, p, li);
            if (nl) ++*nl;
            if (d2 <= r2) fn(ctx, li, d2, hb);
            node = t->escape[node];
            continue;
        }
        if (ni) ++*ni;
        double r2 = radius2_rule(strategy, Kv, radius2, t->nhmax[node]);
        double d2n = d2_point_aabb(t, p, node);
        node = (d2n > r2) ? t->escape[node] : t->left[node];
    }
}

static void traverse_line(const orc_lbvh *t, const double *o, const double *dvec, int strategy,
                          double Kv, double radius2, leaf_fn fn, void *ctx, int64_t *ni,
                          int64_t *nl) {
    const int64_t n = t->n, off = n - 1;
    if (t->root == 0) {
        for (int64_t li = 0; li < n; ++li) {
            double hb = t->lh[li];
            double r2 = radius2_rule(strategy, Kv, radius2, hb);
            double d2 = d2_line_leaf(t, o, dvec, li);
            if (nl) ++*nl;
            if (d2 <= r2) fn(ctx, li, d2, hb);
        }
        return;
    }
    int32_t node = t->root;
    while (node != 0) {
        if (node > off) {
            int64_t li = node - off - 1;
            double hb = t->lh[li];
            double r2 = radius2_rule(strategy, Kv, radius2, hb);
            double d2 = d2_line_leaf(t, o, dvec, li);
            if (nl) ++*nl;
            if (d2 <= r2) fn(ctx, li, d2, hb);
            node = t->escape[node];
            continue;
        }
        if (ni) ++*ni;
        double r2 = radius2_rule(strategy, Kv, radius2, t->nhmax[node]);
        double d2n = d2_line_aabb_lb(t, o, dvec, node);
        node = (d2n > r2) ? t->escape[node] : t->left[node];
    }
}

/* LBVH_find_nearest_h, LinearBVHQueries.jl:118-141: gather traversal whose radius is the
 * running best distance; strict `<` keeps the first (smallest sorted index) of ties. */
static int64_t nearest_leaf(const orc_lbvh *t, const double *p) {
    const int64_t n = t->n, off = n - 1;
    int64_t best = -1;
    double bd = INFINITY;
    if (t->root == 0) {
        for (int64_t li = 0; li < n; ++li) {
            double d2 = d2_point_leaf(t, p, li);
            if (d2 <= bd && d2 < bd) { bd = d2; best = li; }
        }
        return best;
    }
    int32_t node = t->root;
    while (node != 0) {
        if (node > off) {
            int64_t li = node - off - 1;
            double d2 = d2_point_leaf(t, p, li);
            if (d2 <= bd && d2 < bd) { bd = d2; best = li; }
            node = t->escape[node];
            continue;
        }
        double d2n = d2_point_aabb(t, p, node);
        node = (d2n > bd) ? t->escape[node] : t->left[node];
    }
    return best;
}

void orc_nearest(const orc_lbvh *t, int64_t npts, const double *qx, const double *qy,
                 const double *qz, int64_t *idx_out, double *h_out, int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
    PRAGMA_OMP(omp parallel for schedule(dynamic, 64))
    for (int64_t i = 0; i < npts; ++i) {
        double p[3] = {qx[i], qy[i], qz[i]};
        int64_t b = nearest_leaf(t, p);
        if (idx_out) idx_out[i] = b;
        if (h_out) h_out[i] = b < 0 ? NAN : t->lh[b];
    }
}

/* ---- neighbour lists (test hooks; the reference's LBVH_query! records the same set) */
typedef struct { int64_t cnt, base, cap; int32_t *idx; } list_ctx;
static void list_leaf(void *c, int64_t leaf, double d2, double hb) {
    (void)d2; (void)hb;
    list_ctx *l = (list_ctx *)c;
    if (l->idx && l->base + l->cnt < l->cap) l->idx[l->base + l->cnt] = (int32_t)leaf;
    l->cnt++;
}

int64_t orc_point_neighbors(const orc_lbvh *t, int k, int strategy, int64_t npts,
                            const double *qx, const double *qy, const double *qz,
                            const double *ha, int64_t *offsets, int32_t *idx, int64_t cap) {
    const double Kv = orc_kernel_support(k);
    int64_t total = 0;
    for (int64_t i = 0; i < npts; ++i) {
        double p[3] = {qx[i], qy[i], qz[i]};
        double hai = 0.0;
        if (strategy != ORC_SCATTER) {
            if (ha) hai = ha[i];
            else { int64_t b = nearest_leaf(t, p); hai = b < 0 ? NAN : t->lh[b]; }
        }
        double radius = Kv * hai, radius2 = radius * radius;
        list_ctx l = {0, total, cap, idx};
        if (offsets) offsets[i] = total;
        traverse_point(t, p, strategy, Kv, radius2, list_leaf, &l, NULL, NULL);
        total += l.cnt;
    }
    if (offsets) offsets[npts] = total;
    return total;
}

int64_t orc_line_neighbors(const orc_lbvh *t, int k, int strategy, int64_t nl,
                           const double *ox, const double *oy, const double *oz,
                           const double *dx, const double *dy, const double *dz,
                           const double *ha, int64_t *offsets, int32_t *idx, int64_t cap) {
    const double Kv = orc_kernel_support(k);
    int64_t total = 0;
    for (int64_t i = 0; i < nl; ++i) {
        double o[3] = {ox[i], oy[i], oz[i]}, d[3] = {dx[i], dy[i], dz[i]};
        double hai = (strategy != ORC_SCATTER && ha) ? ha[i] : 0.0;
        double radius = Kv * hai, radius2 = radius * radius;
        list_ctx l = {0, total, cap, idx};
        if (offsets) offsets[i] = total;
        traverse_line(t, o, d, strategy, Kv, radius2, list_leaf, &l, NULL, NULL);
        total += l.cnt;
    }
    if (offsets) offsets[nl] = total;
    return total;
}

static void count_leaf(void *c, int64_t leaf, double d2, double hb) {
    (void)leaf; (void)d2; (void)hb;
    ++*(int64_t *)c;
}

/* Visit counts per query (internal nodes tested, leaves tested, leaves accepted) for the
 * traffic model of the traversal; includes the nearest-h pass for gather/symmetric. */
void orc_point_visits(const orc_lbvh *t, int k, int strategy, int64_t npts, const double *qx,
                      const double *qy, const double *qz, int64_t *n_internal, int64_t *n_leaf,
                      int64_t *n_accept, int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
    const double Kv = orc_kernel_support(k);
    PRAGMA_OMP(omp parallel for schedule(dynamic, 64))
    for (int64_t i = 0; i < npts; ++i) {
        double p[3] = {qx[i], qy[i], qz[i]};
        double hai = 0.0;
        if (strategy != ORC_SCATTER) { int64_t b = nearest_leaf(t, p); hai = t->lh[b]; }
        double radius = Kv * hai, radius2 = radius * radius;
        int64_t ni = 0, nl = 0, na = 0;
        traverse_point(t, p, strategy, Kv, radius2, count_leaf, &na, &ni, &nl);
        n_internal[i] = ni; n_leaf[i] = nl; n_accept[i] = na;
    }
}

void orc_line_visits(const orc_lbvh *t, int k, int64_t nl_, const double *ox, const double *oy,
                     const double *oz, const double *dx, const double *dy, const double *dz,
                     int64_t *n_internal, int64_t *n_leaf, int64_t *n_accept, int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
    const double Kv = orc_kernel_support(k);
    PRAGMA_OMP(omp parallel for schedule(dynamic, 64))
    for (int64_t i = 0; i < nl_; ++i) {
        double o[3] = {ox[i], oy[i], oz[i]}, d[3] = {dx[i], dy[i], dz[i]};
        int64_t ni = 0, nl = 0, na = 0;
        traverse_line(t, o, d, ORC_SCATTER, Kv, 0.0, count_leaf, &na, &ni, &nl);
        n_internal[i] = ni; n_leaf[i] = nl; n_accept[i] = na;
    }
}

/* ===================================================================================
 * Point interpolation — KI/single_point_interpolation/kernels/general_interpolation.jl
 * (gather :1-223, scatter :225-444, symmetric :446-669) with the accumulation helpers of
 * accumulations/{scalar,gradient,divergence,curl}_accumulation.jl.  The reference calls
 * each helper separately and every helper recomputes the same W / grad W from the same
 * operands, so evaluating them once per pair gives identical bits.
 * =================================================================================== */
#define ORC_MAXQ 16
typedef struct {
    const orc_lbvh *t;
    const double *m, *rho;
    const double *const *quant;
    const orc_catalog *cat;
    int k, strategy;
    double p[3], ha;
    double S1;
    double sc[ORC_MAXQ];
    double gf[ORC_MAXQ][3], gb[ORC_MAXQ][3], gs[ORC_MAXQ];
    double df[ORC_MAXQ], db[ORC_MAXQ][3], ds[ORC_MAXQ][3];
    double cf[ORC_MAXQ][3], cb[ORC_MAXQ][3], cs[ORC_MAXQ][3];
} point_acc;

static void point_leaf(void *c, int64_t li, double d2, double hb) {
    point_acc *a = (point_acc *)c;
    const orc_lbvh *t = a->t;
    const orc_catalog *cat = a->cat;
    const int k = a->k;
    double dx = a->p[0] - t->lx[li], dy = a->p[1] - t->ly[li], dz = a->p[2] - t->lz[li];
    double dr = sqrt(d2);
    double mb = a->m[li], rhob = a->rho[li];
    double W, g[3];
    if (a->strategy == ORC_SYMMETRIC) {
        W = 0.5 * (orc_W(k, dr, a->ha) + orc_W(k, dr, hb));
    } else {
        W = orc_W(k, dr, a->strategy == ORC_GATHER ? a->ha : hb);
    }
    int need_grad = cat->n_grad + cat->n_div + cat->n_curl > 0;
    if (need_grad) {
        if (a->strategy == ORC_SYMMETRIC) {
            double ga[3], gbv[3];
            orc_gradW(k, dx, dy, dz, a->ha, ga);
            orc_gradW(k, dx, dy, dz, hb, gbv);
            g[0] = 0.5 * (ga[0] + gbv[0]); g[1] = 0.5 * (ga[1] + gbv[1]); g[2] = 0.5 * (ga[2] + gbv[2]);
        } else {
            orc_gradW(k, dx, dy, dz, a->strategy == ORC_GATHER ? a->ha : hb, g);
        }
    }
    double w = mb * W / rhob;                      /* _ShepardNormalization_accumulation */
    a->S1 += w;
    for (int j = 0; j < cat->n_scalar; ++j) a->sc[j] += a->quant[cat->scalar_slots[j]][li] * w;
    if (!need_grad) return;
    double invrho = 1.0 / rhob;
    double bx = mb * invrho * g[0], by = mb * invrho * g[1], bz = mb * invrho * g[2];
    for (int j = 0; j < cat->n_grad; ++j) {
        int s = cat->grad_slots[j];
        if (s < 0) {                               /* _gradient_density_accumulation */
            double mx = mb * g[0], my = mb * g[1], mz = mb * g[2];
            a->gf[j][0] += mx; a->gf[j][1] += my; a->gf[j][2] += mz;
            a->gb[j][0] += mx * invrho; a->gb[j][1] += my * invrho; a->gb[j][2] += mz * invrho;
            a->gs[j] += mb * W;                    /* _density_accumulation */
        } else {                                   /* _gradient_quantity_accumulation */
            double Ab = a->quant[s][li];
            a->gf[j][0] += bx * Ab; a->gf[j][1] += by * Ab; a->gf[j][2] += bz * Ab;
            a->gb[j][0] += bx; a->gb[j][1] += by; a->gb[j][2] += bz;
            a->gs[j] += Ab * w;
        }
    }
    for (int j = 0; j < cat->n_div; ++j) {         /* _divergence_quantity_accumulation */
        double Ax = a->quant[cat->div_slots[3 * j]][li], Ay = a->quant[cat->div_slots[3 * j + 1]][li],
               Az = a->quant[cat->div_slots[3 * j + 2]][li];
        a->df[j] += bx * Ax + by * Ay + bz * Az;
        a->db[j][0] += bx; a->db[j][1] += by; a->db[j][2] += bz;
        a->ds[j][0] += Ax * w; a->ds[j][1] += Ay * w; a->ds[j][2] += Az * w;
    }
    for (int j = 0; j < cat->n_curl; ++j) {        /* _curl_quantity_accumulation */
        double Ax = a->quant[cat->curl_slots[3 * j]][li], Ay = a->quant[cat->curl_slots[3 * j + 1]][li],
               Az = a->quant[cat->curl_slots[3 * j + 2]][li];
        a->cf[j][0] += Ay * bz - Az * by;
        a->cf[j][1] += Az * bx - Ax * bz;
        a->cf[j][2] += Ax * by - Ay * bx;
        a->cb[j][0] += bx; a->cb[j][1] += by; a->cb[j][2] += bz;
        a->cs[j][0] += Ax * w; a->cs[j][1] += Ay * w; a->cs[j][2] += Az * w;
    }
}

/* Finalisation, general_interpolation.jl:574-668, stored in catalog order
 * (PointSamples_interpolation.jl:140-182). */
static void point_finalize(const point_acc *a, int64_t i, double *const *out) {
    const orc_catalog *cat = a->cat;
    int o = 0;
    if (a->S1 == 0.0) {
        int L = cat->n_scalar + 3 * cat->n_grad + cat->n_div + 3 * cat->n_curl;
        for (int j = 0; j < L; ++j) out[j][i] = NAN;
        return;
    }
    double invS1 = 1.0 / a->S1;
    for (int j = 0; j < cat->n_scalar; ++j)
        out[o++][i] = cat->scalar_shepard[j] ? a->sc[j] * invS1 : a->sc[j];
    for (int j = 0; j < cat->n_grad; ++j) {
        double A = a->gs[j] * invS1;
        for (int d = 0; d < 3; ++d) out[o++][i] = a->gf[j][d] - a->gb[j][d] * A;
    }
    for (int j = 0; j < cat->n_div; ++j) {
        double Ax = a->ds[j][0] * invS1, Ay = a->ds[j][1] * invS1, Az = a->ds[j][2] * invS1;
        double b = Ax * a->db[j][0] + Ay * a->db[j][1] + Az * a->db[j][2];
        out[o++][i] = a->df[j] - b;
    }
    for (int j = 0; j < cat->n_curl; ++j) {
        double Ax = a->cs[j][0] * invS1, Ay = a->cs[j][1] * invS1, Az = a->cs[j][2] * invS1;
        double mx = a->cb[j][0], my = a->cb[j][1], mz = a->cb[j][2];
        double bx = Ay * mz - Az * my, by = Az * mx - Ax * mz, bz = Ax * my - Ay * mx;
        out[o++][i] = -(a->cf[j][0] - bx);
        out[o++][i] = -(a->cf[j][1] - by);
        out[o++][i] = -(a->cf[j][2] - bz);
    }
}

/* PointSamples_interpolation(::CPUComputeBackend, ...) per-sample body
 * (PointSamples_interpolation.jl:53-185): nearest-h for gather/symmetric, then the
 * general kernel.  The LBVH must already have been built (initialize_interpolation). */
int orc_point_samples(const orc_lbvh *t, const double *m, const double *rho, int ncol,
                      const double *const *quant, int k, int strategy, const orc_catalog *cat,
                      int64_t npts, const double *qx, const double *qy, const double *qz,
                      double *const *out, int nthreads) {
    (void)ncol;
    if (cat->n_scalar > ORC_MAXQ || cat->n_grad > ORC_MAXQ || cat->n_div > ORC_MAXQ ||
        cat->n_curl > ORC_MAXQ)
        return -1;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
    const double Kv = orc_kernel_support(k);
    PRAGMA_OMP(omp parallel for schedule(dynamic, 16))
    for (int64_t i = 0; i < npts; ++i) {
        point_acc acc;
        point_acc *a = &acc;
        memset(a, 0, sizeof(point_acc));
        a->t = t; a->m = m; a->rho = rho; a->quant = quant; a->cat = cat; a->k = k;
        a->strategy = strategy;
        a->p[0] = qx[i]; a->p[1] = qy[i]; a->p[2] = qz[i];
        double radius2 = 0.0;
        if (strategy != ORC_SCATTER) {
            int64_t b = nearest_leaf(t, a->p);
            a->ha = b < 0 ? NAN : t->lh[b];
            double radius = Kv * a->ha;
            radius2 = radius * radius;
        }
        traverse_point(t, a->p, strategy, Kv, radius2, point_leaf, a, NULL, NULL);
        point_finalize(a, i, out);
    }
    return 0;
}

/* ===================================================================================
 * Line-integrated interpolation — KI/line_integrated_interpolation/
 * (line_integrated_scalar_interpolation.jl:76-214, accumulations :15-42).
 * d2 of a particle from the line can round to a value below zero; Julia's sqrt would
 * throw there, and this restatement (like the GPU) uses sqrt(max(d2, 0)).
 * =================================================================================== */
typedef struct {
    const double *m, *rho;
    const double *const *quant;
    const int32_t *slots;
    int nscal, k, strategy;
    double ha, S1, out[ORC_MAXQ];
} line_acc;

static void line_leaf(void *c, int64_t li, double d2, double hb) {
    line_acc *a = (line_acc *)c;
    double dr = sqrt(d2 > 0.0 ? d2 : 0.0);
    double mb = a->m[li], rhob = a->rho[li];
    double W;
    if (a->strategy == ORC_SYMMETRIC) W = 0.5 * (orc_line_W(a->k, dr, a->ha) + orc_line_W(a->k, dr, hb));
    else W = orc_line_W(a->k, dr, a->strategy == ORC_GATHER ? a->ha : hb);
    double w = mb * W / rhob;
    a->S1 += w;
    for (int j = 0; j < a->nscal; ++j) a->out[j] += a->quant[a->slots[j]][li] * w;
}

int orc_line_samples(const orc_lbvh *t, const double *m, const double *rho, int ncol,
                     const double *const *quant, int k, int strategy, const double *ha,
                     int nscal, const int32_t *slots, const uint8_t *shepard, int64_t nl,
                     const double *ox, const double *oy, const double *oz, const double *dx,
                     const double *dy, const double *dz, double *const *out, int nthreads) {
    (void)ncol;
    if (nscal > ORC_MAXQ) return -1;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
    const double Kv = orc_kernel_support(k);
    (void)line_table(k);
    PRAGMA_OMP(omp parallel for schedule(dynamic, 16))
    for (int64_t i = 0; i < nl; ++i) {
        line_acc a;
        memset(&a, 0, sizeof(a));
        a.m = m; a.rho = rho; a.quant = quant; a.slots = slots; a.nscal = nscal; a.k = k;
        a.strategy = strategy;
        a.ha = (strategy != ORC_SCATTER && ha) ? ha[i] : 0.0;
        double radius = Kv * a.ha, radius2 = radius * radius;
        double o[3] = {ox[i], oy[i], oz[i]}, d[3] = {dx[i], dy[i], dz[i]};
        traverse_line(t, o, d, strategy, Kv, radius2, line_leaf, &a, NULL, NULL);
        if (a.S1 == 0.0) {
            for (int j = 0; j < nscal; ++j) out[j][i] = NAN;
            continue;
        }
        double invS1 = 1.0 / a.S1;
        for (int j = 0; j < nscal; ++j) out[j][i] = shepard[j] ? a.out[j] * invS1 : a.out[j];
    }
    return 0;
}
