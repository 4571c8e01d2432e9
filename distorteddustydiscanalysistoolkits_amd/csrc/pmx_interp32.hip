// pmx_interp32.hip — FP32 compensated interpolation path for gfx950.
//
// Design (one wavefront = one tile of 64 Morton-ordered samples):
//   * Samples are sorted by a 30-bit Morton code of their position in the particle box, so a
//     tile covers a compact region and its 64 traversals overlap.
//   * The wavefront walks the LBVH once for the whole tile (stackless left / escape links,
//     wave-uniform node id, scalar loads): an internal node is entered if any lane's
//     conservative FP32 test needs it, a leaf is kept if any lane accepts it.
//   * Accepted leaves are staged in LDS in batches of 64 together with their payload
//     (position, h, m/rho, the catalog's field columns), loaded cooperatively by the lanes;
//     each lane then accumulates the batch for its own sample.
//   * Neighbour decisions are the reference's FP64 decisions: an FP32 test with a relative
//     guard band decides the clear cases, the band is re-evaluated with the reference's FP64
//     formula on the FP64 particle columns.  Node tests use boxes rounded outwards and hmax
//     rounded up, so pruning is conservative; the accepted set equals the FP64 predicate set.
//   * Kernel weights and field terms are FP32; per-sample sums are compensated (TwoSum).
//     Derivative sums accumulate A_b - A_ref (A_ref = first neighbour's value), which leaves
//     f - A~ b unchanged mathematically and removes the cancellation of large means.
//     Line-integrated (column) sums accumulate in FP64, as does the Shepard division.
// Reference semantics: general_interpolation.jl:1-669 (points), line_integrated_scalar_
// interpolation.jl:123-166 (lines), LinearBVHQueries.jl:118-141 (nearest-h).
#include <hip/hip_runtime.h>

#include <vector>

#include "pmx_internal.h"

namespace pmx {
namespace {

constexpr int kWave = 64;
constexpr float kEps = 4e-6f;     // guard band of FP32 distance / radius comparisons

// ---------------------------------------------------------------------------------------
// FP32 tree derived from the FP64 tree
// ---------------------------------------------------------------------------------------
// neighbouring floats towards -inf / +inf (finite, non-NaN input)
__device__ inline float next_down(float f) {
    if (f == 0.0f) return -__uint_as_float(1u);
    const uint32_t u = __float_as_uint(f);
    return __uint_as_float(f > 0.0f ? u - 1u : u + 1u);
}
__device__ inline float next_up(float f) { return -next_down(-f); }
__device__ inline float f32_down(double d) {
    float f = (float)d;
    if ((double)f > d) f = next_down(f);
    return f;
}
__device__ inline float f32_up(double d) {
    float f = (float)d;
    if ((double)f < d) f = next_up(f);
    return f;
}

__global__ void __launch_bounds__(256) k_prep_nodes(const Node64* __restrict__ n64, int64_t n, Node32* __restrict__ n32,
                                                    int32_t* __restrict__ esc) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x + 1; i < n; i += (int64_t)gridDim.x * 256) {
        const Node64 a = n64[i];
        Node32 b;
        b.mnx = f32_down(a.mn[0]); b.mny = f32_down(a.mn[1]); b.mnz = f32_down(a.mn[2]);
        b.mxx = f32_up(a.mx[0]); b.mxy = f32_up(a.mx[1]); b.mxz = f32_up(a.mx[2]);
        b.hmax = f32_up(a.hmax);
        b.left = a.left;
        n32[i] = b;
        esc[i] = a.escape;
    }
}

// leaf32: {x_hi, y_hi, z_hi, h}; leaflo: {x_lo, y_lo, z_lo, 1/h}; payload column 0 = m/rho,
// columns 1.. = quant columns, column 1+ncol = rho.
__global__ void __launch_bounds__(256) k_prep_leaves(const double* __restrict__ cols, int64_t n, int ncol,
                                                     float4* __restrict__ leaf, float4* __restrict__ leaflo,
                                                     float* __restrict__ pay, int* __restrict__ has_lo) {
    int lo_seen = 0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        const double x = cols[i], y = cols[n + i], z = cols[2 * n + i];
        const double m = cols[3 * n + i], h = cols[4 * n + i], rho = cols[5 * n + i];
        const float xh = (float)x, yh = (float)y, zh = (float)z;
        const float xl = (float)(x - (double)xh), yl = (float)(y - (double)yh), zl = (float)(z - (double)zh);
        lo_seen |= (xl != 0.0f) | (yl != 0.0f) | (zl != 0.0f);
        const float hf = (float)h;
        leaf[i] = make_float4(xh, yh, zh, hf);
        leaflo[i] = make_float4(xl, yl, zl, (float)(1.0 / h));
        pay[i] = (float)(m / rho);
        for (int c = 0; c < ncol; ++c) pay[(int64_t)(1 + c) * n + i] = (float)cols[(int64_t)(6 + c) * n + i];
        pay[(int64_t)(1 + ncol) * n + i] = (float)rho;
    }
    if (lo_seen) atomicOr(has_lo, 1);
}

// 30-bit Morton key of a sample position inside the particle box (clamped), for tiling.
__device__ inline uint64_t spread10(uint32_t v) {
    uint64_t x = v & 0x3ff;
    x = (x | (x << 16)) & 0x30000ffULL;
    x = (x | (x << 8)) & 0x300f00fULL;
    x = (x | (x << 4)) & 0x30c30c3ULL;
    x = (x | (x << 2)) & 0x9249249ULL;
    return x;
}
__device__ inline uint32_t q10(double v, double lo, double hi) {
    double f = hi > lo ? (v - lo) / (hi - lo) : 0.5;
    f = f < 0.0 ? 0.0 : (f > 1.0 ? 1.0 : f);
    return (uint32_t)(f * 1023.0);
}
__global__ void __launch_bounds__(256) k_query_keys(const double* __restrict__ x, const double* __restrict__ y,
                                                    const double* __restrict__ z, int64_t stride_is_packed, int64_t npts,
                                                    const Node64* __restrict__ root, uint64_t* __restrict__ keys,
                                                    int32_t* __restrict__ vals) {
    const Node64 r = root[1];
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < npts; i += (int64_t)gridDim.x * 256) {
        const double px = x[i], py = y[i], pz = z[i];
        uint64_t k = (spread10(q10(px, r.mn[0], r.mx[0])) << 2) | (spread10(q10(py, r.mn[1], r.mx[1])) << 1) |
                     spread10(q10(pz, r.mn[2], r.mx[2]));
        keys[i] = k;
        vals[i] = (int32_t)i;
    }
    (void)stride_is_packed;
}

__global__ void __launch_bounds__(256) k_iota(uint64_t* __restrict__ keys, int32_t* __restrict__ vals, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        keys[i] = 0;
        vals[i] = (int32_t)i;
    }
}

// ---------------------------------------------------------------------------------------
// FP32 kernel shapes (same piecewise polynomials as the FP64 path)
// ---------------------------------------------------------------------------------------
template <int K>
__device__ inline void shape_df32(float q, float& f, float& df) {
    if (K == kM4) {
        const float a = 2.0f - q, b = 1.0f - q;
        if (q < 1.0f) { f = 0.25f * (a * a * a) - b * b * b; df = -0.75f * (a * a) + 3.0f * (b * b); }
        else if (q < 2.0f) { f = 0.25f * (a * a * a); df = -0.75f * (a * a); }
        else { f = 0.0f; df = 0.0f; }
    } else if (K == kM5) {
        const float a = 2.5f - q, b = 1.5f - q, c = 0.5f - q;
        const float a3 = a * a * a, b3 = b * b * b, c3 = c * c * c;
        if (q < 0.5f) { f = (a3 * a - 5.0f * (b3 * b)) + 10.0f * (c3 * c); df = (-4.0f * a3 + 20.0f * b3) - 40.0f * c3; }
        else if (q < 1.5f) { f = a3 * a - 5.0f * (b3 * b); df = -4.0f * a3 + 20.0f * b3; }
        else if (q < 2.5f) { f = a3 * a; df = -4.0f * a3; }
        else { f = 0.0f; df = 0.0f; }
    } else if (K == kM6) {
        const float a = 3.0f - q, b = 2.0f - q, c = 1.0f - q;
        const float a4 = (a * a) * (a * a), b4 = (b * b) * (b * b), c4 = (c * c) * (c * c);
        if (q < 1.0f) { f = (a4 * a - 6.0f * (b4 * b)) + 15.0f * (c4 * c); df = (-5.0f * a4 + 30.0f * b4) - 75.0f * c4; }
        else if (q < 2.0f) { f = a4 * a - 6.0f * (b4 * b); df = -5.0f * a4 + 30.0f * b4; }
        else if (q < 3.0f) { f = a4 * a; df = -5.0f * a4; }
        else { f = 0.0f; df = 0.0f; }
    } else if (K == kC2) {
        if (q < 2.0f) {
            const float u = 1.0f - 0.5f * q, u3 = u * u * u;
            f = (u3 * u) * (2.0f * q + 1.0f);
            df = (u3 * u) * 2.0f - u3 * (4.0f * q + 2.0f);
        } else { f = 0.0f; df = 0.0f; }
    } else if (K == kC4) {
        if (q < 2.0f) {
            const float u = 1.0f - 0.5f * q, u5 = (u * u) * (u * u) * u;
            f = (u5 * u) * (((35.0f / 12.0f) * (q * q) + 3.0f * q) + 1.0f);
            df = (u5 * u) * ((35.0f / 6.0f) * q + 3.0f) - u5 * (((35.0f / 4.0f) * (q * q) + 9.0f * q) + 3.0f);
        } else { f = 0.0f; df = 0.0f; }
    } else {
        if (q < 2.0f) {
            const float u = 1.0f - 0.5f * q, u7 = (u * u) * (u * u) * (u * u) * u;
            f = (u7 * u) * (((4.0f * (q * q * q) + 6.25f * (q * q)) + 4.0f * q) + 1.0f);
            df = (u7 * u) * ((12.0f * (q * q) + 12.5f * q) + 4.0f) -
                 u7 * (((16.0f * (q * q * q) + 25.0f * (q * q)) + 16.0f * q) + 4.0f);
        } else { f = 0.0f; df = 0.0f; }
    }
}

// Compensated accumulator (TwoSum; -ffp-contract=off keeps it exact).
struct CSum {
    float s, c;
    __device__ inline void add(float x) {
        const float t = s + x;
        const float bp = t - s;
        c += (s - (t - bp)) + (x - bp);
        s = t;
    }
    __device__ inline double value() const { return (double)s + (double)c; }
};

// ---------------------------------------------------------------------------------------
// Tile state helpers
// ---------------------------------------------------------------------------------------
struct T32 {
    const Node32* __restrict__ nodes;
    const int32_t* __restrict__ node_esc;
    const float4* __restrict__ leaf;
    const float4* __restrict__ leaflo;
    const int32_t* __restrict__ leaf_esc;
    const double* __restrict__ cols64;     // FP64 x | y | z | m | h | rho | quant
    const float* __restrict__ pay;         // m/rho | quant... | rho
    int64_t n;
    int32_t root;
    int32_t has_lo;
};

__device__ inline uint64_t ballot(bool p) { return __ballot(p); }

__device__ inline int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// conservative FP32 squared distance from a point (hi + lo) to a box rounded outwards,
// scaled down by the guard factor so that it never exceeds the exact FP64 distance
__device__ inline float d2_box32(const float ph[3], const float pl[3], const Node32& b) {
    const float mn[3] = {b.mnx, b.mny, b.mnz}, mx[3] = {b.mxx, b.mxy, b.mxz};
    float s = 0.0f;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float lo_gap = (mn[k] - ph[k]) - pl[k];
        const float hi_gap = (ph[k] - mx[k]) + pl[k];
        const float t = lo_gap > 0.0f ? lo_gap : (hi_gap > 0.0f ? hi_gap : 0.0f);
        s += t * t;
    }
    return s * (1.0f - 2.0f * kEps);
}

// Exact reference point-leaf d2 in FP64 (_squared_distance_point_coords).
__device__ inline double d2_exact64(const double p[3], const T32& T, int64_t li) {
    const int64_t n = T.n;
    double s = 0.0, d;
    d = p[0] - T.cols64[li]; s += d * d;
    d = p[1] - T.cols64[n + li]; s += d * d;
    d = p[2] - T.cols64[2 * n + li]; s += d * d;
    return s;
}

template <int S>
__device__ inline double r2_exact64(double Kv, double radius2, double hb) {
    return radius2_rule<S>(Kv, radius2, hb);
}

// ---------------------------------------------------------------------------------------
// Point queries: nearest-h and the union traversal of a tile
// ---------------------------------------------------------------------------------------
struct PointLane {
    double p64[3];
    float ph[3], pl[3];
    bool valid;
};

__device__ inline PointLane point_lane(const double* qx, const double* qy, const double* qz, bool valid, int64_t qi) {
    PointLane q;
    q.valid = valid;
    q.p64[0] = valid ? qx[qi] : 0.0;
    q.p64[1] = valid ? qy[qi] : 0.0;
    q.p64[2] = valid ? qz[qi] : 0.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) { q.ph[k] = (float)q.p64[k]; q.pl[k] = (float)(q.p64[k] - (double)q.ph[k]); }
    return q;
}

__device__ inline float d2_leaf32(const PointLane& q, const float4& L, const float4& Lo) {
    const float dx = (q.ph[0] - L.x) + (q.pl[0] - Lo.x), dy = (q.ph[1] - L.y) + (q.pl[1] - Lo.y),
                dz = (q.ph[2] - L.z) + (q.pl[2] - Lo.z);
    return dx * dx + dy * dy + dz * dz;
}

// LBVH_find_nearest_h (LinearBVHQueries.jl:118-141) for a tile: exact FP64 argmin of d2
// with ties to the smaller sorted index (leaves are reached in increasing index order and
// only a strictly smaller FP64 distance replaces the best).  FP32 distances only select
// the leaves whose FP64 distance must be computed and the nodes that may be skipped.
__device__ inline int64_t nearest32(const T32& T, const PointLane& q) {
    const int64_t n = T.n;
    int64_t best = -1;
    double best64 = __builtin_inf();
    float best32 = __builtin_inff();
    if (T.root == 0) {
        for (int64_t li = 0; li < n; ++li) {
            const double d2 = d2_exact64(q.p64, T, li);
            if (q.valid && d2 < best64) { best64 = d2; best = li; }
        }
        return best;
    }
    int node = T.root;
    int steps = 0;
    while (node > 0 && node < 2 * n && ++steps <= 2 * n) {
        if (node >= n) {
            const int li = node - (int)n;
            const float d2 = d2_leaf32(q, T.leaf[li], T.leaflo[li]);
            const bool maybe = q.valid && (best < 0 || d2 <= best32 * (1.0f + kEps));
            if (ballot(maybe)) {
                if (maybe) {
                    const double e = d2_exact64(q.p64, T, li);
                    if (e < best64) { best64 = e; best = li; best32 = f32_up(e); }
                }
            }
            node = uniform(T.leaf_esc[li]);
            continue;
        }
        const Node32 nd = T.nodes[node];
        const float d2n = d2_box32(q.ph, q.pl, nd);
        const bool need = q.valid && (best < 0 || !(d2n > best32 * (1.0f + kEps)));
        node = ballot(need) ? uniform(nd.left) : uniform(T.node_esc[node]);
    }
    return best;
}

// Union traversal of a tile for the accepted sets of a strategy
// (LinearBVHTraversalMacros.jl:26-371).  Every leaf accepted by at least one lane is
// appended to the LDS batch (cidx, cmask = lanes accepting); on_batch(ncand) consumes a
// full batch and at the end the remainder.
template <int S, class OnBatch>
__device__ inline void union_points(const T32& T, const PointLane& q, double Kv, double radius2_64, float radius2_32,
                                    int32_t* cidx, uint64_t* cmask, OnBatch&& on_batch) {
    const int64_t n = T.n;
    const float Kvf = (float)Kv;
    const int lane = threadIdx.x;
    int ncand = 0;
    auto visit_leaf = [&](int li) {
        const float4 L = T.leaf[li];
        const float d2 = d2_leaf32(q, L, T.leaflo[li]);
        const float rb = Kvf * L.w, r2b = rb * rb;
        const float r2 = S == kGather ? radius2_32 : (S == kScatter ? r2b : fmaxf(radius2_32, r2b));
        bool acc = q.valid && d2 < r2 * (1.0f - kEps);
        const bool band = q.valid && !acc && !(d2 > r2 * (1.0f + kEps));
        if (ballot(band)) {
            if (band) {
                const double e = d2_exact64(q.p64, T, li);
                acc = e <= r2_exact64<S>(Kv, radius2_64, T.cols64[4 * n + li]);
            }
        }
        const uint64_t m = ballot(acc);
        if (m) {
            if (lane == 0) { cidx[ncand] = li; cmask[ncand] = m; }
            ++ncand;
            if (ncand == kWave) {
                __syncthreads();
                on_batch(ncand);
                ncand = 0;
            }
        }
    };
    if (T.root == 0) {
        for (int li = 0; li < (int)n; ++li) visit_leaf(li);
    } else {
        int node = T.root;
        int steps = 0;
        while (node > 0 && node < 2 * n && ++steps <= 2 * n) {
            if (node >= n) {
                const int li = node - (int)n;
                visit_leaf(li);
                node = uniform(T.leaf_esc[li]);
                continue;
            }
            const Node32 nd = T.nodes[node];
            const float d2n = d2_box32(q.ph, q.pl, nd);
            const float rb = Kvf * nd.hmax, r2b = rb * rb;
            const float r2 = S == kGather ? radius2_32 : (S == kScatter ? r2b : fmaxf(radius2_32, r2b));
            const bool need = q.valid && !(d2n > r2 * (1.0f + kEps));
            node = ballot(need) ? uniform(nd.left) : uniform(T.node_esc[node]);
        }
    }
    if (ncand > 0) {
        __syncthreads();
        on_batch(ncand);
    }
}

// ---------------------------------------------------------------------------------------
// Point kernel: accumulation
// ---------------------------------------------------------------------------------------
constexpr int kMaxF = 12;   // distinct field columns staged per candidate

struct PointCat {
    // distinct payload columns staged in LDS: payload column index (1.. = quant, 1+ncol = rho)
    int nf;
    int fcol[kMaxF];
    int ns, ng, nd, nc;
    int sf[4];        // scalar -> staged field
    int sh[4];
    int gf[2];        // gradient -> staged field (density gradient uses the rho column)
    int df[6];        // divergence components -> staged fields
    int cf[6];        // curl components -> staged fields
    int os, og, od, oc;
};

struct Cand {
    float x, y, z, h;
    float xl, yl, zl, invh;
    float w;
    float f[kMaxF];
};

template <int K, int S>
__global__ void __launch_bounds__(kWave) k_points32(T32 T, PointCat cat, const double* __restrict__ qx,
                                                    const double* __restrict__ qy, const double* __restrict__ qz,
                                                    const int32_t* __restrict__ perm, int64_t npts,
                                                    float* __restrict__ out, int64_t ldo) {
    __shared__ Cand cand[kWave];
    __shared__ uint64_t cmask[kWave];
    __shared__ int32_t cidx[kWave];
    const int lane = threadIdx.x;
    const double Kv = support(K);
    const float sig = norm3<float>(K);
    const int64_t n = T.n;
    const int64_t ntiles = (npts + kWave - 1) / kWave;

    for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int64_t slot = tile * kWave + lane;
        const bool valid = slot < npts;
        const int64_t qi = valid ? perm[slot] : 0;
        const PointLane q = point_lane(qx, qy, qz, valid, qi);

        double ha64 = 0.0, radius2_64 = 0.0;
        float radius2_32 = 0.0f;
        if (S != kScatter) {
            const int64_t best = nearest32(T, q);
            ha64 = best < 0 ? __builtin_nan("") : T.cols64[4 * n + best];
            const double radius = Kv * ha64;
            radius2_64 = radius * radius;
            radius2_32 = (float)radius2_64;
        }
        const float ha32 = (float)ha64;
        const float invha = 1.0f / ha32;
        const float invha3 = invha * invha * invha;
        const float invha4 = (invha * invha) * (invha * invha);

        CSum S1{0, 0}, sc[4] = {}, b[3] = {}, gfa[2][3] = {}, gs[2] = {}, dfa[2] = {}, ds[2][3] = {}, cfa[2][3] = {},
             cs[2][3] = {};
        // shift references of the derivative sums (first accepted neighbour's values)
        float refg[2] = {0, 0}, refd[2][3] = {}, refc[2][3] = {};
        bool have_ref = false;

        auto on_batch = [&](int ncand) {
            if (lane < ncand) {   // cooperative payload load
                const int li = cidx[lane];
                const float4 L = T.leaf[li];
                const float4 Lo = T.leaflo[li];
                Cand& c = cand[lane];
                c.x = L.x; c.y = L.y; c.z = L.z; c.h = L.w;
                c.xl = Lo.x; c.yl = Lo.y; c.zl = Lo.z; c.invh = Lo.w;
                c.w = T.pay[li];
                for (int f = 0; f < cat.nf; ++f) c.f[f] = T.pay[(int64_t)cat.fcol[f] * n + li];
            }
            __syncthreads();
            for (int j = 0; j < ncand; ++j) {
                if (!((cmask[j] >> lane) & 1ull)) continue;
                const Cand& c = cand[j];
                const float dx = (q.ph[0] - c.x) + (q.pl[0] - c.xl), dy = (q.ph[1] - c.y) + (q.pl[1] - c.yl),
                            dz = (q.ph[2] - c.z) + (q.pl[2] - c.zl);
                const float d2 = dx * dx + dy * dy + dz * dz;
                const float r = __builtin_amdgcn_sqrtf(d2);
                float W, gcoef;
                {
                    float fb, dfb;
                    shape_df32<K>(r * c.invh, fb, dfb);
                    const float ib2 = c.invh * c.invh;
                    const float Wb = (ib2 * c.invh) * (sig * fb);
                    const float Gb = (ib2 * ib2) * (sig * dfb);
                    if (S == kScatter) {
                        W = Wb; gcoef = Gb;
                    } else {
                        float fa, dfa;
                        shape_df32<K>(r * invha, fa, dfa);
                        const float Wa = invha3 * (sig * fa);
                        const float Ga = invha4 * (sig * dfa);
                        if (S == kGather) { W = Wa; gcoef = Ga; }
                        else { W = 0.5f * (Wa + Wb); gcoef = 0.5f * (Ga + Gb); }
                    }
                }
                const float w = c.w * W;
                S1.add(w);
#pragma unroll
                for (int s = 0; s < 4; ++s)
                    if (s < cat.ns) sc[s].add(c.f[cat.sf[s]] * w);
                if (cat.ng + cat.nd + cat.nc == 0) continue;
                if (!have_ref) {
#pragma unroll
                    for (int s = 0; s < 2; ++s) {
                        if (s < cat.ng) refg[s] = c.f[cat.gf[s]];
                        if (s < cat.nd)
                            for (int k = 0; k < 3; ++k) refd[s][k] = c.f[cat.df[3 * s + k]];
                        if (s < cat.nc)
                            for (int k = 0; k < 3; ++k) refc[s][k] = c.f[cat.cf[3 * s + k]];
                    }
                    have_ref = true;
                }
                // (m/rho) grad W = (m/rho) f'(q) sigma h^-4 / r * Delta (zero at r == 0)
                const float cg = r > 0.0f ? c.w * gcoef * __builtin_amdgcn_rcpf(r) : 0.0f;
                const float gx = cg * dx, gy = cg * dy, gz = cg * dz;
                b[0].add(gx); b[1].add(gy); b[2].add(gz);
#pragma unroll
                for (int s = 0; s < 2; ++s) {
                    if (s < cat.ng) {
                        const float A = c.f[cat.gf[s]] - refg[s];
                        gfa[s][0].add(A * gx); gfa[s][1].add(A * gy); gfa[s][2].add(A * gz);
                        gs[s].add(A * w);
                    }
                    if (s < cat.nd) {
                        const float Ax = c.f[cat.df[3 * s]] - refd[s][0];
                        const float Ay = c.f[cat.df[3 * s + 1]] - refd[s][1];
                        const float Az = c.f[cat.df[3 * s + 2]] - refd[s][2];
                        dfa[s].add(Ax * gx + Ay * gy + Az * gz);
                        ds[s][0].add(Ax * w); ds[s][1].add(Ay * w); ds[s][2].add(Az * w);
                    }
                    if (s < cat.nc) {
                        const float Ax = c.f[cat.cf[3 * s]] - refc[s][0];
                        const float Ay = c.f[cat.cf[3 * s + 1]] - refc[s][1];
                        const float Az = c.f[cat.cf[3 * s + 2]] - refc[s][2];
                        cfa[s][0].add(Ay * gz - Az * gy);
                        cfa[s][1].add(Az * gx - Ax * gz);
                        cfa[s][2].add(Ax * gy - Ay * gx);
                        cs[s][0].add(Ax * w); cs[s][1].add(Ay * w); cs[s][2].add(Az * w);
                    }
                }
            }
            __syncthreads();
        };
        union_points<S>(T, q, Kv, radius2_64, radius2_32, cidx, cmask, on_batch);

        // ---- finalisation (general_interpolation.jl:574-668) in FP64 from the compensated sums
        if (valid) {
            const double s1 = S1.value();
            if (s1 == 0.0) {
                const float nan = __builtin_nanf("");
                for (int j = 0; j < cat.ns; ++j) out[(cat.os + j) * ldo + qi] = nan;
                for (int j = 0; j < 3 * cat.ng; ++j) out[(cat.og + j) * ldo + qi] = nan;
                for (int j = 0; j < cat.nd; ++j) out[(cat.od + j) * ldo + qi] = nan;
                for (int j = 0; j < 3 * cat.nc; ++j) out[(cat.oc + j) * ldo + qi] = nan;
            } else {
                const double inv = 1.0 / s1;
                const double bx = b[0].value(), by = b[1].value(), bz = b[2].value();
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    if (j >= cat.ns) break;
                    const double v = sc[j].value();
                    out[(cat.os + j) * ldo + qi] = (float)(cat.sh[j] ? v * inv : v);
                }
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    if (j < cat.ng) {
                        const double A = gs[j].value() * inv;
                        out[(cat.og + 3 * j) * ldo + qi] = (float)(gfa[j][0].value() - bx * A);
                        out[(cat.og + 3 * j + 1) * ldo + qi] = (float)(gfa[j][1].value() - by * A);
                        out[(cat.og + 3 * j + 2) * ldo + qi] = (float)(gfa[j][2].value() - bz * A);
                    }
                    if (j < cat.nd) {
                        const double Ax = ds[j][0].value() * inv, Ay = ds[j][1].value() * inv, Az = ds[j][2].value() * inv;
                        out[(cat.od + j) * ldo + qi] = (float)(dfa[j].value() - (Ax * bx + Ay * by + Az * bz));
                    }
                    if (j < cat.nc) {
                        const double Ax = cs[j][0].value() * inv, Ay = cs[j][1].value() * inv, Az = cs[j][2].value() * inv;
                        const double ex = Ay * bz - Az * by, ey = Az * bx - Ax * bz, ez = Ax * by - Ay * bx;
                        out[(cat.oc + 3 * j) * ldo + qi] = (float)(-(cfa[j][0].value() - ex));
                        out[(cat.oc + 3 * j + 1) * ldo + qi] = (float)(-(cfa[j][1].value() - ey));
                        out[(cat.oc + 3 * j + 2) * ldo + qi] = (float)(-(cfa[j][2].value() - ez));
                    }
                }
            }
        }
    }
}

// Neighbour lists / counts of the compensated path (inspection hook): the same union
// traversal and decisions as k_points32, recording accepted leaves in DFS order.
template <int S>
__global__ void __launch_bounds__(kWave) k_points32_nb(T32 T, double Kv, const double* __restrict__ qx,
                                                       const double* __restrict__ qy, const double* __restrict__ qz,
                                                       const double* __restrict__ ha_in, int64_t npts,
                                                       int64_t* __restrict__ counts, const int64_t* __restrict__ offsets,
                                                       int32_t* __restrict__ idx) {
    __shared__ uint64_t cmask[kWave];
    __shared__ int32_t cidx[kWave];
    const int lane = threadIdx.x;
    const int64_t n = T.n;
    const int64_t ntiles = (npts + kWave - 1) / kWave;
    for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int64_t qi = tile * kWave + lane;
        const bool valid = qi < npts;
        const PointLane q = point_lane(qx, qy, qz, valid, valid ? qi : 0);
        double radius2_64 = 0.0;
        if (S != kScatter) {
            double ha;
            if (ha_in) ha = valid ? ha_in[qi] : 0.0;
            else {
                const int64_t best = nearest32(T, q);
                ha = best < 0 ? __builtin_nan("") : T.cols64[4 * n + best];
            }
            const double r = Kv * ha;
            radius2_64 = r * r;
        }
        int64_t cnt = 0;
        int32_t* dst = (valid && offsets) ? idx + offsets[qi] : nullptr;
        auto on_batch = [&](int ncand) {
            for (int j = 0; j < ncand; ++j)
                if ((cmask[j] >> lane) & 1ull) {
                    if (dst) dst[cnt] = cidx[j];
                    ++cnt;
                }
            __syncthreads();
        };
        union_points<S>(T, q, Kv, radius2_64, (float)radius2_64, cidx, cmask, on_batch);
        if (valid) counts[qi] = cnt;
    }
}

__global__ void __launch_bounds__(kWave) k_nearest32(T32 T, const double* __restrict__ qx, const double* __restrict__ qy,
                                                     const double* __restrict__ qz, int64_t npts,
                                                     int64_t* __restrict__ idx) {
    const int64_t ntiles = (npts + kWave - 1) / kWave;
    for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int64_t qi = tile * kWave + threadIdx.x;
        const bool valid = qi < npts;
        const PointLane q = point_lane(qx, qy, qz, valid, valid ? qi : 0);
        const int64_t best = nearest32(T, q);
        if (valid) idx[qi] = best;
    }
}

// ---------------------------------------------------------------------------------------
// Line kernel (scatter; line_integrated_scalar_interpolation.jl:123-166)
// ---------------------------------------------------------------------------------------
struct LineCat32 {
    int n;
    int fcol[8];    // payload column of each output (1.. = quant)
    int shep[8];
};

struct LCand {
    float x, y, z, h;
    float xl, yl, zl, invh;
    float w;
    float f[8];
};

// Exact reference line-leaf d2 in FP64 (_squared_distance_line_coords).
__device__ inline double d2_line_exact64(const double o[3], const double dv[3], const T32& T, int64_t li) {
    const int64_t n = T.n;
    double D2 = 0.0, Dm = 0.0, d;
    d = T.cols64[li] - o[0]; D2 += d * d; Dm += d * dv[0];
    d = T.cols64[n + li] - o[1]; D2 += d * d; Dm += d * dv[1];
    d = T.cols64[2 * n + li] - o[2]; D2 += d * d; Dm += d * dv[2];
    return D2 - Dm * Dm;
}

// FP32 perpendicular distance^2 of a point from a line (cross-product form, no cancellation)
// and |Delta|^2 (for the rounding bound of the reference's D2 - Dm^2 form).
__device__ inline float d2_line32(const float oh[3], const float ol[3], const float dv[3], float x, float y, float z,
                                  float xl, float yl, float zl, float& D2) {
    const float ax = (x - oh[0]) + (xl - ol[0]), ay = (y - oh[1]) + (yl - ol[1]), az = (z - oh[2]) + (zl - ol[2]);
    D2 = ax * ax + ay * ay + az * az;
    const float cx = ay * dv[2] - az * dv[1], cy = az * dv[0] - ax * dv[2], cz = ax * dv[1] - ay * dv[0];
    return cx * cx + cy * cy + cz * cz;
}

// Conservative lower bound of the squared distance between a line and a box (rounded
// outwards).  Axis-aligned directions use the exact rectangle distance in the normal plane;
// other directions use the bounding-sphere bound of the reference with guard factors.
__device__ inline float d2_line_box32(const float oh[3], const float ol[3], const float dv[3], int axis,
                                      const Node32& b) {
    const float mn[3] = {b.mnx, b.mny, b.mnz}, mx[3] = {b.mxx, b.mxy, b.mxz};
    if (axis >= 0) {
        float s = 0.0f;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            if (k == axis) continue;
            const float lo_gap = (mn[k] - oh[k]) - ol[k];
            const float hi_gap = (oh[k] - mx[k]) + ol[k];
            const float t = lo_gap > 0.0f ? lo_gap : (hi_gap > 0.0f ? hi_gap : 0.0f);
            s += t * t;
        }
        return s * (1.0f - 2.0f * kEps);
    }
    float c[3], R2 = 0.0f;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        c[k] = 0.5f * (mn[k] + mx[k]);
        const float hh = 0.5f * (mx[k] - mn[k]);
        R2 += hh * hh;
    }
    float D2;
    const float dc2 = d2_line32(oh, ol, dv, c[0], c[1], c[2], 0.0f, 0.0f, 0.0f, D2);
    const float dc = __builtin_amdgcn_sqrtf(dc2) * (1.0f - kEps);
    const float R = __builtin_amdgcn_sqrtf(R2) * (1.0f + kEps) + 1e-30f;
    if (dc <= R) return 0.0f;
    const float g = (dc - R) * (1.0f - kEps);
    return g * g;
}

__device__ inline float lin_lut32(float q, float inv_dq, const float* __restrict__ I) {
    const float idxf = q * inv_dq + 1.0f;
    int i = (int)idxf;
    i = i < 1 ? 1 : (i > kLutN - 1 ? kLutN - 1 : i);
    const float t = idxf - (float)i;
    return I[i - 1] * (1.0f - t) + I[i] * t;
}

struct LineLane {
    double o64[3], d64[3];
    float oh[3], ol[3], dv[3];
    int axis;     // index of the direction axis for axis-aligned lines, else -1
    bool valid;
};

__device__ inline LineLane line_lane(const double* o, const double* dvec, int64_t nl, bool valid, int64_t qi) {
    LineLane L;
    L.valid = valid;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        L.o64[k] = valid ? o[k * nl + qi] : 0.0;
        L.d64[k] = valid ? dvec[k * nl + qi] : (k == 2 ? 1.0 : 0.0);
        L.oh[k] = (float)L.o64[k];
        L.ol[k] = (float)(L.o64[k] - (double)L.oh[k]);
        L.dv[k] = (float)L.d64[k];
    }
    L.axis = -1;
    if (L.d64[0] == 0.0 && L.d64[1] == 0.0 && (L.d64[2] == 1.0 || L.d64[2] == -1.0)) L.axis = 2;
    else if (L.d64[0] == 0.0 && L.d64[2] == 0.0 && (L.d64[1] == 1.0 || L.d64[1] == -1.0)) L.axis = 1;
    else if (L.d64[1] == 0.0 && L.d64[2] == 0.0 && (L.d64[0] == 1.0 || L.d64[0] == -1.0)) L.axis = 0;
    return L;
}

// Union traversal of a tile of lines, scatter rule r = K h (LinearBVHTraversalMacros.jl:539-644).
template <class OnBatch>
__device__ inline void union_lines(const T32& T, const LineLane& q, double Kv, int32_t* cidx, uint64_t* cmask,
                                   OnBatch&& on_batch) {
    const int64_t n = T.n;
    const float Kvf = (float)Kv;
    const int lane = threadIdx.x;
    int ncand = 0;
    auto visit_leaf = [&](int li) {
        const float4 L = T.leaf[li];
        const float4 Lo = T.leaflo[li];
        float D2;
        const float d2 = d2_line32(q.oh, q.ol, q.dv, L.x, L.y, L.z, Lo.x, Lo.y, Lo.z, D2);
        const float rb = Kvf * L.w, r2 = rb * rb;
        // band: FP32 relative error of d2 plus the FP64 rounding of the reference's D2 - Dm^2
        const float tol = r2 * kEps + 1e-14f * D2;
        bool acc = q.valid && d2 < r2 - tol;
        const bool band = q.valid && !acc && !(d2 > r2 + tol);
        if (ballot(band)) {
            if (band) {
                const double e = d2_line_exact64(q.o64, q.d64, T, li);
                const double rr = Kv * T.cols64[4 * n + li];
                acc = e <= rr * rr;
            }
        }
        const uint64_t m = ballot(acc);
        if (m) {
            if (lane == 0) { cidx[ncand] = li; cmask[ncand] = m; }
            ++ncand;
            if (ncand == kWave) {
                __syncthreads();
                on_batch(ncand);
                ncand = 0;
            }
        }
    };
    if (T.root == 0) {
        for (int li = 0; li < (int)n; ++li) visit_leaf(li);
    } else {
        int node = T.root;
        int steps = 0;
        while (node > 0 && node < 2 * n && ++steps <= 2 * n) {
            if (node >= n) {
                const int li = node - (int)n;
                visit_leaf(li);
                node = uniform(T.leaf_esc[li]);
                continue;
            }
            const Node32 nd = T.nodes[node];
            const float d2n = d2_line_box32(q.oh, q.ol, q.dv, q.axis, nd);
            const float rb = Kvf * nd.hmax;
            const bool need = q.valid && !(d2n > rb * rb * (1.0f + kEps));
            node = ballot(need) ? uniform(nd.left) : uniform(T.node_esc[node]);
        }
    }
    if (ncand > 0) {
        __syncthreads();
        on_batch(ncand);
    }
}

template <int K>
__global__ void __launch_bounds__(kWave) k_lines32(T32 T, LineCat32 cat, const double* __restrict__ o,
                                                   const double* __restrict__ dvec, const int32_t* __restrict__ perm,
                                                   int64_t nl, const float* __restrict__ lut, float inv_dq,
                                                   float* __restrict__ out, int64_t ldo) {
    __shared__ LCand cand[kWave];
    __shared__ uint64_t cmask[kWave];
    __shared__ int32_t cidx[kWave];
    const int lane = threadIdx.x;
    const double Kv = support(K);
    const float Kvf = (float)Kv;
    const int64_t n = T.n;
    const int64_t ntiles = (nl + kWave - 1) / kWave;

    for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int64_t slot = tile * kWave + lane;
        const bool valid = slot < nl;
        const int64_t qi = valid ? perm[slot] : 0;
        const LineLane q = line_lane(o, dvec, nl, valid, qi);
        double S1 = 0.0, acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        auto on_batch = [&](int ncand) {
            if (lane < ncand) {
                const int li = cidx[lane];
                const float4 L = T.leaf[li];
                const float4 Lo = T.leaflo[li];
                LCand& c = cand[lane];
                c.x = L.x; c.y = L.y; c.z = L.z; c.h = L.w;
                c.xl = Lo.x; c.yl = Lo.y; c.zl = Lo.z; c.invh = Lo.w;
                c.w = T.pay[li];
                for (int f = 0; f < cat.n; ++f) c.f[f] = T.pay[(int64_t)cat.fcol[f] * n + li];
            }
            __syncthreads();
            for (int j = 0; j < ncand; ++j) {
                if (!((cmask[j] >> lane) & 1ull)) continue;
                const LCand& c = cand[j];
                float D2;
                const float d2 = d2_line32(q.oh, q.ol, q.dv, c.x, c.y, c.z, c.xl, c.yl, c.zl, D2);
                const float qq = __builtin_amdgcn_sqrtf(d2) * c.invh;
                const float I = qq >= Kvf ? 0.0f : lin_lut32(qq, inv_dq, lut);
                const double wd = (double)(c.w * (c.invh * I));
                S1 += wd;
#pragma unroll
                for (int f = 0; f < 8; ++f)
                    if (f < cat.n) acc[f] += (double)c.f[f] * wd;
            }
            __syncthreads();
        };
        union_lines(T, q, Kv, cidx, cmask, on_batch);
        if (valid) {
            if (S1 == 0.0) {
                for (int f = 0; f < cat.n; ++f) out[f * ldo + qi] = __builtin_nanf("");
            } else {
                const double inv = 1.0 / S1;
#pragma unroll
                for (int f = 0; f < 8; ++f)
                    if (f < cat.n) out[f * ldo + qi] = (float)(cat.shep[f] ? acc[f] * inv : acc[f]);
            }
        }
    }
}

__global__ void __launch_bounds__(kWave) k_lines32_nb(T32 T, double Kv, const double* __restrict__ o,
                                                      const double* __restrict__ dvec, int64_t nl,
                                                      int64_t* __restrict__ counts, const int64_t* __restrict__ offsets,
                                                      int32_t* __restrict__ idx) {
    __shared__ uint64_t cmask[kWave];
    __shared__ int32_t cidx[kWave];
    const int lane = threadIdx.x;
    const int64_t ntiles = (nl + kWave - 1) / kWave;
    for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int64_t qi = tile * kWave + lane;
        const bool valid = qi < nl;
        const LineLane q = line_lane(o, dvec, nl, valid, valid ? qi : 0);
        int64_t cnt = 0;
        int32_t* dst = (valid && offsets) ? idx + offsets[qi] : nullptr;
        auto on_batch = [&](int ncand) {
            for (int j = 0; j < ncand; ++j)
                if ((cmask[j] >> lane) & 1ull) {
                    if (dst) dst[cnt] = cidx[j];
                    ++cnt;
                }
            __syncthreads();
        };
        union_lines(T, q, Kv, cidx, cmask, on_batch);
        if (valid) counts[qi] = cnt;
    }
}

T32 tree32(pmx_ctx* c) {
    T32 T;
    T.nodes = c->nodes32.as<Node32>();
    T.node_esc = c->node_esc.as<int32_t>();
    T.leaf = c->leaf32.as<float4>();
    T.leaflo = c->leaf32.as<float4>() + c->tree.n;
    T.leaf_esc = c->leaf_esc.as<int32_t>();
    T.cols64 = c->cols64.as<double>();
    T.pay = c->pay32.as<float>();
    T.n = c->tree.n;
    T.root = c->tree.root;
    T.has_lo = 1;
    return T;
}

// Morton order of the samples (queries): permutation into c->scratch_b
int sort_queries(pmx_ctx* c, const double* x, const double* y, const double* z, int64_t npts, int32_t** perm) {
    PMX_TRY(ensure(c, c->scratch_a, 8 * (size_t)npts + 4 * (size_t)npts + 64));
    PMX_TRY(ensure(c, c->scratch_b, 8 * (size_t)npts + 4 * (size_t)npts + 64));
    uint64_t* k0 = c->scratch_a.as<uint64_t>();
    int32_t* v0 = reinterpret_cast<int32_t*>(k0 + npts);
    uint64_t* k1 = c->scratch_b.as<uint64_t>();
    int32_t* v1 = reinterpret_cast<int32_t*>(k1 + npts);
    const dim3 g(grid_for(npts, 256, 8192)), b(256);
    if (c->tree.n >= 2) {
        hipLaunchKernelGGL(k_query_keys, g, b, 0, c->stream, x, y, z, 0, npts, c->nodes64.as<Node64>(), k0, v0);
        bool alt = false;
        PMX_TRY(radix_sort_pairs(c, k0, v0, npts, k1, v1, 0, 32, &alt));
        *perm = alt ? v1 : v0;
    } else {
        hipLaunchKernelGGL(k_iota, g, b, 0, c->stream, k0, v0, npts);
        *perm = v0;
    }
    PMX_HIP(c, hipGetLastError());
    return PMX_OK;
}

template <int K>
void launch_points32(pmx_ctx* c, int S, const PointCat& cat, const double* qx, const double* qy, const double* qz,
                     const int32_t* perm, int64_t npts, float* out, int64_t ldo) {
    const int64_t ntiles = (npts + kWave - 1) / kWave;
    const dim3 g((unsigned)std::min<int64_t>(ntiles, 1 << 20)), b(kWave);
    T32 T = tree32(c);
    if (S == kGather)
        hipLaunchKernelGGL((k_points32<K, kGather>), g, b, 0, c->stream, T, cat, qx, qy, qz, perm, npts, out, ldo);
    else if (S == kScatter)
        hipLaunchKernelGGL((k_points32<K, kScatter>), g, b, 0, c->stream, T, cat, qx, qy, qz, perm, npts, out, ldo);
    else
        hipLaunchKernelGGL((k_points32<K, kSymmetric>), g, b, 0, c->stream, T, cat, qx, qy, qz, perm, npts, out, ldo);
}

template <int K>
void launch_lines32(pmx_ctx* c, const LineCat32& cat, const double* o, const double* d, const int32_t* perm,
                    int64_t nl, float* out, int64_t ldo) {
    const int64_t ntiles = (nl + kWave - 1) / kWave;
    const dim3 g((unsigned)std::min<int64_t>(ntiles, 1 << 20)), b(kWave);
    const float inv_dq = (float)(1.0 / c->lut_dq[K]);
    hipLaunchKernelGGL(k_lines32<K>, g, b, 0, c->stream, tree32(c), cat, o, d, perm, nl,
                       c->lut32.as<float>() + (size_t)K * kLutN, inv_dq, out, ldo);
}

}  // namespace

int prepare_f32(pmx_ctx* c) {
    if (c->tree.f32_ready) return PMX_OK;
    const int64_t n = c->tree.n;
    const int ncol = c->tree.ncol;
    PMX_TRY(ensure(c, c->nodes32, sizeof(Node32) * (size_t)(n + 1)));
    PMX_TRY(ensure(c, c->node_esc, 4 * (size_t)(n + 1)));
    PMX_TRY(ensure(c, c->leaf32, 2 * sizeof(float4) * (size_t)n));
    PMX_TRY(ensure(c, c->pay32, 4 * (size_t)n * (2 + ncol)));
    PMX_TRY(ensure(c, c->misc, sizeof(double) * (6 * 1024 + 8) + 64));
    int* has_lo = reinterpret_cast<int*>(c->misc.as<double>() + 6 * 1024 + 7);
    const dim3 g(grid_for(n, 256, 16384)), b(256);
    if (n >= 2)
        hipLaunchKernelGGL(k_prep_nodes, g, b, 0, c->stream, c->nodes64.as<Node64>(), n, c->nodes32.as<Node32>(),
                           c->node_esc.as<int32_t>());
    hipLaunchKernelGGL(k_prep_leaves, g, b, 0, c->stream, c->cols64.as<double>(), n, ncol, c->leaf32.as<float4>(),
                       c->leaf32.as<float4>() + n, c->pay32.as<float>(), has_lo);
    PMX_HIP(c, hipGetLastError());
    c->tree.f32_ready = true;
    return PMX_OK;
}

int interp_points_f32c(pmx_ctx* c, int strategy, const CatChunk& ch, const double* qx, const double* qy,
                       const double* qz, int64_t npts, float* out, int64_t ldo) {
    const int ncol = c->tree.ncol;
    PointCat cat{};
    auto field = [&](int paycol) {
        for (int f = 0; f < cat.nf; ++f)
            if (cat.fcol[f] == paycol) return f;
        cat.fcol[cat.nf] = paycol;
        return cat.nf++;
    };
    cat.ns = ch.ns; cat.ng = ch.ng; cat.nd = ch.nd; cat.nc = ch.nc;
    for (int j = 0; j < ch.ns; ++j) { cat.sf[j] = field(1 + ch.ss[j]); cat.sh[j] = ch.sh[j]; }
    for (int j = 0; j < ch.ng; ++j) cat.gf[j] = field(ch.gs[j] < 0 ? 1 + ncol : 1 + ch.gs[j]);
    for (int j = 0; j < 3 * ch.nd; ++j) cat.df[j] = field(1 + ch.ds[j]);
    for (int j = 0; j < 3 * ch.nc; ++j) cat.cf[j] = field(1 + ch.cs[j]);
    if (cat.nf > kMaxF) return set_error(c, PMX_ERR_ARGUMENT, "catalog chunk uses too many distinct columns");
    cat.os = ch.os; cat.og = ch.og; cat.od = ch.od; cat.oc = ch.oc;
    int32_t* perm = nullptr;
    PMX_TRY(sort_queries(c, qx, qy, qz, npts, &perm));
    switch (c->tree.kernel) {
    case kM4: launch_points32<kM4>(c, strategy, cat, qx, qy, qz, perm, npts, out, ldo); break;
    case kM5: launch_points32<kM5>(c, strategy, cat, qx, qy, qz, perm, npts, out, ldo); break;
    case kM6: launch_points32<kM6>(c, strategy, cat, qx, qy, qz, perm, npts, out, ldo); break;
    case kC2: launch_points32<kC2>(c, strategy, cat, qx, qy, qz, perm, npts, out, ldo); break;
    case kC4: launch_points32<kC4>(c, strategy, cat, qx, qy, qz, perm, npts, out, ldo); break;
    default: launch_points32<kC6>(c, strategy, cat, qx, qy, qz, perm, npts, out, ldo); break;
    }
    PMX_HIP(c, hipGetLastError());
    return PMX_OK;
}

int interp_lines_f32c(pmx_ctx* c, const int* slots, const int* shep, int ncols, const double* o, const double* d,
                      int64_t nl, float* out, int64_t ldo) {
    int32_t* perm = nullptr;
    PMX_TRY(sort_queries(c, o, o + nl, o + 2 * nl, nl, &perm));
    for (int base = 0; base < ncols || (ncols == 0 && base == 0); base += 8) {
        LineCat32 cat{};
        cat.n = ncols - base < 8 ? ncols - base : 8;
        for (int j = 0; j < cat.n; ++j) { cat.fcol[j] = 1 + slots[base + j]; cat.shep[j] = shep[base + j]; }
        float* ob = out + (int64_t)base * ldo;
        switch (c->tree.kernel) {
        case kM4: launch_lines32<kM4>(c, cat, o, d, perm, nl, ob, ldo); break;
        case kM5: launch_lines32<kM5>(c, cat, o, d, perm, nl, ob, ldo); break;
        case kM6: launch_lines32<kM6>(c, cat, o, d, perm, nl, ob, ldo); break;
        case kC2: launch_lines32<kC2>(c, cat, o, d, perm, nl, ob, ldo); break;
        case kC4: launch_lines32<kC4>(c, cat, o, d, perm, nl, ob, ldo); break;
        default: launch_lines32<kC6>(c, cat, o, d, perm, nl, ob, ldo); break;
        }
        if (ncols == 0) break;
    }
    PMX_HIP(c, hipGetLastError());
    return PMX_OK;
}

int nearest_f32c(pmx_ctx* c, const double* qx, const double* qy, const double* qz, int64_t npts, int64_t* idx) {
    PMX_TRY(prepare_f32(c));
    const int64_t ntiles = (npts + kWave - 1) / kWave;
    hipLaunchKernelGGL(k_nearest32, dim3((unsigned)std::min<int64_t>(ntiles, 1 << 20)), dim3(kWave), 0, c->stream,
                       tree32(c), qx, qy, qz, npts, idx);
    PMX_HIP(c, hipGetLastError());
    return PMX_OK;
}

int neighbors_points_f32c(pmx_ctx* c, int strategy, const double* qx, const double* qy, const double* qz,
                          const double* ha, int64_t npts, int64_t* counts, const int64_t* offsets, int32_t* idx) {
    PMX_TRY(prepare_f32(c));
    const int64_t ntiles = (npts + kWave - 1) / kWave;
    const dim3 g((unsigned)std::min<int64_t>(ntiles, 1 << 20)), b(kWave);
    const double Kv = support(c->tree.kernel);
    T32 T = tree32(c);
    if (strategy == kGather)
        hipLaunchKernelGGL(k_points32_nb<kGather>, g, b, 0, c->stream, T, Kv, qx, qy, qz, ha, npts, counts, offsets, idx);
    else if (strategy == kScatter)
        hipLaunchKernelGGL(k_points32_nb<kScatter>, g, b, 0, c->stream, T, Kv, qx, qy, qz, ha, npts, counts, offsets, idx);
    else
        hipLaunchKernelGGL(k_points32_nb<kSymmetric>, g, b, 0, c->stream, T, Kv, qx, qy, qz, ha, npts, counts, offsets,
                           idx);
    PMX_HIP(c, hipGetLastError());
    return PMX_OK;
}

int neighbors_lines_f32c(pmx_ctx* c, const double* o, const double* d, int64_t nl, int64_t* counts,
                         const int64_t* offsets, int32_t* idx) {
    PMX_TRY(prepare_f32(c));
    const int64_t ntiles = (nl + kWave - 1) / kWave;
    hipLaunchKernelGGL(k_lines32_nb, dim3((unsigned)std::min<int64_t>(ntiles, 1 << 20)), dim3(kWave), 0, c->stream,
                       tree32(c), support(c->tree.kernel), o, d, nl, counts, offsets, idx);
    PMX_HIP(c, hipGetLastError());
    return PMX_OK;
}

}  // namespace pmx
