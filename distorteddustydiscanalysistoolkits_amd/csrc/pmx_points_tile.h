// pmx_points_tile.h — the FP32-compensated sample-tile procedure (one wavefront per tile of
// 64 Morton-ordered samples) and its helpers, shared by the per-call kernels and the batched
// persistent kernel.  Reference semantics: general_interpolation.jl:1-669 (accumulation and
// finalisation), LinearBVHQueries.jl:118-141 (nearest-h), LinearBVHTraversalMacros.jl:26-371
// (radius rules).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pmx_internal.h"
#include "pmx_tree32.h"

namespace pmx {
namespace pts {

constexpr int kWave = 64;
constexpr float kEps = 4e-6f;     // guard band of FP32 distance / radius comparisons
constexpr int kMaxF = 9;          // distinct field columns staged per leaf (points)
constexpr int kMaxD = 7;          // distinct derivative columns staged per leaf (1 + 3 + 3)
constexpr int kQueue = 128;       // accepted-pair queue of a sample tile (two chunks of 64)

// ---------------------------------------------------------------------------------------
// FP32 kernel shapes: f(q) and p(q) = f'(q) / q (kernels/M4_spline.jl .. C6_Wendland.jl).
// p is written without the cancelling differences of the reference's derivative forms
// (e.g. M4: -3/4 (2-q)^2 + 3 (1-q)^2 = q (-3 + 9/4 q)), so the FP32 gradient keeps its
// relative accuracy for close neighbours, and grad W = sigma h^-5 p(q) Delta needs no 1/r.
// ---------------------------------------------------------------------------------------
template <int K>
__device__ inline void shape_fp32(float q, float e, float& f, float& p) {
    // e = support - q, formed in FP64 by the caller: the edge factor (support - q) keeps its
    // relative accuracy for neighbours near the support radius, where FP32 q would not
    const float rq = q > 0.0f ? __builtin_amdgcn_rcpf(q) : 0.0f;
    if (K == kM4) {
        const float a = e, b = 1.0f - q;
        if (q < 1.0f) { f = 0.25f * (a * a * a) - b * b * b; p = -3.0f + 2.25f * q; }
        else if (q < 2.0f) { f = 0.25f * (a * a * a); p = (-0.75f * (a * a)) * rq; }
        else { f = 0.0f; p = 0.0f; }
    } else if (K == kM5) {
        const float a = e, b = 1.5f - q, c = 0.5f - q;
        const float a3 = a * a * a, b3 = b * b * b, c3 = c * c * c;
        if (q < 0.5f) { f = (a3 * a - 5.0f * (b3 * b)) + 10.0f * (c3 * c); p = -30.0f + 24.0f * (q * q); }
        else if (q < 1.5f) { f = a3 * a - 5.0f * (b3 * b); p = (-4.0f * a3 + 20.0f * b3) * rq; }
        else if (q < 2.5f) { f = a3 * a; p = (-4.0f * a3) * rq; }
        else { f = 0.0f; p = 0.0f; }
    } else if (K == kM6) {
        const float a = e, b = 2.0f - q, c = 1.0f - q;
        const float a4 = (a * a) * (a * a), b4 = (b * b) * (b * b), c4 = (c * c) * (c * c);
        if (q < 1.0f) { f = (a4 * a - 6.0f * (b4 * b)) + 15.0f * (c4 * c); p = -120.0f + (q * q) * (120.0f - 50.0f * q); }
        else if (q < 2.0f) { f = a4 * a - 6.0f * (b4 * b); p = (-5.0f * a4 + 30.0f * b4) * rq; }
        else if (q < 3.0f) { f = a4 * a; p = (-5.0f * a4) * rq; }
        else { f = 0.0f; p = 0.0f; }
    } else if (K == kC2) {
        if (q < 2.0f) {
            const float u = 0.5f * e, u3 = u * u * u;
            f = (u3 * u) * (2.0f * q + 1.0f);
            p = -5.0f * u3;
        } else { f = 0.0f; p = 0.0f; }
    } else if (K == kC4) {
        if (q < 2.0f) {
            const float u = 0.5f * e, u5 = (u * u) * (u * u) * u;
            f = (u5 * u) * (((35.0f / 12.0f) * (q * q) + 3.0f * q) + 1.0f);
            p = (-14.0f / 3.0f) * u5 * (1.0f + 2.5f * q);
        } else { f = 0.0f; p = 0.0f; }
    } else {
        if (q < 2.0f) {
            const float u = 0.5f * e, u7 = (u * u) * (u * u) * (u * u) * u;
            f = (u7 * u) * (((4.0f * (q * q * q) + 6.25f * (q * q)) + 4.0f * q) + 1.0f);
            p = -5.5f * u7 * ((4.0f * q + 3.5f) * q + 1.0f);
        } else { f = 0.0f; p = 0.0f; }
    }
}

// The same shapes in FP64 for derivative fields: their sums cancel (a nearly solenoidal
// velocity field has a divergence far below its individual pair terms), so the pair terms
// are formed in FP64.  1/q from the FP32 reciprocal and two FP64 Newton steps.
__device__ inline double rcp64(double x) {
    double y = (double)__builtin_amdgcn_rcpf((float)x);
    y = y * (2.0 - x * y);
    return y * (2.0 - x * y);
}

template <int K>
__device__ inline void shape_fp64(double q, double e, double& f, double& p) {
    const double rq = q > 0.0 ? rcp64(q) : 0.0;
    if (K == kM4) {
        const double a = e, b = 1.0 - q;
        if (q < 1.0) { f = 0.25 * (a * a * a) - b * b * b; p = -3.0 + 2.25 * q; }
        else if (q < 2.0) { f = 0.25 * (a * a * a); p = (-0.75 * (a * a)) * rq; }
        else { f = 0.0; p = 0.0; }
    } else if (K == kM5) {
        const double a = e, b = 1.5 - q, c = 0.5 - q;
        const double a3 = a * a * a, b3 = b * b * b, c3 = c * c * c;
        if (q < 0.5) { f = (a3 * a - 5.0 * (b3 * b)) + 10.0 * (c3 * c); p = -30.0 + 24.0 * (q * q); }
        else if (q < 1.5) { f = a3 * a - 5.0 * (b3 * b); p = (-4.0 * a3 + 20.0 * b3) * rq; }
        else if (q < 2.5) { f = a3 * a; p = (-4.0 * a3) * rq; }
        else { f = 0.0; p = 0.0; }
    } else if (K == kM6) {
        const double a = e, b = 2.0 - q, c = 1.0 - q;
        const double a4 = (a * a) * (a * a), b4 = (b * b) * (b * b), c4 = (c * c) * (c * c);
        if (q < 1.0) { f = (a4 * a - 6.0 * (b4 * b)) + 15.0 * (c4 * c); p = -120.0 + (q * q) * (120.0 - 50.0 * q); }
        else if (q < 2.0) { f = a4 * a - 6.0 * (b4 * b); p = (-5.0 * a4 + 30.0 * b4) * rq; }
        else if (q < 3.0) { f = a4 * a; p = (-5.0 * a4) * rq; }
        else { f = 0.0; p = 0.0; }
    } else if (K == kC2) {
        if (q < 2.0) {
            const double u = 0.5 * e, u3 = u * u * u;
            f = (u3 * u) * (2.0 * q + 1.0);
            p = -5.0 * u3;
        } else { f = 0.0; p = 0.0; }
    } else if (K == kC4) {
        if (q < 2.0) {
            const double u = 0.5 * e, u5 = (u * u) * (u * u) * u;
            f = (u5 * u) * (((35.0 / 12.0) * (q * q) + 3.0 * q) + 1.0);
            p = (-14.0 / 3.0) * u5 * (1.0 + 2.5 * q);
        } else { f = 0.0; p = 0.0; }
    } else {
        if (q < 2.0) {
            const double u = 0.5 * e, u7 = (u * u) * (u * u) * (u * u) * u;
            f = (u7 * u) * (((4.0 * (q * q * q) + 6.25 * (q * q)) + 4.0 * q) + 1.0);
            p = -5.5 * u7 * ((4.0 * q + 3.5) * q + 1.0);
        } else { f = 0.0; p = 0.0; }
    }
}

// ---------------------------------------------------------------------------------------
// Points
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

// conservative FP32 squared distance from a point (hi + lo) to an outward-rounded box,
// scaled down so it never exceeds the exact distance
__device__ inline float d2_box32(const PointLane& q, const Node32& b) {
    const float mn[3] = {b.mnx, b.mny, b.mnz}, mx[3] = {b.mxx, b.mxy, b.mxz};
    float s = 0.0f;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float lo_gap = (mn[k] - q.ph[k]) - q.pl[k];
        const float hi_gap = (q.ph[k] - mx[k]) + q.pl[k];
        const float t = lo_gap > 0.0f ? lo_gap : (hi_gap > 0.0f ? hi_gap : 0.0f);
        s += t * t;
    }
    return s * (1.0f - 2.0f * kEps);
}

// Exact reference point-leaf d2 in FP64 (_squared_distance_point_coords).
__device__ inline double d2_exact64(const double p[3], const T32& T, int64_t li) {
    double px, py, pz;
    pos64(T, li, px, py, pz);
    double s = 0.0, d;
    d = p[0] - px; s += d * d;
    d = p[1] - py; s += d * d;
    d = p[2] - pz; s += d * d;
    return s;
}

// Positions of a leaf run staged in LDS (CAP leaves at most).
template <int CAP>
struct BPosT {
    float x[CAP], y[CAP], z[CAP], h[CAP];
    float xl[CAP], yl[CAP], zl[CAP], invh[CAP];
};
using BPos = BPosT<kRunScalar>;

template <int CAP>
__device__ inline void stage_pos(const T32& T, BPosT<CAP>& P, int lo, int cnt) {
    const int lane = threadIdx.x;
    if (lane < cnt) {
        const float4 L = T.leaf[lo + lane];
        const float4 Lo = T.leaflo[lo + lane];
        P.x[lane] = L.x; P.y[lane] = L.y; P.z[lane] = L.z; P.h[lane] = L.w;
        P.xl[lane] = Lo.x; P.yl[lane] = Lo.y; P.zl[lane] = Lo.z; P.invh[lane] = Lo.w;
    }
}

template <int CAP>
__device__ inline float d2_pos32(const PointLane& q, const BPosT<CAP>& P, int j, float& dx, float& dy, float& dz) {
    dx = (q.ph[0] - P.x[j]) + (q.pl[0] - P.xl[j]);
    dy = (q.ph[1] - P.y[j]) + (q.pl[1] - P.yl[j]);
    dz = (q.ph[2] - P.z[j]) + (q.pl[2] - P.zl[j]);
    return dx * dx + dy * dy + dz * dz;
}

// Squared distance from a point to an FP32 node box (FP32, no guard: a heuristic only).
__device__ inline float d2_point_box_plain(const float c[3], const Node32& b) {
    const float mn[3] = {b.mnx, b.mny, b.mnz}, mx[3] = {b.mxx, b.mxy, b.mxz};
    float s = 0.0f;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float t = c[k] < mn[k] ? mn[k] - c[k] : (c[k] > mx[k] ? c[k] - mx[k] : 0.0f);
        s += t * t;
    }
    return s;
}

// The leaf run of at most T.bucket particles reached by descending from the root towards the
// point c, at each node into the child whose box is closer (a leaf child is its position).
// The right child of a node is the escape link of its left child.  Wave-uniform.
__device__ inline void seed_run(const T32& T, const float c[3], int& lo, int& cnt) {
    const int n = (int)T.n;
    int node = T.root;
    Node32 nd = T.nodes[node];
    for (int depth = 0; depth < 64; ++depth) {
        if (uniform(nd.cnt) <= T.bucket) break;
        const int l = uniform(nd.left);
        float dl, dr;
        int r;
        Node32 nl{}, nr{};
        if (l >= n) {
            const float4 L = T.leaf[l - n];
            const float p[3] = {L.x - c[0], L.y - c[1], L.z - c[2]};
            dl = p[0] * p[0] + p[1] * p[1] + p[2] * p[2];
            r = uniform(T.leaf_esc[l - n]);
        } else {
            nl = T.nodes[l];
            dl = d2_point_box_plain(c, nl);
            r = uniform(nl.escape);
        }
        if (r >= n) {
            const float4 R = T.leaf[r - n];
            const float p[3] = {R.x - c[0], R.y - c[1], R.z - c[2]};
            dr = p[0] * p[0] + p[1] * p[1] + p[2] * p[2];
        } else {
            nr = T.nodes[r];
            dr = d2_point_box_plain(c, nr);
        }
        const bool right = uniform(dr < dl ? 1 : 0) != 0;
        const int next = right ? r : l;
        if (next >= n) {                       // a leaf: the run of that one particle
            lo = next - n;
            cnt = 1;
            return;
        }
        nd = right ? nr : nl;
    }
    lo = uniform(nd.lo);
    cnt = uniform(nd.cnt);
}

// LBVH_find_nearest_h (LinearBVHQueries.jl:118-141) for a tile: exact FP64 argmin of d2,
// ties to the smaller sorted index.  The reference walks from the root with no bound and takes
// a strictly smaller distance; here the search first tests the leaf run nearest to the tile's
// centre (a greedy descent), so the bound prunes the walk from its first node.  A candidate
// replaces the best when its FP64 distance is smaller, or equal with a smaller index, so the
// result is the lexicographic minimum of (d2, index) whatever the visiting order; nodes at the
// best distance are never pruned (guard band), so ties are all seen.
template <int CAP>
__device__ inline int64_t nearest32(const T32& T, const PointLane& q, BPosT<CAP>& P) {
    int64_t best = -1;
    double best64 = __builtin_inf();
    float best32 = __builtin_inff();
    auto need = [&](const Node32& nd) {
        return q.valid && (best < 0 || !(d2_box32(q, nd) > best32 * (1.0f + kEps)));
    };
    auto bucket = [&](int lo, int cnt) {
        stage_pos(T, P, lo, cnt);
        __syncthreads();
        for (int j = 0; j < cnt; ++j) {
            float dx, dy, dz;
            const float d2 = d2_pos32(q, P, j, dx, dy, dz);
            const bool maybe = q.valid && (best < 0 || d2 <= best32 * (1.0f + kEps));
            if (ballot(maybe)) {
                if (maybe) {
                    const double e = d2_exact64(q.p64, T, lo + j);
                    if (e < best64 || (e == best64 && lo + j < best)) { best64 = e; best = lo + j; best32 = f32_up(e); }
                }
            }
        }
        __syncthreads();
    };
    if (!ballot(q.valid)) return -1;
    if (T.root != 0) {
        // centre of the tile's valid samples
        float lo3[3], hi3[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            lo3[k] = q.valid ? q.ph[k] : __builtin_inff();
            hi3[k] = q.valid ? q.ph[k] : -__builtin_inff();
        }
#pragma unroll
        for (int o = 1; o < kWave; o <<= 1) {
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                lo3[k] = fminf(lo3[k], __shfl_xor(lo3[k], o));
                hi3[k] = fmaxf(hi3[k], __shfl_xor(hi3[k], o));
            }
        }
        float c[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) c[k] = 0.5f * (lo3[k] + hi3[k]);      // equal on every lane
        int slo = 0, scnt = 0;
        seed_run(T, c, slo, scnt);
        bucket(slo, scnt);
    }
    walk(T, true, need, bucket);
    return best;
}

// Accept test of one staged leaf with the strategy's radius rule; band -> exact FP64.
template <int S>
__device__ inline bool accept_point(const T32& T, const PointLane& q, float d2, float hb32, int64_t li, float Kvf,
                                    double Kv, float radius2_32, double radius2_64) {
    const float rb = Kvf * hb32, r2b = rb * rb;
    const float r2 = S == kGather ? radius2_32 : (S == kScatter ? r2b : fmaxf(radius2_32, r2b));
    bool acc = q.valid && d2 < r2 * (1.0f - kEps);
    const bool band = q.valid && !acc && !(d2 > r2 * (1.0f + kEps));
    if (ballot(band)) {
        if (band) {
            const double e = d2_exact64(q.p64, T, li);
            acc = e <= radius2_rule<S>(Kv, radius2_64, h64(T, li));
        }
    }
    return acc;
}

template <int S>
__device__ inline bool node_need_point(const PointLane& q, const Node32& nd, float Kvf, float radius2_32) {
    const float rb = Kvf * nd.hmax, r2b = rb * rb;
    const float r2 = S == kGather ? radius2_32 : (S == kScatter ? r2b : fmaxf(radius2_32, r2b));
    return q.valid && !(d2_box32(q, nd) > r2 * (1.0f + kEps));
}

// The box of a tile's valid samples (FP64 coordinates rounded outwards to FP32) and the largest
// squared search radius of its samples (gather / symmetric; 0 for scatter).  Wave-uniform.
struct TileBox {
    float mn[3], mx[3];
    float r2;
};

template <int S>
__device__ inline TileBox tile_box(const PointLane& q, float radius2_32) {
    double lo[3], hi[3];
    float r2 = q.valid && S != kScatter ? radius2_32 : 0.0f;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        lo[k] = q.valid ? q.p64[k] : __builtin_inf();
        hi[k] = q.valid ? q.p64[k] : -__builtin_inf();
    }
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            lo[k] = fmin(lo[k], __shfl_xor(lo[k], o));
            hi[k] = fmax(hi[k], __shfl_xor(hi[k], o));
        }
        r2 = fmaxf(r2, __shfl_xor(r2, o));      // a lane without particle (NaN) loses to any number
    }
    TileBox B;
#pragma unroll
    for (int k = 0; k < 3; ++k) { B.mn[k] = f32_down(lo[k]); B.mx[k] = f32_up(hi[k]); }
    B.r2 = r2;
    return B;
}

// Can staged particle j be accepted by any sample of the tile?  Its distance from the tile's box
// against the largest radius a pair with it can have, conservative as d2_box32 (a rejected
// particle is further than its radius rule from every sample).
template <int S, int CAP>
__device__ inline bool reaches_tile(const TileBox& B, const BPosT<CAP>& P, int j, float Kvf) {
    float s = 0.0f;
    const float xh[3] = {P.x[j], P.y[j], P.z[j]}, xl[3] = {P.xl[j], P.yl[j], P.zl[j]};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float lo_gap = (B.mn[k] - xh[k]) - xl[k];
        const float hi_gap = (xh[k] - B.mx[k]) + xl[k];
        const float t = lo_gap > 0.0f ? lo_gap : (hi_gap > 0.0f ? hi_gap : 0.0f);
        s += t * t;
    }
    const float rb = Kvf * P.h[j], r2b = rb * rb;
    const float r2 = S == kGather ? B.r2 : (S == kScatter ? r2b : fmaxf(B.r2, r2b));
    return !(s * (1.0f - 2.0f * kEps) > r2 * (1.0f + kEps));
}

// Scalars are staged from the FP32 payload; the fields under a derivative are staged in FP64
// and each sample rounds A_b - A_ref to FP32 with its own FP64 reference A_ref (the shift
// leaves every derivative estimate unchanged and keeps the FP32 rounding relative to the
// local variation instead of the field's magnitude).
struct PointCat {
    int ns, ng, nd, nc;
    int scol[2], sh[2];   // payload column of each scalar (1.. = quant, 1+ncol = rho)
    int nder;
    int dcol[kMaxF];      // FP64 column of each derivative field (5 = rho, 6.. = quant)
    int gf;               // gradient -> derivative field
    int df[3];
    int cf[3];
    int os, og, od, oc;
    // derivative launches: the sums held per sample (LDS rows) and each logical sum's row
    // (-1: not held).  When the curl and divergence fields coincide, the curl's sums of A w
    // reuse the divergence's rows (cs_alias).
    int nacc, cs_alias;
    int amap[20];
};

// LDS of one sample tile: the staged payload of a leaf run, the tile's samples (read by the
// lanes that process their pairs), the queue of accepted (sample, particle) pairs and the
// FP64 sums of every sample.  Scalar catalogs (DERIV false) use a fixed layout; derivative
// catalogs (DERIV true) a layout sized by the catalog in dynamic LDS.
template <bool DERIV>
struct BTile;

template <>
struct BTile<false> {
    float w[kRunScalar];
    double ih64[kRunScalar];   // 1 / h in FP64
    float s[2][kRunScalar];
    float ph[3][kWave], pl[3][kWave];
    double invha[kWave];
    double acc[3][kWave];
    uint16_t pair[kQueue];
};

// Derivative tiles: views of the dynamic LDS, sized per launch by the catalog (deriv_tile_lds):
// staged m/rho, 1/h and derivative fields of a leaf run; the tile's samples; the held sums;
// the pair queue.
template <>
struct BTile<true> {
    double* w64;
    double* ih64;
    double* d;        // [nder][64]
    double* invha;    // [64]
    double* acc;      // [nacc][64]
    float* s;         // [2][64]
    float* ph;        // [3][64]
    float* pl;        // [3][64]
    uint16_t* pair;   // [kQueue]
    __device__ BTile(void* base, int nder, int nacc) {
        char* p = static_cast<char*>(base);
        w64 = reinterpret_cast<double*>(p); p += 8 * kRunDeriv;
        ih64 = reinterpret_cast<double*>(p); p += 8 * kRunDeriv;
        d = reinterpret_cast<double*>(p); p += 8 * kRunDeriv * (size_t)nder;
        invha = reinterpret_cast<double*>(p); p += 8 * kWave;
        acc = reinterpret_cast<double*>(p); p += 8 * kWave * (size_t)nacc;
        s = reinterpret_cast<float*>(p); p += 4 * 2 * kRunDeriv;
        ph = reinterpret_cast<float*>(p); p += 4 * 3 * kWave;
        pl = reinterpret_cast<float*>(p); p += 4 * 3 * kWave;
        pair = reinterpret_cast<uint16_t*>(p);
    }
};

__host__ __device__ inline size_t deriv_tile_lds(int nder, int nacc) {
    return 8 * (size_t)(2 * kRunDeriv + kRunDeriv * nder + kWave + kWave * nacc) + 4 * (size_t)(2 * kRunDeriv + 6 * kWave) +
           2 * (size_t)kQueue;
}


// accumulator slots: 0 = sum W (Shepard), 1.. = scalars, then the derivative sums
enum : int { kA_S1 = 0, kA_SC = 1, kA_B = 3, kA_GF = 6, kA_GS = 9, kA_DF = 10, kA_DS = 11, kA_CF = 14, kA_CS = 17 };
constexpr int kAccMax = 20;

// Splitting of heavy tiles (gather / symmetric).  A sample whose nearest particle has a large
// h gathers from a huge sphere (a halo sample above a dense midplane can take 10^5 - 10^6
// pairs), and one wavefront would walk it alone.  A tile that has tested `budget` leaves takes
// a split slot and hands every further needed subtree of at most `item_max` particles to an
// item queue instead of walking it; the items are walked by all wavefronts of a second
// kernel, which add their FP64 sums into the slot, and a third kernel finalises the slot's
// samples.  Slots or items beyond the capacities are walked in place.
struct PointSplit {
    int budget;                 // leaves a tile tests before it splits (0 = never split)
    int item_max;               // largest subtree handed over as one item
    int cap_items, cap_slots;
    unsigned int* counters;     // [0] items queued, [1] slots taken, [2] items started
    int2* items;                // {slot, subtree root}
    int* slot_tile;             // slot -> tile
    int* slot_unit;             // slot -> work unit (batches; may be null)
    int unit;                   // work unit of the tile being walked
    int* slot_near;             // [slot][64] nearest particle of each lane
    double* slot_acc;           // [slot][kAccMax][64] FP64 sums
};

// Tile walkers.  Both give the per-lane state of a tile walk and the steps of a staged leaf
// run: setup(nearest), need(node), run(lo, cnt), sum(a) (the lane's FP64 sums), spent.
template <int K, int S, bool DERIV>
struct PointWalk;

// Scalar catalogs: every lane tests the run's particles against its own sample and appends
// the accepted pairs to an LDS queue; the wavefront then evaluates 64 queued pairs at a time,
// one pair per lane, adding FP32 weights into the samples' FP64 LDS sums.  Only accepted pairs
// cost kernel evaluations, however unevenly they fall on the lanes.
// Separation |r| of a queued pair from its square: the hardware FP64 reciprocal square root
// (relative error ~2^-23) and one Newton step (~1e-14), instead of the IEEE square root's longer
// sequence; d2 + 1e-300 equals d2 for any d2 above ~1e-284, and a particle on the sample gets
// r ~ 1e-150 (q = 0 to the kernel).
__device__ inline double pair_dist(double d2) {
    const double dd = d2 + 1e-300;
    const double y = __builtin_amdgcn_rsq(dd);
    const double r0 = dd * y;
    return fma(fma(-r0, r0, dd), 0.5 * y, r0);
}

template <int K, int S>
struct PointWalk<K, S, false> {
    static constexpr int kAcc = 3;
    const T32& T;
    const PointCat& cat;
    BPos& P;
    BTile<false>& Y;
    PointLane q;
    int lane;
    double Kv, radius2_64;
    float Kvf, sig, radius2_32;
    int qhead, qtail;           // wave-uniform queue positions (pairs appended / processed)
    int spent;                  // leaves tested so far (wave-uniform)
    TileBox box;                // the tile's samples and their largest radius (wave-uniform)

    __device__ PointWalk(const T32& T_, const PointCat& cat_, BPos& P_, BTile<false>& Y_, const PointLane& q_)
        : T(T_), cat(cat_), P(P_), Y(Y_), q(q_), lane(threadIdx.x), Kv(support(K)), radius2_64(0.0),
          Kvf((float)support(K)), sig(norm3<float>(K)), radius2_32(0.0f), qhead(0), qtail(0), spent(0) {}

    // lane state in LDS; nearest < 0 for scatter (or no particle)
    __device__ void setup(int64_t nearest) {
        double invha64 = 0.0;
        if (S != kScatter) {
            const double ha64 = nearest < 0 ? __builtin_nan("") : h64(T, nearest);
            const double radius = Kv * ha64;
            radius2_64 = radius * radius;
            radius2_32 = (float)radius2_64;
            invha64 = 1.0 / ha64;
        }
        box = tile_box<S>(q, radius2_32);
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 3; ++k) { Y.ph[k][lane] = q.ph[k]; Y.pl[k][lane] = q.pl[k]; }
        Y.invha[lane] = invha64;
#pragma unroll
        for (int a = 0; a < kAcc; ++a) Y.acc[a][lane] = 0.0;
    }

    __device__ bool need(const Node32& nd) const { return node_need_point<S>(q, nd, Kvf, radius2_32); }


    __device__ double sum(int a) const { return Y.acc[a][lane]; }

    // one pair per lane: pairs qhead .. qhead + cnt - 1 of the queue
    __device__ void process(int cnt) {
        __syncthreads();
        if (lane < cnt) {
            const int pr = Y.pair[(qhead + lane) & (kQueue - 1)];
            const int sa = pr & (kWave - 1), j = pr >> 6;
            const float qh0 = Y.ph[0][sa], qh1 = Y.ph[1][sa], qh2 = Y.ph[2][sa];
            const float ql0 = Y.pl[0][sa], ql1 = Y.pl[1][sa], ql2 = Y.pl[2][sa];
            // separation in FP64 from the hi / lo parts (both differences exact in FP64)
            const double ex = ((double)qh0 - (double)P.x[j]) + ((double)ql0 - (double)P.xl[j]);
            const double ey = ((double)qh1 - (double)P.y[j]) + ((double)ql1 - (double)P.yl[j]);
            const double ez = ((double)qh2 - (double)P.z[j]) + ((double)ql2 - (double)P.zl[j]);
            const double rd = pair_dist(ex * ex + ey * ey + ez * ez);
            const float invh = P.invh[j];
            float W;
            {
                const double qb = rd * Y.ih64[j];
                float fb, pb;
                shape_fp32<K>((float)qb, (float)(Kv - qb), fb, pb);
                const float Wb = ((invh * invh) * invh) * (sig * fb);
                if (S == kScatter) {
                    W = Wb;
                } else {
                    const double ia64 = Y.invha[sa];
                    const float ia = (float)ia64;
                    const double qa = rd * ia64;
                    float fa, pa;
                    shape_fp32<K>((float)qa, (float)(Kv - qa), fa, pa);
                    const float Wa = (ia * ia * ia) * (sig * fa);
                    W = S == kGather ? Wa : 0.5f * (Wa + Wb);
                }
            }
            const float w = Y.w[j] * W;
            atomicAdd(&Y.acc[kA_S1][sa], (double)w);
            if (cat.ns > 0) atomicAdd(&Y.acc[kA_SC][sa], (double)(Y.s[0][j] * w));
            if (cat.ns > 1) atomicAdd(&Y.acc[kA_SC + 1][sa], (double)(Y.s[1][j] * w));
        }
        qhead += cnt;
        __syncthreads();
    }

    // a run of leaves lo .. lo+cnt-1: stage; test the particles that can reach the tile's box
    // against every lane's sample; queue the accepted pairs and process them 64 at a time
    __device__ __attribute__((always_inline)) void run(int lo, int cnt) {
        const int64_t n = T.n;
        stage_pos(T, P, lo, cnt);
        if (lane < cnt) {
            const int64_t li = lo + lane;
            Y.w[lane] = T.pay[li];
            Y.ih64[lane] = T.ih64[li];
            for (int f = 0; f < cat.ns; ++f) Y.s[f][lane] = T.pay[(int64_t)cat.scol[f] * n + li];
        }
        __syncthreads();
        const uint64_t cand = ballot(lane < cnt && reaches_tile<S>(box, P, lane, Kvf));
        for (uint64_t cm = cand; cm; cm &= cm - 1) {
            const int j = __builtin_ctzll(cm);
            float dx, dy, dz;
            const float d2 = d2_pos32(q, P, j, dx, dy, dz);
            const bool acc = accept_point<S>(T, q, d2, P.h[j], lo + j, Kvf, Kv, radius2_32, radius2_64);
            const uint64_t m = ballot(acc);
            if (!m) continue;
            if (acc) {
                const int at = (qtail + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                                       __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u))) &
                               (kQueue - 1);
                Y.pair[at] = (uint16_t)(lane | (j << 6));
            }
            qtail += __builtin_popcountll(m);
            if (qtail - qhead >= kWave) process(kWave);
        }
        if (qtail > qhead) process(qtail - qhead);
        spent += cnt;
    }
};

// Derivative catalogs: the same pair queue with FP64 pair terms (W, sigma h^-5 f'(q)/q, the
// separation and the fields; FP32 terms left 4e-6 on a nearly solenoidal velocity field,
// whose divergence is far below its pair terms) into FP64 LDS sums.  With FP64 terms the
// sums need no per-sample field reference: the cancellation of f - A~ b costs digits of an
// FP64 result.  Keeping the sums in registers instead (each lane evaluating its own accepted
// particles) needs 255 VGPRs, one wavefront per SIMD; the LDS holds only the sums the catalog
// uses.
template <int K, int S>
struct PointWalk<K, S, true> {
    static constexpr int kAcc = kAccMax;
    const T32& T;
    const PointCat& cat;
    BPosT<kRunDeriv>& P;
    BTile<true>& Y;
    PointLane q;
    int lane;
    double Kv, radius2_64;
    float Kvf, radius2_32;
    int qhead, qtail;           // wave-uniform queue positions (pairs appended / processed)
    int spent;                  // leaves tested so far (wave-uniform)
    TileBox box;                // the tile's samples and their largest radius (wave-uniform)

    __device__ PointWalk(const T32& T_, const PointCat& cat_, BPosT<kRunDeriv>& P_, BTile<true>& Y_, const PointLane& q_)
        : T(T_), cat(cat_), P(P_), Y(Y_), q(q_), lane(threadIdx.x), Kv(support(K)), radius2_64(0.0),
          Kvf((float)support(K)), radius2_32(0.0f), qhead(0), qtail(0), spent(0) {}

    __device__ void setup(int64_t nearest) {
        double invha64 = 0.0;
        if (S != kScatter) {
            const double ha64 = nearest < 0 ? __builtin_nan("") : h64(T, nearest);
            const double radius = Kv * ha64;
            radius2_64 = radius * radius;
            radius2_32 = (float)radius2_64;
            invha64 = 1.0 / ha64;
        }
        box = tile_box<S>(q, radius2_32);
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 3; ++k) { Y.ph[k * kWave + lane] = q.ph[k]; Y.pl[k * kWave + lane] = q.pl[k]; }
        Y.invha[lane] = invha64;
        for (int a = 0; a < cat.nacc; ++a) Y.acc[a * kWave + lane] = 0.0;
    }

    __device__ bool need(const Node32& nd) const { return node_need_point<S>(q, nd, Kvf, radius2_32); }


    __device__ double sum(int a) const { return cat.amap[a] < 0 ? 0.0 : Y.acc[cat.amap[a] * kWave + lane]; }

    __device__ void add(int a, int sa, double v) { atomicAdd(&Y.acc[cat.amap[a] * kWave + sa], v); }

    __device__ void process(int cnt) {
        __syncthreads();
        if (lane < cnt) {
            const int pr = Y.pair[(qhead + lane) & (kQueue - 1)];
            const int sa = pr & (kWave - 1), j = pr >> 6;
            const float qh0 = Y.ph[sa], qh1 = Y.ph[kWave + sa], qh2 = Y.ph[2 * kWave + sa];
            const float ql0 = Y.pl[sa], ql1 = Y.pl[kWave + sa], ql2 = Y.pl[2 * kWave + sa];
            // separation in FP64 from the hi / lo parts (both differences exact in FP64)
            const double ex = ((double)qh0 - (double)P.x[j]) + ((double)ql0 - (double)P.xl[j]);
            const double ey = ((double)qh1 - (double)P.y[j]) + ((double)ql1 - (double)P.yl[j]);
            const double ez = ((double)qh2 - (double)P.z[j]) + ((double)ql2 - (double)P.zl[j]);
            const double rd = pair_dist(ex * ex + ey * ey + ez * ez);
            const double sig64 = norm3<double>(K);
            double W, G;
            {
                const double ihb = Y.ih64[j];
                const double qb = rd * ihb;
                double fb, pb;
                shape_fp64<K>(qb, Kv - qb, fb, pb);
                const double ib2 = ihb * ihb;
                const double Wb = (ib2 * ihb) * (sig64 * fb);
                const double Gb = ((ib2 * ib2) * ihb) * (sig64 * pb);
                if (S == kScatter) {
                    W = Wb; G = Gb;
                } else {
                    const double ia = Y.invha[sa];
                    const double qa = rd * ia;
                    double fa, pa;
                    shape_fp64<K>(qa, Kv - qa, fa, pa);
                    const double ia2 = ia * ia;
                    const double Wa = (ia2 * ia) * (sig64 * fa);
                    const double Ga = ((ia2 * ia2) * ia) * (sig64 * pa);
                    if (S == kGather) { W = Wa; G = Ga; }
                    else { W = 0.5 * (Wa + Wb); G = 0.5 * (Ga + Gb); }
                }
            }
            const double wb = Y.w64[j];
            const double w = wb * W;
            add(kA_S1, sa, w);
            if (cat.ns > 0) add(kA_SC, sa, (double)Y.s[j] * w);
            if (cat.ns > 1) add(kA_SC + 1, sa, (double)Y.s[kRunDeriv + j] * w);
            // (m/rho) grad W = (m/rho) sigma h^-5 (f'(q)/q) Delta
            const double cg = wb * G;
            const double gx = cg * ex, gy = cg * ey, gz = cg * ez;
            add(kA_B, sa, gx);
            add(kA_B + 1, sa, gy);
            add(kA_B + 2, sa, gz);
            if (cat.ng) {
                const double A = Y.d[cat.gf * kRunDeriv + j];
                add(kA_GF, sa, A * gx);
                add(kA_GF + 1, sa, A * gy);
                add(kA_GF + 2, sa, A * gz);
                add(kA_GS, sa, A * w);
            }
            if (cat.nd) {
                const double Ax = Y.d[cat.df[0] * kRunDeriv + j], Ay = Y.d[cat.df[1] * kRunDeriv + j],
                             Az = Y.d[cat.df[2] * kRunDeriv + j];
                add(kA_DF, sa, (Ax * gx + Ay * gy) + Az * gz);
                add(kA_DS, sa, Ax * w);
                add(kA_DS + 1, sa, Ay * w);
                add(kA_DS + 2, sa, Az * w);
            }
            if (cat.nc) {
                const double Ax = Y.d[cat.cf[0] * kRunDeriv + j], Ay = Y.d[cat.cf[1] * kRunDeriv + j],
                             Az = Y.d[cat.cf[2] * kRunDeriv + j];
                add(kA_CF, sa, Ay * gz - Az * gy);
                add(kA_CF + 1, sa, Az * gx - Ax * gz);
                add(kA_CF + 2, sa, Ax * gy - Ay * gx);
                if (!cat.cs_alias) {
                    add(kA_CS, sa, Ax * w);
                    add(kA_CS + 1, sa, Ay * w);
                    add(kA_CS + 2, sa, Az * w);
                }
            }
        }
        qhead += cnt;
        __syncthreads();
    }

    __device__ __attribute__((always_inline)) void run(int lo, int cnt) {
        const int64_t n = T.n;
        stage_pos(T, P, lo, cnt);
        if (lane < cnt) {
            const int64_t li = lo + lane;
            Y.w64[lane] = T.w64[li];
            Y.ih64[lane] = T.ih64[li];
            for (int f = 0; f < cat.ns; ++f) Y.s[f * kRunDeriv + lane] = T.pay[(int64_t)cat.scol[f] * n + li];
            for (int f = 0; f < cat.nder; ++f)
                Y.d[f * kRunDeriv + lane] = col64(T, cat.dcol[f], li);
        }
        __syncthreads();
        const uint64_t cand = ballot(lane < cnt && reaches_tile<S>(box, P, lane, Kvf));
        for (uint64_t cm = cand; cm; cm &= cm - 1) {
            const int j = __builtin_ctzll(cm);
            float dx, dy, dz;
            const float d2 = d2_pos32(q, P, j, dx, dy, dz);
            const bool acc = accept_point<S>(T, q, d2, P.h[j], lo + j, Kvf, Kv, radius2_32, radius2_64);
            const uint64_t m = ballot(acc);
            if (!m) continue;
            if (acc) {
                const int at = (qtail + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                                       __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u))) &
                               (kQueue - 1);
                Y.pair[at] = (uint16_t)(lane | (j << 6));
            }
            qtail += __builtin_popcountll(m);
            if (qtail - qhead >= kWave) process(kWave);
        }
        if (qtail > qhead) process(qtail - qhead);
        spent += cnt;
    }
};

// FP64 finalisation of one sample (general_interpolation.jl:574-668) from its sums acc(a).
template <bool DERIV, class Acc>
__device__ inline void finalise_point(const PointCat& cat, Acc&& acc, float* __restrict__ out, int64_t ldo, int64_t qi) {
    const double s1 = acc(kA_S1);
    if (s1 == 0.0) {
        const float nan = __builtin_nanf("");
        for (int j = 0; j < cat.ns; ++j) out[(cat.os + j) * ldo + qi] = nan;
        for (int j = 0; j < 3 * cat.ng; ++j) out[(cat.og + j) * ldo + qi] = nan;
        for (int j = 0; j < cat.nd; ++j) out[(cat.od + j) * ldo + qi] = nan;
        for (int j = 0; j < 3 * cat.nc; ++j) out[(cat.oc + j) * ldo + qi] = nan;
        return;
    }
    const double inv = 1.0 / s1;
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        if (j >= cat.ns) break;
        const double v = acc(kA_SC + j);
        out[(cat.os + j) * ldo + qi] = (float)(cat.sh[j] ? v * inv : v);
    }
    if (!DERIV) return;
    const double bx = acc(kA_B), by = acc(kA_B + 1), bz = acc(kA_B + 2);
    if (cat.ng) {
        const double A = acc(kA_GS) * inv;
        out[cat.og * ldo + qi] = (float)(acc(kA_GF) - bx * A);
        out[(cat.og + 1) * ldo + qi] = (float)(acc(kA_GF + 1) - by * A);
        out[(cat.og + 2) * ldo + qi] = (float)(acc(kA_GF + 2) - bz * A);
    }
    if (cat.nd) {
        const double Ax = acc(kA_DS) * inv, Ay = acc(kA_DS + 1) * inv, Az = acc(kA_DS + 2) * inv;
        out[cat.od * ldo + qi] = (float)(acc(kA_DF) - (Ax * bx + Ay * by + Az * bz));
    }
    if (cat.nc) {
        const double Ax = acc(kA_CS) * inv, Ay = acc(kA_CS + 1) * inv, Az = acc(kA_CS + 2) * inv;
        const double ex = Ay * bz - Az * by, ey = Az * bx - Ax * bz, ez = Ax * by - Ay * bx;
        out[cat.oc * ldo + qi] = (float)(-(acc(kA_CF) - ex));
        out[(cat.oc + 1) * ldo + qi] = (float)(-(acc(kA_CF + 1) - ey));
        out[(cat.oc + 2) * ldo + qi] = (float)(-(acc(kA_CF + 2) - ez));
    }
}

__device__ inline unsigned int lane0_atomic_add(unsigned int* p, unsigned int v) {
    unsigned int r = 0;
    if (threadIdx.x == 0) r = atomicAdd(p, v);
    return (unsigned int)__builtin_amdgcn_readfirstlane((int)r);
}

// One tile of 64 samples (a wavefront; perm gives the samples of tile t).  Each lane first
// finds its nearest-h particle (gather / symmetric).  The union walk
// then stages leaf runs in LDS; every lane tests the run's particles against its own sample
// (FP32 with a guard band, exact FP64 inside the band) and appends the accepted pairs to a
// queue, and the wavefront works through the queue 64 pairs at a time, one pair per lane,
// adding the FP32 contributions into the samples' FP64 accumulators (LDS atomics).  Only
// accepted pairs cost kernel evaluations, however unevenly they fall on the lanes.  A heavy
// tile may split (PointSplit); otherwise the FP64 finalisation writes the outputs.
template <int K, int S, bool DERIV>
__device__ inline void points_tile(const T32& T, const PointCat& cat, const double* __restrict__ qx,
                                   const double* __restrict__ qy, const double* __restrict__ qz,
                                   const int32_t* __restrict__ perm, int64_t npts, float* __restrict__ out,
                                   int64_t ldo, int64_t tile, BPosT<run_cap(DERIV)>& P,
                                   BTile<DERIV>& Y, const PointSplit& sp) {
    const int lane = threadIdx.x;
    const int64_t slot = tile * kWave + lane;
    const bool valid = slot < npts;
    const int64_t qi = valid ? perm[slot] : 0;
    const PointLane q = point_lane(qx, qy, qz, valid, qi);
    int64_t nearest = -1;
    if (S != kScatter) nearest = nearest32(T, q, P);

    PointWalk<K, S, DERIV> W(T, cat, P, Y, q);
    W.setup(nearest);
    int split = -1;             // split slot; -2 once no slot was free
    auto need = [&](const Node32& nd) { return W.need(nd); };
    auto bucket = [&](int lo, int cnt) { W.run(lo, cnt); };
    auto take = [&](const Node32&, int cnt) { return cnt <= T.bucket; };
    auto emit = [&](int node, int cnt) -> bool {
        if (S == kScatter || sp.budget <= 0 || cnt > sp.item_max || split == -2) return false;
        if (split < 0) {
            if (W.spent < sp.budget) return false;
            const unsigned int s_ = lane0_atomic_add(&sp.counters[1], 1u);
            if (s_ >= (unsigned int)sp.cap_slots) { split = -2; return false; }
            split = (int)s_;
        }
        const unsigned int it = lane0_atomic_add(&sp.counters[0], 1u);
        if (it >= (unsigned int)sp.cap_items) return false;
        if (lane == 0) sp.items[it] = make_int2(split, node);
        return true;
    };
    if (ballot(valid)) {
        if (T.root == 0) bucket(0, (int)T.n);
        else walk_sub(T, T.root, 0, need, bucket, take, emit);
    }
    __syncthreads();
    if (split >= 0) {
        // the items add to these sums; the finishing kernel writes the outputs
        double* acc = sp.slot_acc + (int64_t)split * kAccMax * kWave;
#pragma unroll
        for (int a = 0; a < PointWalk<K, S, DERIV>::kAcc; ++a) acc[a * kWave + lane] = W.sum(a);
        sp.slot_near[(int64_t)split * kWave + lane] = valid ? (int)nearest : -1;
        if (lane == 0) {
            sp.slot_tile[split] = (int)tile;
            if (sp.slot_unit) sp.slot_unit[split] = sp.unit;
        }
    } else if (valid) {
        finalise_point<DERIV>(cat, [&](int a) { return W.sum(a); }, out, ldo, qi);
    }
    __syncthreads();
}

// One queued item of a split tile: the walk of one subtree for the tile's 64 samples; the
// sums are added into the slot (FP64 atomics).
template <int K, int S, bool DERIV>
__device__ inline void points_item(const T32& T, const PointCat& cat, const double* __restrict__ qx,
                                   const double* __restrict__ qy, const double* __restrict__ qz,
                                   const int32_t* __restrict__ perm, int64_t npts, const PointSplit& sp,
                                   int64_t item, BPosT<run_cap(DERIV)>& P, BTile<DERIV>& Y) {
    const int lane = threadIdx.x;
    const int2 it = sp.items[item];
    const int split = uniform(it.x), node = uniform(it.y);
    const int64_t tile = uniform(sp.slot_tile[split]);
    const int64_t slot = tile * kWave + lane;
    const bool valid = slot < npts;
    const int64_t qi = valid ? perm[slot] : 0;
    const PointLane q = point_lane(qx, qy, qz, valid, qi);
    const int64_t nearest = valid ? (int64_t)sp.slot_near[(int64_t)split * kWave + lane] : -1;
    PointWalk<K, S, DERIV> W(T, cat, P, Y, q);
    W.setup(nearest);
    auto need = [&](const Node32& nd) { return W.need(nd); };
    auto bucket = [&](int lo, int cnt) { W.run(lo, cnt); };
    auto take = [&](const Node32&, int cnt) { return cnt <= T.bucket; };
    const int stop = uniform(T.nodes[node].escape);
    walk_sub(T, node, stop, need, bucket, take, [](int, int) { return false; });
    __syncthreads();
    double* acc = sp.slot_acc + (int64_t)split * kAccMax * kWave;
#pragma unroll
    for (int a = 0; a < PointWalk<K, S, DERIV>::kAcc; ++a) {
        const double v = W.sum(a);
        if (v != 0.0) unsafeAtomicAdd(&acc[a * kWave + lane], v);
    }
    __syncthreads();
}

// Outputs of one split slot from its FP64 sums.
template <bool DERIV>
__device__ inline void points_finish(const PointCat& cat, const int32_t* __restrict__ perm, int64_t npts,
                                     float* __restrict__ out, int64_t ldo, const PointSplit& sp, int split) {
    const int lane = threadIdx.x;
    const int64_t tile = sp.slot_tile[split];
    const int64_t slot = tile * kWave + lane;
    if (slot >= npts) return;
    const int64_t qi = perm[slot];
    const double* acc = sp.slot_acc + (int64_t)split * kAccMax * kWave;
    finalise_point<DERIV>(cat, [&](int a) { return acc[a * kWave + lane]; }, out, ldo, qi);
}

}  // namespace pts

// The derivative / scalar launches one catalog chunk is split into (pmx_interp32.hip).
std::vector<pts::PointCat> point_cats(const CatChunk& ch);

}  // namespace pmx
