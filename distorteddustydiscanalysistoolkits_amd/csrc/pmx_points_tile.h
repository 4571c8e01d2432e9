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
constexpr int kMaxBucket = 64;
constexpr float kEps = 4e-6f;     // guard band of FP32 distance / radius comparisons
constexpr int kMaxF = 9;          // distinct field columns staged per leaf (points)
constexpr int kMaxD = 7;          // distinct derivative columns staged per leaf (1 + 3 + 3)

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
    const int64_t n = T.n;
    double s = 0.0, d;
    d = p[0] - T.cols64[li]; s += d * d;
    d = p[1] - T.cols64[n + li]; s += d * d;
    d = p[2] - T.cols64[2 * n + li]; s += d * d;
    return s;
}

// Positions of a bucket staged in LDS.
struct BPos {
    float x[kMaxBucket], y[kMaxBucket], z[kMaxBucket], h[kMaxBucket];
    float xl[kMaxBucket], yl[kMaxBucket], zl[kMaxBucket], invh[kMaxBucket];
};

__device__ inline void stage_pos(const T32& T, BPos& P, int lo, int cnt) {
    const int lane = threadIdx.x;
    if (lane < cnt) {
        const float4 L = T.leaf[lo + lane];
        const float4 Lo = T.leaflo[lo + lane];
        P.x[lane] = L.x; P.y[lane] = L.y; P.z[lane] = L.z; P.h[lane] = L.w;
        P.xl[lane] = Lo.x; P.yl[lane] = Lo.y; P.zl[lane] = Lo.z; P.invh[lane] = Lo.w;
    }
}

__device__ inline float d2_pos32(const PointLane& q, const BPos& P, int j, float& dx, float& dy, float& dz) {
    dx = (q.ph[0] - P.x[j]) + (q.pl[0] - P.xl[j]);
    dy = (q.ph[1] - P.y[j]) + (q.pl[1] - P.yl[j]);
    dz = (q.ph[2] - P.z[j]) + (q.pl[2] - P.zl[j]);
    return dx * dx + dy * dy + dz * dz;
}

// LBVH_find_nearest_h (LinearBVHQueries.jl:118-141) for a tile: exact FP64 argmin of d2,
// ties to the smaller sorted index (leaves come in increasing index order and only a
// strictly smaller FP64 distance replaces the best).
__device__ inline int64_t nearest32(const T32& T, const PointLane& q, BPos& P) {
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
                    if (e < best64) { best64 = e; best = lo + j; best32 = f32_up(e); }
                }
            }
        }
        __syncthreads();
    };
    walk(T, ballot(q.valid) != 0, need, bucket);
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
            acc = e <= radius2_rule<S>(Kv, radius2_64, T.cols64[4 * T.n + li]);
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
};

struct BPay {
    float w[kMaxBucket];
    double ih64[kMaxBucket];   // 1 / h in FP64
    float s[2][kMaxBucket];
    double d[kMaxD][kMaxBucket];
};

// One tile of 64 Morton-ordered samples (a wavefront): nearest-h (gather / symmetric), the
// union walk, FP32 compensated accumulation and the FP64 finalisation.  P and Y are the
// wavefront's LDS staging areas.
template <int K, int S, bool DERIV>
__device__ inline void points_tile(const T32& T, const PointCat& cat, const double* __restrict__ qx,
                                   const double* __restrict__ qy, const double* __restrict__ qz,
                                   const int32_t* __restrict__ perm, int64_t npts, float* __restrict__ out,
                                   int64_t ldo, int64_t tile, BPos& P, BPay& Y) {
    const int lane = threadIdx.x;
    const double Kv = support(K);
    const float Kvf = (float)Kv;
    const float sig = norm3<float>(K);
    const int64_t n = T.n;

    const int64_t slot = tile * kWave + lane;
    const bool valid = slot < npts;
    const int64_t qi = valid ? perm[slot] : 0;
    const PointLane q = point_lane(qx, qy, qz, valid, qi);

    double radius2_64 = 0.0, invha64 = 0.0;
    float radius2_32 = 0.0f, invha = 0.0f;
    int64_t nearest = -1;
    if (S != kScatter) {
        nearest = nearest32(T, q, P);
        const double ha64 = nearest < 0 ? __builtin_nan("") : T.cols64[4 * n + nearest];
        const double radius = Kv * ha64;
        radius2_64 = radius * radius;
        radius2_32 = (float)radius2_64;
        invha64 = 1.0 / ha64;
        invha = (float)invha64;
    }
    const float invha3 = invha * invha * invha;
    const float invha5 = (invha * invha) * (invha * invha) * invha;

    CSum S1{0, 0}, sc[2] = {}, b[3] = {}, gf[3] = {}, gs{0, 0}, df{0, 0}, ds[3] = {}, cf[3] = {}, cs[3] = {};
    // per-sample FP64 reference values of the derivative fields: the nearest particle
    // (gather / symmetric) or the first accepted neighbour (scatter)
    double refg = 0.0, refd[3] = {0.0, 0.0, 0.0}, refc[3] = {0.0, 0.0, 0.0};
    bool have_ref = false;
    if (DERIV && nearest >= 0) {
        if (cat.ng) refg = T.cols64[(int64_t)cat.dcol[cat.gf] * n + nearest];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            if (cat.nd) refd[k] = T.cols64[(int64_t)cat.dcol[cat.df[k]] * n + nearest];
            if (cat.nc) refc[k] = T.cols64[(int64_t)cat.dcol[cat.cf[k]] * n + nearest];
        }
        have_ref = true;
    }

    auto need = [&](const Node32& nd) { return node_need_point<S>(q, nd, Kvf, radius2_32); };
    auto bucket = [&](int lo, int cnt) {
        stage_pos(T, P, lo, cnt);
        if (lane < cnt) {
            const int64_t li = lo + lane;
            Y.w[lane] = T.pay[li];
            Y.ih64[lane] = 1.0 / T.cols64[4 * n + li];
            for (int f = 0; f < cat.ns; ++f) Y.s[f][lane] = T.pay[(int64_t)cat.scol[f] * n + li];
            if (DERIV)
                for (int f = 0; f < cat.nder; ++f) Y.d[f][lane] = T.cols64[(int64_t)cat.dcol[f] * n + li];
        }
        __syncthreads();
        for (int j = 0; j < cnt; ++j) {
            float dx, dy, dz;
            const float d2 = d2_pos32(q, P, j, dx, dy, dz);
            if (!accept_point<S>(T, q, d2, P.h[j], lo + j, Kvf, Kv, radius2_32, radius2_64)) continue;
            const float invh = P.invh[j];
            // separation in FP64 from the hi / lo parts (both differences exact in FP64)
            const double ex = ((double)q.ph[0] - (double)P.x[j]) + ((double)q.pl[0] - (double)P.xl[j]);
            const double ey = ((double)q.ph[1] - (double)P.y[j]) + ((double)q.pl[1] - (double)P.yl[j]);
            const double ez = ((double)q.ph[2] - (double)P.z[j]) + ((double)q.pl[2] - (double)P.zl[j]);
            const double rd = sqrt(ex * ex + ey * ey + ez * ez);
            float W, G;       // W and sigma h^-5 f'(q)/q
            {
                const double qb = rd * Y.ih64[j];
                float fb, pb;
                shape_fp32<K>((float)qb, (float)(Kv - qb), fb, pb);
                const float ib2 = invh * invh;
                const float Wb = (ib2 * invh) * (sig * fb);
                const float Gb = ((ib2 * ib2) * invh) * (sig * pb);
                if (S == kScatter) {
                    W = Wb; G = Gb;
                } else {
                    const double qa = rd * invha64;
                    float fa, pa;
                    shape_fp32<K>((float)qa, (float)(Kv - qa), fa, pa);
                    const float Wa = invha3 * (sig * fa);
                    const float Ga = invha5 * (sig * pa);
                    if (S == kGather) { W = Wa; G = Ga; }
                    else { W = 0.5f * (Wa + Wb); G = 0.5f * (Ga + Gb); }
                }
            }
            const float wb = Y.w[j];
            const float w = wb * W;
            S1.add(w);
            if (cat.ns > 0) sc[0].add(Y.s[0][j] * w);
            if (cat.ns > 1) sc[1].add(Y.s[1][j] * w);
            if (!DERIV) continue;
            if (!have_ref) {
                if (cat.ng) refg = Y.d[cat.gf][j];
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    if (cat.nd) refd[k] = Y.d[cat.df[k]][j];
                    if (cat.nc) refc[k] = Y.d[cat.cf[k]][j];
                }
                have_ref = true;
            }
            // (m/rho) grad W = (m/rho) sigma h^-5 (f'(q)/q) Delta
            const float cg = wb * G;
            const float gx = cg * dx, gy = cg * dy, gz = cg * dz;
            b[0].add(gx); b[1].add(gy); b[2].add(gz);
            if (cat.ng) {
                const float A = (float)(Y.d[cat.gf][j] - refg);
                gf[0].add(A * gx); gf[1].add(A * gy); gf[2].add(A * gz);
                gs.add(A * w);
            }
            if (cat.nd) {
                const float Ax = (float)(Y.d[cat.df[0]][j] - refd[0]), Ay = (float)(Y.d[cat.df[1]][j] - refd[1]),
                            Az = (float)(Y.d[cat.df[2]][j] - refd[2]);
                df.add(Ax * gx + Ay * gy + Az * gz);
                ds[0].add(Ax * w); ds[1].add(Ay * w); ds[2].add(Az * w);
            }
            if (cat.nc) {
                const float Ax = (float)(Y.d[cat.cf[0]][j] - refc[0]), Ay = (float)(Y.d[cat.cf[1]][j] - refc[1]),
                            Az = (float)(Y.d[cat.cf[2]][j] - refc[2]);
                cf[0].add(Ay * gz - Az * gy);
                cf[1].add(Az * gx - Ax * gz);
                cf[2].add(Ax * gy - Ay * gx);
                cs[0].add(Ax * w); cs[1].add(Ay * w); cs[2].add(Az * w);
            }
        }
        __syncthreads();
    };
    walk(T, ballot(valid) != 0, need, bucket);

    // finalisation (general_interpolation.jl:574-668) in FP64 from the compensated sums
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
            if (cat.ns > 0) out[cat.os * ldo + qi] = (float)(cat.sh[0] ? sc[0].value() * inv : sc[0].value());
            if (cat.ns > 1) out[(cat.os + 1) * ldo + qi] = (float)(cat.sh[1] ? sc[1].value() * inv : sc[1].value());
            if (DERIV) {
                const double bx = b[0].value(), by = b[1].value(), bz = b[2].value();
                if (cat.ng) {
                    const double A = gs.value() * inv;
                    out[cat.og * ldo + qi] = (float)(gf[0].value() - bx * A);
                    out[(cat.og + 1) * ldo + qi] = (float)(gf[1].value() - by * A);
                    out[(cat.og + 2) * ldo + qi] = (float)(gf[2].value() - bz * A);
                }
                if (cat.nd) {
                    const double Ax = ds[0].value() * inv, Ay = ds[1].value() * inv, Az = ds[2].value() * inv;
                    out[cat.od * ldo + qi] = (float)(df.value() - (Ax * bx + Ay * by + Az * bz));
                }
                if (cat.nc) {
                    const double Ax = cs[0].value() * inv, Ay = cs[1].value() * inv, Az = cs[2].value() * inv;
                    const double ex = Ay * bz - Az * by, ey = Az * bx - Ax * bz, ez = Ax * by - Ay * bx;
                    out[cat.oc * ldo + qi] = (float)(-(cf[0].value() - ex));
                    out[(cat.oc + 1) * ldo + qi] = (float)(-(cf[1].value() - ey));
                    out[(cat.oc + 2) * ldo + qi] = (float)(-(cf[2].value() - ez));
                }
            }
        }
    }

}

}  // namespace pts
}  // namespace pmx
