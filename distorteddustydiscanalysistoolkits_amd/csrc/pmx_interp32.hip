// pmx_interp32.hip — FP32 compensated interpolation path for gfx950.
//
// One wavefront = one tile of 64 samples taken in Morton order of their positions.
//   * The wavefront walks the LBVH once for the whole tile (stackless left / escape links,
//     wave-uniform node ids): a node is entered if any lane's conservative FP32 test needs
//     it.  A node covering at most `bucket` leaves is not descended: its leaves, contiguous
//     in Morton order, are loaded by the lanes together (coalesced) into LDS with their
//     payload, and every lane then tests and accumulates them for its own sample.
//   * Neighbour decisions equal the reference's FP64 decisions: an FP32 test with a relative
//     guard band decides the clear cases; inside the band the reference's FP64 formula is
//     evaluated on the FP64 particle columns.  Node boxes are rounded outwards and hmax up,
//     so pruning is conservative: the accepted set is the FP64 predicate set.
//   * Accepted (sample, particle) pairs go through an LDS queue and are evaluated 64 at a
//     time, one pair per lane, into FP64 LDS sums.  Scalar catalogs weigh with FP32 kernel
//     shapes; derivative catalogs form their pair terms in FP64 (W, grad W, separation,
//     A_b - A_ref with A_ref the sample's nearest particle, or first accepted neighbour for
//     scatter, which leaves f - A~ b unchanged and removes the cancellation of large means).
//     Heavy tiles hand their remaining subtrees to an item queue walked by all wavefronts.
//     Line maps of one parallel set use the particle-centric tiles (pmx_lines_tiled.hip);
//     other lines test pairs in FP32 and evaluate accepted pairs in FP64.  All Shepard
//     divisions are FP64.
// Reference semantics: general_interpolation.jl:1-669 (points), line_integrated_scalar_
// interpolation.jl:123-166 (lines), LinearBVHQueries.jl:118-141 (nearest-h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "pmx_internal.h"
#include "pmx_tree32.h"
#include "pmx_points_tile.h"

namespace pmx {


namespace {

using namespace pts;

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

template <int K, int S, bool DERIV>
__global__ void __launch_bounds__(kWave) k_points32(T32 T, PointCat cat, const double* __restrict__ qx,
                                                    const double* __restrict__ qy, const double* __restrict__ qz,
                                                    const int32_t* __restrict__ perm,
                                                    int64_t npts,
                                                    float* __restrict__ out, int64_t ldo, PointSplit sp) {
    __shared__ BPosT<run_cap(DERIV)> P;
    const int64_t ntiles = (npts + kWave - 1) / kWave;
    if constexpr (DERIV) {
        extern __shared__ double dyn_tile[];        // deriv_tile_lds(cat.nder, cat.nacc) bytes
        BTile<true> Y(dyn_tile, cat.nder, cat.nacc);
        for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x)
            points_tile<K, S, true>(T, cat, qx, qy, qz, perm, npts, out, ldo, tile, P, Y, sp);
    } else {
        __shared__ BTile<false> Y;
        for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x)
            points_tile<K, S, false>(T, cat, qx, qy, qz, perm, npts, out, ldo, tile, P, Y, sp);
    }
}

// Items of split tiles, taken from the queue by every wavefront until it is empty.
template <int K, int S, bool DERIV>
__global__ void __launch_bounds__(kWave) k_points32_items(T32 T, PointCat cat, const double* __restrict__ qx,
                                                          const double* __restrict__ qy,
                                                          const double* __restrict__ qz,
                                                          const int32_t* __restrict__ perm, int64_t npts,
                                                          PointSplit sp) {
    __shared__ BPosT<run_cap(DERIV)> P;
    const unsigned int nitems = min(sp.counters[0], (unsigned int)sp.cap_items);
    if constexpr (DERIV) {
        extern __shared__ double dyn_tile[];
        BTile<true> Y(dyn_tile, cat.nder, cat.nacc);
        for (;;) {
            const unsigned int item = lane0_atomic_add(&sp.counters[2], 1u);
            if (item >= nitems) break;
            points_item<K, S, true>(T, cat, qx, qy, qz, perm, npts, sp, item, P, Y);
        }
    } else {
        __shared__ BTile<false> Y;
        for (;;) {
            const unsigned int item = lane0_atomic_add(&sp.counters[2], 1u);
            if (item >= nitems) break;
            points_item<K, S, false>(T, cat, qx, qy, qz, perm, npts, sp, item, P, Y);
        }
    }
}

template <bool DERIV>
__global__ void __launch_bounds__(kWave) k_points32_finish(PointCat cat, const int32_t* __restrict__ perm,
                                                           int64_t npts, float* __restrict__ out, int64_t ldo,
                                                           PointSplit sp) {
    const int nslots = (int)min(sp.counters[1], (unsigned int)sp.cap_slots);
    for (int split = blockIdx.x; split < nslots; split += gridDim.x)
        points_finish<DERIV>(cat, perm, npts, out, ldo, sp, split);
}

__global__ void k_zero_u32(unsigned int* __restrict__ p, int n) {
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0u;
}


// Neighbour lists / counts of the compensated path (inspection hook): same walk and
// decisions as k_points32, recording accepted leaves in increasing sorted index.
template <int S>
__global__ void __launch_bounds__(kWave) k_points32_nb(T32 T, double Kv, const double* __restrict__ qx,
                                                       const double* __restrict__ qy, const double* __restrict__ qz,
                                                       const double* __restrict__ ha_in, int64_t npts,
                                                       int64_t* __restrict__ counts, const int64_t* __restrict__ offsets,
                                                       int32_t* __restrict__ idx) {
    __shared__ BPos P;
    const int lane = threadIdx.x;
    const float Kvf = (float)Kv;
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
                const int64_t best = nearest32(T, q, P);
                ha = best < 0 ? __builtin_nan("") : h64(T, best);
            }
            const double r = Kv * ha;
            radius2_64 = r * r;
        }
        const float radius2_32 = (float)radius2_64;
        int64_t cnt_acc = 0;
        int32_t* dst = (valid && offsets) ? idx + offsets[qi] : nullptr;
        auto need = [&](const Node32& nd) { return node_need_point<S>(q, nd, Kvf, radius2_32); };
        auto bucket = [&](int lo, int cnt) {
            stage_pos(T, P, lo, cnt);
            __syncthreads();
            for (int j = 0; j < cnt; ++j) {
                float dx, dy, dz;
                const float d2 = d2_pos32(q, P, j, dx, dy, dz);
                if (accept_point<S>(T, q, d2, P.h[j], lo + j, Kvf, Kv, radius2_32, radius2_64)) {
                    if (dst) dst[cnt_acc] = lo + j;
                    ++cnt_acc;
                }
            }
            __syncthreads();
        };
        walk(T, ballot(valid) != 0, need, bucket);
        if (valid) counts[qi] = cnt_acc;
    }
}

__global__ void __launch_bounds__(kWave) k_nearest32(T32 T, const double* __restrict__ qx, const double* __restrict__ qy,
                                                     const double* __restrict__ qz, int64_t npts,
                                                     int64_t* __restrict__ idx) {
    __shared__ BPos P;
    const int64_t ntiles = (npts + kWave - 1) / kWave;
    for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int64_t qi = tile * kWave + threadIdx.x;
        const bool valid = qi < npts;
        const PointLane q = point_lane(qx, qy, qz, valid, valid ? qi : 0);
        const int64_t best = nearest32(T, q, P);
        if (valid) idx[qi] = best;
    }
}

// ---------------------------------------------------------------------------------------
// Lines (scatter; line_integrated_scalar_interpolation.jl:123-166)
// ---------------------------------------------------------------------------------------
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

// Exact reference line-leaf d2 in FP64 (_squared_distance_line_coords).
__device__ inline double d2_line_exact64(const double o[3], const double dv[3], const T32& T, int64_t li) {
    double px, py, pz;
    pos64(T, li, px, py, pz);
    double D2 = 0.0, Dm = 0.0, d;
    d = px - o[0]; D2 += d * d; Dm += d * dv[0];
    d = py - o[1]; D2 += d * d; Dm += d * dv[1];
    d = pz - o[2]; D2 += d * d; Dm += d * dv[2];
    return D2 - Dm * Dm;
}

// FP32 perpendicular distance^2 of a point from a line (cross-product form, no cancellation)
// and |Delta|^2 (for the rounding bound of the reference's D2 - Dm^2 form).
__device__ inline float d2_line32(const LineLane& q, float x, float y, float z, float xl, float yl, float zl,
                                  float& D2) {
    const float ax = (x - q.oh[0]) + (xl - q.ol[0]), ay = (y - q.oh[1]) + (yl - q.ol[1]),
                az = (z - q.oh[2]) + (zl - q.ol[2]);
    D2 = ax * ax + ay * ay + az * az;
    const float cx = ay * q.dv[2] - az * q.dv[1], cy = az * q.dv[0] - ax * q.dv[2], cz = ax * q.dv[1] - ay * q.dv[0];
    return cx * cx + cy * cy + cz * cz;
}

// Conservative lower bound of the squared distance between a line and an outward-rounded
// box: axis-aligned directions use the exact rectangle distance in the normal plane, other
// directions the reference's bounding-sphere bound with guard factors.
__device__ inline float d2_line_box32(const LineLane& q, const Node32& b) {
    const float mn[3] = {b.mnx, b.mny, b.mnz}, mx[3] = {b.mxx, b.mxy, b.mxz};
    if (q.axis >= 0) {
        float s = 0.0f;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            if (k == q.axis) continue;
            const float lo_gap = (mn[k] - q.oh[k]) - q.ol[k];
            const float hi_gap = (q.oh[k] - mx[k]) + q.ol[k];
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
    const float dc2 = d2_line32(q, c[0], c[1], c[2], 0.0f, 0.0f, 0.0f, D2);
    const float dc = __builtin_amdgcn_sqrtf(dc2) * (1.0f - kEps);
    const float R = __builtin_amdgcn_sqrtf(R2) * (1.0f + kEps) + 1e-30f;
    if (dc <= R) return 0.0f;
    const float gap = (dc - R) * (1.0f - kEps);
    return gap * gap;
}

__device__ inline bool accept_line(const T32& T, const LineLane& q, float d2, float D2, float hb32, int64_t li,
                                   float Kvf, double Kv) {
    const float rb = Kvf * hb32, r2 = rb * rb;
    // band: FP32 relative error of d2 plus the FP64 rounding of the reference's D2 - Dm^2
    const float tol = r2 * kEps + 1e-14f * D2;
    bool acc = q.valid && d2 < r2 - tol;
    const bool band = q.valid && !acc && !(d2 > r2 + tol);
    if (ballot(band)) {
        if (band) {
            const double e = d2_line_exact64(q.o64, q.d64, T, li);
            const double rr = Kv * h64(T, li);
            acc = e <= rr * rr;
        }
    }
    return acc;
}

struct LineCat32 {
    int n;
    int fcol[8];    // payload column of each output (1.. = quant)
    int shep[8];
};

struct BLine {
    double w64[kMaxRunLines];
    double invh64[kMaxRunLines];
    float f[8][kMaxRunLines];
};

// H2: weight Inorm / h^2 instead of Inorm / h (compile-time, as in the tiled line kernel).
template <int K, bool H2>
__global__ void __launch_bounds__(kWave) k_lines32(T32 T, LineCat32 cat, const double* __restrict__ o,
                                                   const double* __restrict__ dvec, const int32_t* __restrict__ perm,
                                                   int64_t nl, const double* __restrict__ lut, double dq,
                                                   float* __restrict__ out, int64_t ldo) {
    __shared__ BPosT<kMaxRunLines> P;
    __shared__ BLine Y;
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
        auto need = [&](const Node32& nd) {
            const float rb = Kvf * nd.hmax;
            return q.valid && !(d2_line_box32(q, nd) > rb * rb * (1.0f + kEps));
        };
        auto bucket = [&](int lo, int cnt) {
            stage_pos(T, P, lo, cnt);
            if (lane < cnt) {
                const int64_t li = lo + lane;
                Y.w64[lane] = T.w64[li];
                Y.invh64[lane] = 1.0 / h64(T, li);
                for (int f = 0; f < cat.n; ++f) Y.f[f][lane] = T.pay[(int64_t)cat.fcol[f] * n + li];
            }
            __syncthreads();
            for (int j = 0; j < cnt; ++j) {
                float D2;
                const float d2 = d2_line32(q, P.x[j], P.y[j], P.z[j], P.xl[j], P.yl[j], P.zl[j], D2);
                if (!accept_line(T, q, d2, D2, P.h[j], lo + j, Kvf, Kv)) continue;
                // line_integrated_kernel_function: invh * Inorm(r * invh), evaluated in FP64 with
                // the perpendicular distance from the hi + lo particle position
                const double ax = ((double)P.x[j] - q.o64[0]) + (double)P.xl[j];
                const double ay = ((double)P.y[j] - q.o64[1]) + (double)P.yl[j];
                const double az = ((double)P.z[j] - q.o64[2]) + (double)P.zl[j];
                const double cx = ay * q.d64[2] - az * q.d64[1], cy = az * q.d64[0] - ax * q.d64[2],
                             cz = ax * q.d64[1] - ay * q.d64[0];
                const double invh = Y.invh64[j];
                const double qq = sqrt(cx * cx + cy * cy + cz * cz) * invh;
                const double I = qq >= Kv ? 0.0 : lin_lut_f64(qq, dq, lut);
                const double w = Y.w64[j] * (H2 ? (invh * I) * invh : invh * I);
                S1 += w;
#pragma unroll
                for (int f = 0; f < 8; ++f)
                    if (f < cat.n) acc[f] += (double)Y.f[f][j] * w;
            }
            __syncthreads();
        };
        walk(T, ballot(valid) != 0, need, bucket);
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
    __shared__ BPosT<kMaxRunLines> P;
    const float Kvf = (float)Kv;
    const int64_t ntiles = (nl + kWave - 1) / kWave;
    for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int64_t qi = tile * kWave + threadIdx.x;
        const bool valid = qi < nl;
        const LineLane q = line_lane(o, dvec, nl, valid, valid ? qi : 0);
        int64_t cnt_acc = 0;
        int32_t* dst = (valid && offsets) ? idx + offsets[qi] : nullptr;
        auto need = [&](const Node32& nd) {
            const float rb = Kvf * nd.hmax;
            return q.valid && !(d2_line_box32(q, nd) > rb * rb * (1.0f + kEps));
        };
        auto bucket = [&](int lo, int cnt) {
            stage_pos(T, P, lo, cnt);
            __syncthreads();
            for (int j = 0; j < cnt; ++j) {
                float D2;
                const float d2 = d2_line32(q, P.x[j], P.y[j], P.z[j], P.xl[j], P.yl[j], P.zl[j], D2);
                if (accept_line(T, q, d2, D2, P.h[j], lo + j, Kvf, Kv)) {
                    if (dst) dst[cnt_acc] = lo + j;
                    ++cnt_acc;
                }
            }
            __syncthreads();
        };
        walk(T, ballot(valid) != 0, need, bucket);
        if (valid) counts[qi] = cnt_acc;
    }
}

unsigned tile_grid(int64_t nq) {
    const int64_t ntiles = (nq + kWave - 1) / kWave;
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>(ntiles, 1 << 20));
}

// The three launches of one sample pass: tiles (heavy ones split), the queued items of the
// split tiles, the outputs of the split tiles.
template <int K, int S, bool DERIV>
void launch_points32_s(pmx_ctx* c, const PointCat& cat, const double* qx, const double* qy, const double* qz,
                       const int32_t* perm, int64_t npts, float* out, int64_t ldo,
                       const PointSplit& sp) {
    const dim3 g(tile_grid(npts)), b(kWave);
    T32 T = tree32(c, c->tune.bucket_points, run_cap(DERIV));
    const bool split = S != kScatter && sp.budget > 0;
    const size_t lds = DERIV ? deriv_tile_lds(cat.nder, cat.nacc) : 0;
    if (split) hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, c->stream, sp.counters, 4);
    hipLaunchKernelGGL((k_points32<K, S, DERIV>), g, b, lds, c->stream, T, cat, qx, qy, qz, perm, npts, out, ldo,
                       sp);
    if (!split) return;
    hipLaunchKernelGGL((k_points32_items<K, S, DERIV>), dim3(c->ncu * 32), b, lds, c->stream, T, cat, qx, qy, qz,
                       perm, npts, sp);
    hipLaunchKernelGGL(k_points32_finish<DERIV>, dim3(std::min(sp.cap_slots, 4096)), b, 0, c->stream, cat, perm, npts,
                       out, ldo, sp);
}

template <int K, bool DERIV>
void launch_points32(pmx_ctx* c, int S, const PointCat& cat, const double* qx, const double* qy, const double* qz,
                     const int32_t* perm, int64_t npts, float* out, int64_t ldo,
                     const PointSplit& sp) {
    if (S == kGather) launch_points32_s<K, kGather, DERIV>(c, cat, qx, qy, qz, perm, npts, out, ldo, sp);
    else if (S == kScatter) launch_points32_s<K, kScatter, DERIV>(c, cat, qx, qy, qz, perm, npts, out, ldo, sp);
    else launch_points32_s<K, kSymmetric, DERIV>(c, cat, qx, qy, qz, perm, npts, out, ldo, sp);
}

template <int K>
void launch_points32_k(pmx_ctx* c, int S, const PointCat& cat, const double* qx, const double* qy, const double* qz,
                       const int32_t* perm, int64_t npts, float* out, int64_t ldo,
                       const PointSplit& sp) {
    if (cat.ng + cat.nd + cat.nc > 0) launch_points32<K, true>(c, S, cat, qx, qy, qz, perm, npts, out, ldo, sp);
    else launch_points32<K, false>(c, S, cat, qx, qy, qz, perm, npts, out, ldo, sp);
}

// Split-tile storage of the context: counters, item queue and slots (FP64 sums of up to
// kSplitSlots tiles); budget 0 when splitting is off.
int point_split(pmx_ctx* c, int64_t npts, int strategy, PointSplit* sp) {
    std::memset(sp, 0, sizeof(*sp));
    if (strategy == kScatter || c->split_budget <= 0) return PMX_OK;
    const int64_t ntiles = (npts + kWave - 1) / kWave;
    sp->budget = c->split_budget;
    sp->item_max = std::max(c->split_item, 2 * c->tune.bucket_points);
    sp->cap_items = kSplitItems;
    sp->cap_slots = (int)std::max<int64_t>(1, std::min<int64_t>(ntiles, kSplitSlots));
    const size_t b_cnt = 256, b_items = 8 * (size_t)sp->cap_items, b_tile = 4 * (size_t)sp->cap_slots,
                 b_near = 4 * (size_t)sp->cap_slots * kWave, b_acc = 8 * (size_t)sp->cap_slots * kAccMax * kWave;
    auto up = [](size_t v) { return (v + 255) & ~(size_t)255; };
    PMX_TRY(ensure(c, c->split, b_cnt + up(b_items) + up(b_tile) + up(b_near) + b_acc));
    char* base = c->split.as<char>();
    sp->counters = reinterpret_cast<unsigned int*>(base);
    sp->items = reinterpret_cast<int2*>(base + b_cnt);
    sp->slot_tile = reinterpret_cast<int*>(base + b_cnt + up(b_items));
    sp->slot_near = reinterpret_cast<int*>(base + b_cnt + up(b_items) + up(b_tile));
    sp->slot_acc = reinterpret_cast<double*>(base + b_cnt + up(b_items) + up(b_tile) + up(b_near));
    return PMX_OK;
}

template <int K>
void launch_lines32(pmx_ctx* c, const LineCat32& cat, const double* o, const double* d, const int32_t* perm,
                    int64_t nl, float* out, int64_t ldo) {
    auto kern = c->line_inv_h2 ? k_lines32<K, true> : k_lines32<K, false>;
    hipLaunchKernelGGL(kern, dim3(tile_grid(nl)), dim3(kWave), 0, c->stream, tree32(c, c->tune.bucket_lines, kMaxRunLines), cat, o, d,
                       perm, nl, c->lut64.as<double>() + (size_t)K * kLutN, c->lut_dq[K], out, ldo);
}

}  // namespace

T32 tree32(pmx_ctx* c, int bucket, int cap) {
    T32 T;
    const size_t payf = (size_t)c->tree.n * (2 + c->tree.ncol);
    T.nodes = c->nodes32.as<Node32>();
    T.leaf = c->leaf32.as<float4>();
    T.leaflo = c->leaf32.as<float4>() + c->tree.n;
    T.leaf_esc = c->leaf_esc.as<int32_t>();
    T.cols64 = c->tree.f32in ? nullptr : c->cols64.as<double>();
    T.f32in = c->tree.f32in ? 1 : 0;
    T.ncol = c->tree.ncol;
    T.pay = c->pay32.as<float>();
    T.w64 = reinterpret_cast<const double*>(c->pay32.as<float>() + payf + (payf & 1));
    T.ih64 = T.w64 + c->tree.n;
    T.n = c->tree.n;
    T.root = c->tree.root;
    T.bucket = std::max(kMinRun, std::min(bucket, cap));
    return T;
}

// Morton order of the samples (queries): permutation in c->scratch_a or c->scratch_b
int sort_queries(pmx_ctx* c, const double* x, const double* y, const double* z, int64_t npts, int32_t** perm,
                 const LineWork* ws) {
    uint64_t *k0, *k1;
    int32_t *v0, *v1;
    if (ws) {
        k0 = ws->keys0; v0 = ws->vals0; k1 = ws->keys1; v1 = ws->vals1;
    } else {
        PMX_TRY(ensure(c, c->scratch_a, 8 * (size_t)npts + 4 * (size_t)npts + 64));
        PMX_TRY(ensure(c, c->scratch_b, 8 * (size_t)npts + 4 * (size_t)npts + 64));
        k0 = c->scratch_a.as<uint64_t>();
        v0 = reinterpret_cast<int32_t*>(k0 + npts);
        k1 = c->scratch_b.as<uint64_t>();
        v1 = reinterpret_cast<int32_t*>(k1 + npts);
    }
    const dim3 g(grid_for(npts, 256, 8192)), b(256);
    if (c->tree.n >= 2) {
        hipLaunchKernelGGL(k_query_keys, g, b, 0, c->stream, x, y, z, 0, npts, c->nodes64.as<Node64>(), k0, v0);
        bool alt = false;
        PMX_TRY(radix_sort_pairs(c, k0, v0, npts, k1, v1, 0, 32, &alt, ws ? ws->hist : nullptr));
        *perm = alt ? v1 : v0;
    } else {
        hipLaunchKernelGGL(k_iota, g, b, 0, c->stream, k0, v0, npts);
        *perm = v0;
    }
    PMX_HIP(c, hipGetLastError());
    return PMX_OK;
}

// The FP32 tree (nodes32, leaf32, pay32) is written by the build itself (refit and column
// gather); this only checks that it belongs to the current tree.
int prepare_f32(pmx_ctx* c) {
    if (!c->tree.f32_ready) return set_error(c, PMX_ERR_STATE, "FP32 tree data missing: rebuild the LBVH");
    return PMX_OK;
}

// The launches of one catalog chunk: sub-chunks of <= 2 scalars, 1 gradient, 1 divergence and
// 1 curl, each with its derivative fields and the LDS rows of the sums it holds.
std::vector<PointCat> point_cats(const CatChunk& ch) {
    std::vector<PointCat> cats;
    const int nsub = std::max({(ch.ns + 1) / 2, ch.ng, ch.nd, ch.nc, 1});
    for (int k = 0; k < nsub; ++k) {
        PointCat cat{};
        auto dfield = [&](int col64) {
            for (int f = 0; f < cat.nder; ++f)
                if (cat.dcol[f] == col64) return f;
            cat.dcol[cat.nder] = col64;
            return cat.nder++;
        };
        cat.ns = std::max(0, std::min(2, ch.ns - 2 * k));
        cat.ng = k < ch.ng ? 1 : 0;
        cat.nd = k < ch.nd ? 1 : 0;
        cat.nc = k < ch.nc ? 1 : 0;
        for (int j = 0; j < cat.ns; ++j) { cat.scol[j] = 1 + ch.ss[2 * k + j]; cat.sh[j] = ch.sh[2 * k + j]; }
        if (cat.ng) cat.gf = dfield(ch.gs[k] < 0 ? 5 : 6 + ch.gs[k]);
        if (cat.nd) for (int j = 0; j < 3; ++j) cat.df[j] = dfield(6 + ch.ds[3 * k + j]);
        if (cat.nc) for (int j = 0; j < 3; ++j) cat.cf[j] = dfield(6 + ch.cs[3 * k + j]);
        cat.os = ch.os + 2 * k;
        cat.og = ch.og + 3 * k;
        cat.od = ch.od + k;
        cat.oc = ch.oc + 3 * k;
        // LDS rows of the sums a derivative launch holds
        for (int a = 0; a < kAccMax; ++a) cat.amap[a] = -1;
        int rows = 0;
        auto hold = [&](int a, int cnt) { for (int j = 0; j < cnt; ++j) cat.amap[a + j] = rows++; };
        hold(kA_S1, 1);
        hold(kA_SC, cat.ns);
        hold(kA_B, 3);
        if (cat.ng) { hold(kA_GF, 3); hold(kA_GS, 1); }
        if (cat.nd) { hold(kA_DF, 1); hold(kA_DS, 3); }
        if (cat.nc) {
            hold(kA_CF, 3);
            cat.cs_alias = cat.nd && cat.cf[0] == cat.df[0] && cat.cf[1] == cat.df[1] && cat.cf[2] == cat.df[2];
            if (cat.cs_alias)
                for (int j = 0; j < 3; ++j) cat.amap[kA_CS + j] = cat.amap[kA_DS + j];
            else
                hold(kA_CS, 3);
        }
        cat.nacc = rows;
        if (cat.ns + cat.ng + cat.nd + cat.nc == 0 && k > 0) continue;
        cats.push_back(cat);
    }
    return cats;
}

int interp_points_f32c(pmx_ctx* c, int strategy, const CatChunk& ch, const int32_t* perm,
                       const double* qx, const double* qy, const double* qz, int64_t npts, float* out, int64_t ldo) {
    for (const PointCat& cat : point_cats(ch)) {
        PointSplit sp;
        PMX_TRY(point_split(c, npts, strategy, &sp));
        switch (c->tree.kernel) {
        case kM4: launch_points32_k<kM4>(c, strategy, cat, qx, qy, qz, perm, npts, out, ldo, sp); break;
        case kM5: launch_points32_k<kM5>(c, strategy, cat, qx, qy, qz, perm, npts, out, ldo, sp); break;
        case kM6: launch_points32_k<kM6>(c, strategy, cat, qx, qy, qz, perm, npts, out, ldo, sp); break;
        case kC2: launch_points32_k<kC2>(c, strategy, cat, qx, qy, qz, perm, npts, out, ldo, sp); break;
        case kC4: launch_points32_k<kC4>(c, strategy, cat, qx, qy, qz, perm, npts, out, ldo, sp); break;
        default: launch_points32_k<kC6>(c, strategy, cat, qx, qy, qz, perm, npts, out, ldo, sp); break;
        }
    }
    PMX_HIP(c, hipGetLastError());
    return PMX_OK;
}

int interp_lines_f32c(pmx_ctx* c, const int* slots, const int* shep, int ncols, const int32_t* perm,
                      const double* o, const double* d, int64_t nl, float* out, int64_t ldo) {
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
    hipLaunchKernelGGL(k_nearest32, dim3(tile_grid(npts)), dim3(kWave), 0, c->stream, tree32(c, c->tune.bucket_points, kRunScalar), qx, qy,
                       qz, npts, idx);
    PMX_HIP(c, hipGetLastError());
    return PMX_OK;
}

int neighbors_points_f32c(pmx_ctx* c, int strategy, const double* qx, const double* qy, const double* qz,
                          const double* ha, int64_t npts, int64_t* counts, const int64_t* offsets, int32_t* idx) {
    PMX_TRY(prepare_f32(c));
    const dim3 g(tile_grid(npts)), b(kWave);
    const double Kv = support(c->tree.kernel);
    T32 T = tree32(c, c->tune.bucket_points, kRunScalar);
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
    hipLaunchKernelGGL(k_lines32_nb, dim3(tile_grid(nl)), dim3(kWave), 0, c->stream, tree32(c, c->tune.bucket_lines, kMaxRunLines),
                       support(c->tree.kernel), o, d, nl, counts, offsets, idx);
    PMX_HIP(c, hipGetLastError());
    return PMX_OK;
}

}  // namespace pmx
