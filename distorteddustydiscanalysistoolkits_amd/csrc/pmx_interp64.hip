// pmx_interp64.hip — plain FP64 interpolation path: one thread per sample, 256-thread
// workgroups, grid-stride loop, stackless left/escape traversal, FP64 everywhere.  This is
// the reference's GPU design (Partia.jl/ext/CUDAExt/KernelInterpolation/grid_interpolation/
// PointSamples_interpolation.jl:22-99, LineSamples_interpolation.jl:27-90) restated for
// gfx950; it evaluates every quantity in the reference's FP64 operation order and is the
// baseline the compensated FP32 path is measured against.
#include <hip/hip_runtime.h>

#include "pmx_internal.h"

namespace pmx {
namespace {

constexpr int kBlock = 256;

struct TreeView {
    const Node64* __restrict__ nodes;
    const Leaf64* __restrict__ leaf;
    const int32_t* __restrict__ leaf_esc;
    int64_t n;
    int32_t root;
};

struct PartView {
    const double* __restrict__ m;
    const double* __restrict__ rho;
    const double* __restrict__ quant;   // column j at quant + j * n
    int64_t n;
};

// Traversal statistics for the traffic model (instrumented kernels only).
struct Stats {
    int64_t internal, leaves;
};

// Stackless DFS over left / escape links (LinearBVHTraversalMacros.jl:26-371); leaves are
// reached in increasing sorted index, the reference's accumulation order.
template <int S, typename Visit>
__device__ inline void traverse_point(const TreeView& T, const double p[3], double Kv, double radius2, Visit& visit,
                                      Stats* st = nullptr) {
    const int64_t n = T.n;
    if (T.root == 0) {
        for (int64_t li = 0; li < n; ++li) {
            const Leaf64 L = T.leaf[li];
            double r2 = radius2_rule<S>(Kv, radius2, L.h);
            double d2 = d2_point_leaf(p, L);
            if (d2 <= r2) visit(li, d2, L.h);
        }
        return;
    }
    int64_t node = T.root;
    int64_t steps = 0;
    const int64_t max_steps = 2 * n;   // a DFS visits each of the 2n-1 nodes at most once
    while (node > 0 && node < 2 * n && ++steps <= max_steps) {
        if (node >= n) {
            const int64_t li = node - n;
            const Leaf64 L = T.leaf[li];
            double r2 = radius2_rule<S>(Kv, radius2, L.h);
            double d2 = d2_point_leaf(p, L);
            if (st) ++st->leaves;
            if (d2 <= r2) visit(li, d2, L.h);
            node = T.leaf_esc[li];
            continue;
        }
        if (st) ++st->internal;
        const Node64 nd = T.nodes[node];
        double r2 = radius2_rule<S>(Kv, radius2, nd.hmax);
        double d2n = d2_point_aabb(p, nd);
        node = (d2n > r2) ? nd.escape : nd.left;
    }
}

template <int S, typename Visit>
__device__ inline void traverse_line(const TreeView& T, const double o[3], const double dv[3], double Kv,
                                     double radius2, Visit& visit, Stats* st = nullptr) {
    const int64_t n = T.n;
    if (T.root == 0) {
        for (int64_t li = 0; li < n; ++li) {
            const Leaf64 L = T.leaf[li];
            double r2 = radius2_rule<S>(Kv, radius2, L.h);
            double d2 = d2_line_leaf(o, dv, L);
            if (d2 <= r2) visit(li, d2, L.h);
        }
        return;
    }
    int64_t node = T.root;
    int64_t steps = 0;
    const int64_t max_steps = 2 * n;
    while (node > 0 && node < 2 * n && ++steps <= max_steps) {
        if (node >= n) {
            const int64_t li = node - n;
            const Leaf64 L = T.leaf[li];
            double r2 = radius2_rule<S>(Kv, radius2, L.h);
            double d2 = d2_line_leaf(o, dv, L);
            if (st) ++st->leaves;
            if (d2 <= r2) visit(li, d2, L.h);
            node = T.leaf_esc[li];
            continue;
        }
        if (st) ++st->internal;
        const Node64 nd = T.nodes[node];
        double r2 = radius2_rule<S>(Kv, radius2, nd.hmax);
        double d2n = d2_line_aabb_lb(o, dv, nd);
        node = (d2n > r2) ? nd.escape : nd.left;
    }
}

// LBVH_find_nearest (LinearBVHQueries.jl:79-95 / 118-141): the pruning radius is the
// running best d2; strict `<` keeps the smallest sorted index among equal distances.
__device__ inline int64_t nearest_leaf(const TreeView& T, const double p[3], Stats* st = nullptr) {
    const int64_t n = T.n;
    int64_t best = -1;
    double bd = __builtin_inf();
    if (T.root == 0) {
        for (int64_t li = 0; li < n; ++li) {
            double d2 = d2_point_leaf(p, T.leaf[li]);
            if (d2 < bd) { bd = d2; best = li; }
        }
        return best;
    }
    int64_t node = T.root;
    int64_t steps = 0;
    const int64_t max_steps = 2 * n;
    while (node > 0 && node < 2 * n && ++steps <= max_steps) {
        if (node >= n) {
            const int64_t li = node - n;
            double d2 = d2_point_leaf(p, T.leaf[li]);
            if (st) ++st->leaves;
            if (d2 < bd) { bd = d2; best = li; }
            node = T.leaf_esc[li];
            continue;
        }
        if (st) ++st->internal;
        const Node64 nd = T.nodes[node];
        double d2n = d2_point_aabb(p, nd);
        node = (d2n > bd) ? nd.escape : nd.left;
    }
    return best;
}

// Per-sample accumulators of _general_quantity_interpolate_kernel
// (general_interpolation.jl:468-488), sized for one catalog chunk.
struct Acc {
    double S1;
    double sc[4];
    double gf[2][3], gb[2][3], gs[2];
    double df[2], db[2][3], ds[2][3];
    double cf[2][3], cb[2][3], cs[2][3];
};

template <int K, int S>
struct PointVisit {
    const TreeView& T;
    const PartView& P;
    const CatChunk& cat;
    const double* p;
    double ha;
    bool need_grad;
    Acc& a;
    // One accepted neighbour: general_interpolation.jl:495-569 with the accumulation helpers
    // of accumulations/{scalar,gradient,divergence,curl}_accumulation.jl.
    __device__ inline void operator()(int64_t li, double d2, double hb) {
        const Leaf64 L = T.leaf[li];
        const double dx = p[0] - L.x, dy = p[1] - L.y, dz = p[2] - L.z;
        const double dr = sqrt(d2);
        const double mb = P.m[li], rhob = P.rho[li];
        double W;
        if (S == kSymmetric) W = 0.5 * (W_f64<K>(dr, ha) + W_f64<K>(dr, hb));
        else W = W_f64<K>(dr, S == kGather ? ha : hb);
        const double w = mb * W / rhob;
        a.S1 += w;
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if (j < cat.ns) a.sc[j] += P.quant[(int64_t)cat.ss[j] * P.n + li] * w;
        if (!need_grad) return;
        double g[3];
        if (S == kSymmetric) {
            double ga[3], gb[3];
            gradW_f64<K>(dx, dy, dz, ha, ga);
            gradW_f64<K>(dx, dy, dz, hb, gb);
            g[0] = 0.5 * (ga[0] + gb[0]);
            g[1] = 0.5 * (ga[1] + gb[1]);
            g[2] = 0.5 * (ga[2] + gb[2]);
        } else {
            gradW_f64<K>(dx, dy, dz, S == kGather ? ha : hb, g);
        }
        const double invrho = 1.0 / rhob;
        const double bx = mb * invrho * g[0], by = mb * invrho * g[1], bz = mb * invrho * g[2];
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            if (j >= cat.ng) break;
            const int s = cat.gs[j];
            if (s < 0) {
                const double mx = mb * g[0], my = mb * g[1], mz = mb * g[2];
                a.gf[j][0] += mx; a.gf[j][1] += my; a.gf[j][2] += mz;
                a.gb[j][0] += mx * invrho; a.gb[j][1] += my * invrho; a.gb[j][2] += mz * invrho;
                a.gs[j] += mb * W;
            } else {
                const double Ab = P.quant[(int64_t)s * P.n + li];
                a.gf[j][0] += bx * Ab; a.gf[j][1] += by * Ab; a.gf[j][2] += bz * Ab;
                a.gb[j][0] += bx; a.gb[j][1] += by; a.gb[j][2] += bz;
                a.gs[j] += Ab * w;
            }
        }
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            if (j >= cat.nd) break;
            const double Ax = P.quant[(int64_t)cat.ds[3 * j] * P.n + li];
            const double Ay = P.quant[(int64_t)cat.ds[3 * j + 1] * P.n + li];
            const double Az = P.quant[(int64_t)cat.ds[3 * j + 2] * P.n + li];
            a.df[j] += bx * Ax + by * Ay + bz * Az;
            a.db[j][0] += bx; a.db[j][1] += by; a.db[j][2] += bz;
            a.ds[j][0] += Ax * w; a.ds[j][1] += Ay * w; a.ds[j][2] += Az * w;
        }
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            if (j >= cat.nc) break;
            const double Ax = P.quant[(int64_t)cat.cs[3 * j] * P.n + li];
            const double Ay = P.quant[(int64_t)cat.cs[3 * j + 1] * P.n + li];
            const double Az = P.quant[(int64_t)cat.cs[3 * j + 2] * P.n + li];
            a.cf[j][0] += Ay * bz - Az * by;
            a.cf[j][1] += Az * bx - Ax * bz;
            a.cf[j][2] += Ax * by - Ay * bx;
            a.cb[j][0] += bx; a.cb[j][1] += by; a.cb[j][2] += bz;
            a.cs[j][0] += Ax * w; a.cs[j][1] += Ay * w; a.cs[j][2] += Az * w;
        }
    }
};

// Finalisation (general_interpolation.jl:574-668) and catalog-ordered store
// (PointSamples_interpolation.jl:140-182).
__device__ inline void finalize(const Acc& a, const CatChunk& cat, double* out, int64_t ldo, int64_t i) {
    if (a.S1 == 0.0) {
        const double nan = __builtin_nan("");
        for (int j = 0; j < cat.ns; ++j) out[(cat.os + j) * ldo + i] = nan;
        for (int j = 0; j < 3 * cat.ng; ++j) out[(cat.og + j) * ldo + i] = nan;
        for (int j = 0; j < cat.nd; ++j) out[(cat.od + j) * ldo + i] = nan;
        for (int j = 0; j < 3 * cat.nc; ++j) out[(cat.oc + j) * ldo + i] = nan;
        return;
    }
    const double invS1 = 1.0 / a.S1;
#pragma unroll
    for (int j = 0; j < 4; ++j)
        if (j < cat.ns) out[(cat.os + j) * ldo + i] = cat.sh[j] ? a.sc[j] * invS1 : a.sc[j];
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        if (j >= cat.ng) break;
        const double A = a.gs[j] * invS1;
        for (int d = 0; d < 3; ++d) out[(cat.og + 3 * j + d) * ldo + i] = a.gf[j][d] - a.gb[j][d] * A;
    }
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        if (j >= cat.nd) break;
        const double Ax = a.ds[j][0] * invS1, Ay = a.ds[j][1] * invS1, Az = a.ds[j][2] * invS1;
        const double b = Ax * a.db[j][0] + Ay * a.db[j][1] + Az * a.db[j][2];
        out[(cat.od + j) * ldo + i] = a.df[j] - b;
    }
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        if (j >= cat.nc) break;
        const double Ax = a.cs[j][0] * invS1, Ay = a.cs[j][1] * invS1, Az = a.cs[j][2] * invS1;
        const double mx = a.cb[j][0], my = a.cb[j][1], mz = a.cb[j][2];
        const double bx = Ay * mz - Az * my, by = Az * mx - Ax * mz, bz = Ax * my - Ay * mx;
        out[(cat.oc + 3 * j) * ldo + i] = -(a.cf[j][0] - bx);
        out[(cat.oc + 3 * j + 1) * ldo + i] = -(a.cf[j][1] - by);
        out[(cat.oc + 3 * j + 2) * ldo + i] = -(a.cf[j][2] - bz);
    }
}

template <int K, int S>
__global__ void __launch_bounds__(kBlock) k_points(TreeView T, PartView P, CatChunk cat, const double* __restrict__ qx,
                                                   const double* __restrict__ qy, const double* __restrict__ qz,
                                                   int64_t npts, double* __restrict__ out, int64_t ldo,
                                                   const double* __restrict__ ha_in, double* __restrict__ ha_out) {
    const double Kv = support(K);
    const bool need_grad = cat.ng + cat.nd + cat.nc > 0;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < npts; i += (int64_t)gridDim.x * kBlock) {
        const double p[3] = {qx[i], qy[i], qz[i]};
        double ha = 0.0, radius2 = 0.0;
        if (S != kScatter) {
            if (ha_in) {
                ha = ha_in[i];
            } else {
                const int64_t b = nearest_leaf(T, p);
                ha = b < 0 ? __builtin_nan("") : T.leaf[b].h;
            }
            if (ha_out) ha_out[i] = ha;
            const double radius = Kv * ha;
            radius2 = radius * radius;
        }
        Acc a = {};
        PointVisit<K, S> v{T, P, cat, p, ha, need_grad, a};
        traverse_point<S>(T, p, Kv, radius2, v);
        finalize(a, cat, out, ldo, i);
    }
}

// Line-integrated scatter kernel (line_integrated_scalar_interpolation.jl:123-166):
// d2 below zero from rounding is clamped before sqrt (the reference's CPU sqrt would throw).
struct LineCat {
    int n;
    int slot[8];
    int shep[8];
    int h2;         // weight Inorm / h^2 instead of Inorm / h
};

template <int K>
struct LineVisit {
    const PartView& P;
    const LineCat& cat;
    const double* lut;
    double dq;
    double S1;
    double acc[8];
    __device__ inline void operator()(int64_t li, double d2, double hb) {
        const double dr = sqrt(d2 > 0.0 ? d2 : 0.0);
        double W = lineW_f64(K, dr, hb, dq, lut);
        if (cat.h2) W = W * (1.0 / hb);
        const double w = P.m[li] * W / P.rho[li];
        S1 += w;
#pragma unroll
        for (int j = 0; j < 8; ++j)
            if (j < cat.n) acc[j] += P.quant[(int64_t)cat.slot[j] * P.n + li] * w;
    }
};

template <int K>
__global__ void __launch_bounds__(kBlock) k_lines(TreeView T, PartView P, LineCat cat, const double* __restrict__ o,
                                                  const double* __restrict__ dvec, int64_t nl, const double* __restrict__ lut,
                                                  double dq, double* __restrict__ out, int64_t ldo) {
    const double Kv = support(K);
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < nl; i += (int64_t)gridDim.x * kBlock) {
        const double oo[3] = {o[i], o[nl + i], o[2 * nl + i]};
        const double dd[3] = {dvec[i], dvec[nl + i], dvec[2 * nl + i]};
        LineVisit<K> v{P, cat, lut, dq, 0.0, {0, 0, 0, 0, 0, 0, 0, 0}};
        traverse_line<kScatter>(T, oo, dd, Kv, 0.0, v);
        if (v.S1 == 0.0) {
            for (int j = 0; j < cat.n; ++j) out[j * ldo + i] = __builtin_nan("");
            continue;
        }
        const double invS1 = 1.0 / v.S1;
        for (int j = 0; j < cat.n; ++j) out[j * ldo + i] = cat.shep[j] ? v.acc[j] * invS1 : v.acc[j];
    }
}

// ---- neighbour lists and nearest (inspection hooks) ----
struct CountVisit {
    int64_t cnt;
    int32_t* dst;
    __device__ inline void operator()(int64_t li, double, double) {
        if (dst) dst[cnt] = (int32_t)li;
        ++cnt;
    }
};

template <int S>
__global__ void __launch_bounds__(kBlock) k_point_nb(TreeView T, double Kv, const double* __restrict__ qx,
                                                     const double* __restrict__ qy, const double* __restrict__ qz,
                                                     const double* __restrict__ ha, int64_t npts, int64_t* __restrict__ counts,
                                                     const int64_t* __restrict__ offsets, int32_t* __restrict__ idx) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < npts; i += (int64_t)gridDim.x * kBlock) {
        const double p[3] = {qx[i], qy[i], qz[i]};
        double radius2 = 0.0;
        if (S != kScatter) {
            double hai;
            if (ha) hai = ha[i];
            else {
                const int64_t b = nearest_leaf(T, p);
                hai = b < 0 ? __builtin_nan("") : T.leaf[b].h;
            }
            const double r = Kv * hai;
            radius2 = r * r;
        }
        CountVisit v{0, offsets ? idx + offsets[i] : nullptr};
        traverse_point<S>(T, p, Kv, radius2, v);
        counts[i] = v.cnt;
    }
}

template <int S>
__global__ void __launch_bounds__(kBlock) k_line_nb(TreeView T, double Kv, const double* __restrict__ o,
                                                    const double* __restrict__ dvec, const double* __restrict__ ha,
                                                    int64_t nl, int64_t* __restrict__ counts,
                                                    const int64_t* __restrict__ offsets, int32_t* __restrict__ idx) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < nl; i += (int64_t)gridDim.x * kBlock) {
        const double oo[3] = {o[i], o[nl + i], o[2 * nl + i]};
        const double dd[3] = {dvec[i], dvec[nl + i], dvec[2 * nl + i]};
        double radius2 = 0.0;
        if (S != kScatter && ha) { const double r = Kv * ha[i]; radius2 = r * r; }
        CountVisit v{0, offsets ? idx + offsets[i] : nullptr};
        traverse_line<S>(T, oo, dd, Kv, radius2, v);
        counts[i] = v.cnt;
    }
}

__global__ void __launch_bounds__(kBlock) k_nearest(TreeView T, const double* __restrict__ qx,
                                                    const double* __restrict__ qy, const double* __restrict__ qz,
                                                    int64_t npts, int64_t* __restrict__ idx) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < npts; i += (int64_t)gridDim.x * kBlock) {
        const double p[3] = {qx[i], qy[i], qz[i]};
        idx[i] = nearest_leaf(T, p);
    }
}

template <int K>
__global__ void __launch_bounds__(kBlock) k_kernel_eval(const double* __restrict__ q, int64_t n, double* __restrict__ f,
                                                        double* __restrict__ df) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
        f[i] = shape_f64<K>(q[i]);
        df[i] = dshape_f64<K>(q[i]);
    }
}

// Visit counts of the FP64 traversal: [0] nearest-h internal, [1] nearest-h leaves,
// [2] internal nodes tested, [3] leaves tested, [4] leaves accepted (sums over samples).
__device__ inline void block_add_stats(const int64_t v[5], unsigned long long* tot) {
    __shared__ int64_t s[5][kBlock];
    for (int k = 0; k < 5; ++k) s[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int k = 0; k < 5; ++k) s[k][threadIdx.x] += s[k][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < 5) atomicAdd(&tot[threadIdx.x], (unsigned long long)s[threadIdx.x][0]);
}

template <int S>
__global__ void __launch_bounds__(kBlock) k_point_stats(TreeView T, double Kv, const double* __restrict__ qx,
                                                        const double* __restrict__ qy, const double* __restrict__ qz,
                                                        int64_t npts, unsigned long long* __restrict__ tot) {
    int64_t v[5] = {0, 0, 0, 0, 0};
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < npts; i += (int64_t)gridDim.x * kBlock) {
        const double p[3] = {qx[i], qy[i], qz[i]};
        double radius2 = 0.0;
        if (S != kScatter) {
            Stats sn{0, 0};
            const int64_t b = nearest_leaf(T, p, &sn);
            v[0] += sn.internal;
            v[1] += sn.leaves;
            const double r = Kv * (b < 0 ? __builtin_nan("") : T.leaf[b].h);
            radius2 = r * r;
        }
        Stats st{0, 0};
        CountVisit cv{0, nullptr};
        traverse_point<S>(T, p, Kv, radius2, cv, &st);
        v[2] += st.internal;
        v[3] += st.leaves;
        v[4] += cv.cnt;
    }
    block_add_stats(v, tot);
}

__global__ void __launch_bounds__(kBlock) k_line_stats(TreeView T, double Kv, const double* __restrict__ o,
                                                       const double* __restrict__ dvec, int64_t nl,
                                                       unsigned long long* __restrict__ tot) {
    int64_t v[5] = {0, 0, 0, 0, 0};
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < nl; i += (int64_t)gridDim.x * kBlock) {
        const double oo[3] = {o[i], o[nl + i], o[2 * nl + i]};
        const double dd[3] = {dvec[i], dvec[nl + i], dvec[2 * nl + i]};
        Stats st{0, 0};
        CountVisit cv{0, nullptr};
        traverse_line<kScatter>(T, oo, dd, Kv, 0.0, cv, &st);
        v[2] += st.internal;
        v[3] += st.leaves;
        v[4] += cv.cnt;
    }
    block_add_stats(v, tot);
}

TreeView tree_view(pmx_ctx* c) {
    return TreeView{c->nodes64.as<Node64>(), c->leaf64.as<Leaf64>(), c->leaf_esc.as<int32_t>(), c->tree.n,
                    c->tree.root};
}
PartView part_view(pmx_ctx* c) {
    const int64_t n = c->tree.n;
    const double* cols = c->cols64.as<double>();
    return PartView{cols + 3 * n, cols + 5 * n, cols + 6 * n, n};
}

template <int K>
void launch_points_k(pmx_ctx* c, int S, const CatChunk& cat, const double* qx, const double* qy, const double* qz,
                     int64_t npts, double* out, int64_t ldo, const double* ha_in, double* ha_out) {
    const dim3 g(grid_for(npts, kBlock, 1 << 16)), b(kBlock);
    TreeView T = tree_view(c);
    PartView P = part_view(c);
    if (S == kGather)
        hipLaunchKernelGGL((k_points<K, kGather>), g, b, 0, c->stream, T, P, cat, qx, qy, qz, npts, out, ldo, ha_in, ha_out);
    else if (S == kScatter)
        hipLaunchKernelGGL((k_points<K, kScatter>), g, b, 0, c->stream, T, P, cat, qx, qy, qz, npts, out, ldo, ha_in, ha_out);
    else
        hipLaunchKernelGGL((k_points<K, kSymmetric>), g, b, 0, c->stream, T, P, cat, qx, qy, qz, npts, out, ldo, ha_in, ha_out);
}

template <int K>
void launch_lines_k(pmx_ctx* c, const LineCat& cat, const double* o, const double* d, int64_t nl, double* out,
                    int64_t ldo) {
    const dim3 g(grid_for(nl, kBlock, 1 << 16)), b(kBlock);
    hipLaunchKernelGGL(k_lines<K>, g, b, 0, c->stream, tree_view(c), part_view(c), cat, o, d, nl,
                       c->lut64.as<double>() + (size_t)K * kLutN, c->lut_dq[K], out, ldo);
}

}  // namespace

int interp_points_f64(pmx_ctx* c, int strategy, const CatChunk& cat, const double* qx, const double* qy,
                      const double* qz, int64_t npts, double* out, int64_t ldo, const double* ha_in, double* ha_out) {
    PMX_TRY(ensure_leaf64(c));
    switch (c->tree.kernel) {
    case kM4: launch_points_k<kM4>(c, strategy, cat, qx, qy, qz, npts, out, ldo, ha_in, ha_out); break;
    case kM5: launch_points_k<kM5>(c, strategy, cat, qx, qy, qz, npts, out, ldo, ha_in, ha_out); break;
    case kM6: launch_points_k<kM6>(c, strategy, cat, qx, qy, qz, npts, out, ldo, ha_in, ha_out); break;
    case kC2: launch_points_k<kC2>(c, strategy, cat, qx, qy, qz, npts, out, ldo, ha_in, ha_out); break;
    case kC4: launch_points_k<kC4>(c, strategy, cat, qx, qy, qz, npts, out, ldo, ha_in, ha_out); break;
    default: launch_points_k<kC6>(c, strategy, cat, qx, qy, qz, npts, out, ldo, ha_in, ha_out); break;
    }
    PMX_HIP(c, hipGetLastError());
    return PMX_OK;
}

int interp_lines_f64(pmx_ctx* c, int strategy, const int* slots, const int* shep, int ncols, const double* o,
                     const double* d, int64_t nl, double* out, int64_t ldo) {
    PMX_TRY(ensure_leaf64(c));
    (void)strategy;
    for (int base = 0; base < ncols || (ncols == 0 && base == 0); base += 8) {
        LineCat cat{};
        cat.n = ncols - base < 8 ? ncols - base : 8;
        for (int j = 0; j < cat.n; ++j) { cat.slot[j] = slots[base + j]; cat.shep[j] = shep[base + j]; }
        cat.h2 = c->line_inv_h2;
        double* ob = out + (int64_t)base * ldo;
        switch (c->tree.kernel) {
        case kM4: launch_lines_k<kM4>(c, cat, o, d, nl, ob, ldo); break;
        case kM5: launch_lines_k<kM5>(c, cat, o, d, nl, ob, ldo); break;
        case kM6: launch_lines_k<kM6>(c, cat, o, d, nl, ob, ldo); break;
        case kC2: launch_lines_k<kC2>(c, cat, o, d, nl, ob, ldo); break;
        case kC4: launch_lines_k<kC4>(c, cat, o, d, nl, ob, ldo); break;
        default: launch_lines_k<kC6>(c, cat, o, d, nl, ob, ldo); break;
        }
        if (ncols == 0) break;
    }
    PMX_HIP(c, hipGetLastError());
    return PMX_OK;
}

int nearest_f64(pmx_ctx* c, const double* qx, const double* qy, const double* qz, int64_t npts, int64_t* idx) {
    PMX_TRY(ensure_leaf64(c));
    hipLaunchKernelGGL(k_nearest, dim3(grid_for(npts, kBlock, 1 << 16)), dim3(kBlock), 0, c->stream, tree_view(c), qx,
                       qy, qz, npts, idx);
    PMX_HIP(c, hipGetLastError());
    return PMX_OK;
}

int neighbors_points_f64(pmx_ctx* c, int strategy, const double* qx, const double* qy, const double* qz,
                         const double* ha, int64_t npts, int64_t* counts, const int64_t* offsets, int32_t* idx) {
    PMX_TRY(ensure_leaf64(c));
    const dim3 g(grid_for(npts, kBlock, 1 << 16)), b(kBlock);
    const double Kv = support(c->tree.kernel);
    TreeView T = tree_view(c);
    if (strategy == kGather)
        hipLaunchKernelGGL(k_point_nb<kGather>, g, b, 0, c->stream, T, Kv, qx, qy, qz, ha, npts, counts, offsets, idx);
    else if (strategy == kScatter)
        hipLaunchKernelGGL(k_point_nb<kScatter>, g, b, 0, c->stream, T, Kv, qx, qy, qz, ha, npts, counts, offsets, idx);
    else
        hipLaunchKernelGGL(k_point_nb<kSymmetric>, g, b, 0, c->stream, T, Kv, qx, qy, qz, ha, npts, counts, offsets, idx);
    PMX_HIP(c, hipGetLastError());
    return PMX_OK;
}

int neighbors_lines_f64(pmx_ctx* c, int strategy, const double* o, const double* d, const double* ha, int64_t nl,
                        int64_t* counts, const int64_t* offsets, int32_t* idx) {
    PMX_TRY(ensure_leaf64(c));
    const dim3 g(grid_for(nl, kBlock, 1 << 16)), b(kBlock);
    const double Kv = support(c->tree.kernel);
    TreeView T = tree_view(c);
    if (strategy == kGather)
        hipLaunchKernelGGL(k_line_nb<kGather>, g, b, 0, c->stream, T, Kv, o, d, ha, nl, counts, offsets, idx);
    else if (strategy == kScatter)
        hipLaunchKernelGGL(k_line_nb<kScatter>, g, b, 0, c->stream, T, Kv, o, d, ha, nl, counts, offsets, idx);
    else
        hipLaunchKernelGGL(k_line_nb<kSymmetric>, g, b, 0, c->stream, T, Kv, o, d, ha, nl, counts, offsets, idx);
    PMX_HIP(c, hipGetLastError());
    return PMX_OK;
}

int traversal_stats_f64(pmx_ctx* c, int kind, int strategy, const double* a, const double* b, const double* cc,
                        int64_t n, unsigned long long* tot) {
    PMX_TRY(ensure_leaf64(c));
    const dim3 g(grid_for(n, kBlock, 1 << 16)), bl(kBlock);
    const double Kv = support(c->tree.kernel);
    TreeView T = tree_view(c);
    if (kind == 1) {
        hipLaunchKernelGGL(k_line_stats, g, bl, 0, c->stream, T, Kv, a, b, n, tot);
    } else if (strategy == kGather) {
        hipLaunchKernelGGL(k_point_stats<kGather>, g, bl, 0, c->stream, T, Kv, a, b, cc, n, tot);
    } else if (strategy == kScatter) {
        hipLaunchKernelGGL(k_point_stats<kScatter>, g, bl, 0, c->stream, T, Kv, a, b, cc, n, tot);
    } else {
        hipLaunchKernelGGL(k_point_stats<kSymmetric>, g, bl, 0, c->stream, T, Kv, a, b, cc, n, tot);
    }
    PMX_HIP(c, hipGetLastError());
    return PMX_OK;
}

int kernel_eval_f64(pmx_ctx* c, int kernel, const double* q, int64_t n, double* f, double* df) {
    const dim3 g(grid_for(n, kBlock, 4096)), b(kBlock);
    switch (kernel) {
    case kM4: hipLaunchKernelGGL(k_kernel_eval<kM4>, g, b, 0, c->stream, q, n, f, df); break;
    case kM5: hipLaunchKernelGGL(k_kernel_eval<kM5>, g, b, 0, c->stream, q, n, f, df); break;
    case kM6: hipLaunchKernelGGL(k_kernel_eval<kM6>, g, b, 0, c->stream, q, n, f, df); break;
    case kC2: hipLaunchKernelGGL(k_kernel_eval<kC2>, g, b, 0, c->stream, q, n, f, df); break;
    case kC4: hipLaunchKernelGGL(k_kernel_eval<kC4>, g, b, 0, c->stream, q, n, f, df); break;
    default: hipLaunchKernelGGL(k_kernel_eval<kC6>, g, b, 0, c->stream, q, n, f, df); break;
    }
    PMX_HIP(c, hipGetLastError());
    return PMX_OK;
}

}  // namespace pmx
