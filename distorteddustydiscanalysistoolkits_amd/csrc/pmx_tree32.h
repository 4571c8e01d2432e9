// pmx_tree32.h — FP32 view of the LBVH and the wave-uniform walk shared by the
// FP32-compensated kernels (sample tiles and line-map tiles).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pmx_internal.h"

namespace pmx {

struct T32 {
    const Node32* __restrict__ nodes;
    const float4* __restrict__ leaf;
    const float4* __restrict__ leaflo;
    const int32_t* __restrict__ leaf_esc;
    const double* __restrict__ cols64;     // FP64 x | y | z | m | h | rho | quant (FP64 input only)
    const float* __restrict__ pay;         // m/rho | quant... | rho
    const double* __restrict__ w64;        // m/rho in FP64
    const double* __restrict__ ih64;       // 1/h in FP64
    int64_t n;
    int32_t root;
    int32_t bucket;
    int32_t f32in;                         // FP32 input: FP64 values come from leaf / pay exactly
    int32_t ncol;                          // quant columns
};

// FP64 particle values in Morton order.  An FP32 input's values are exact in the FP32 leaf
// (x y z h, residual zero) and payload (quant..., rho) arrays; an FP64 input reads cols64.
// The branch is on a kernel argument (wavefront-uniform).
__device__ inline void pos64(const T32& T, int64_t p, double& x, double& y, double& z) {
    if (T.f32in) {
        const float4 H = T.leaf[p];
        x = (double)H.x;
        y = (double)H.y;
        z = (double)H.z;
    } else {
        x = T.cols64[p];
        y = T.cols64[T.n + p];
        z = T.cols64[2 * T.n + p];
    }
}
__device__ inline double h64(const T32& T, int64_t p) {
    return T.f32in ? (double)T.leaf[p].w : T.cols64[4 * T.n + p];
}
// Column c = 5 (rho) or 6 + j (quant j).
__device__ inline double col64(const T32& T, int c, int64_t p) {
    if (!T.f32in) return T.cols64[(int64_t)c * T.n + p];
    const int64_t row = c == 5 ? 1 + T.ncol : c - 5;      // payload: m/rho | quant... | rho
    return (double)T.pay[row * T.n + p];
}

__device__ inline uint64_t ballot(bool p) { return __ballot(p); }
// Barrier of one wavefront over its LDS traffic: a wavefront's LDS operations complete in
// issue order, so ordering them in the compiler is enough (workgroups of several independent
// wavefronts must not use __syncthreads inside per-wavefront work).
__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
__device__ inline int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Inclusive prefix sum over the 64 lanes with DPP moves (no LDS): shifts of 1, 2, 4 and 8 lanes
// inside each row of 16 (lanes whose source lies before the row add 0), then lane 15 of each row
// into the next row and lane 31 into rows 2 and 3.
__device__ inline uint32_t wave_inclusive_add(uint32_t v) {
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, false);   // row_shr:1
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, false);   // row_shr:2
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, false);   // row_shr:4
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, false);   // row_shr:8
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false);   // row_bcast:15
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false);   // row_bcast:31
    return v;
}

// Union walk of the tile over left / escape links, from node `start` until the walk leaves
// its subtree (escape link `stop`; 0 from the root).  need(nd) is each lane's conservative
// node test; bucket(lo, cnt) consumes a run of leaves (lo .. lo+cnt-1) in Morton order.
// take(nd, cnt) says whether a needed node is handed over as one run instead of descended;
// emit(node, cnt) may hand a needed node's whole subtree to other wavefronts instead
// (returns true when it did; wave-uniform).
template <class Need, class Bucket, class Take, class Emit>
__device__ inline void walk_sub(const T32& T, int start, int stop, Need&& need, Bucket&& bucket, Take&& take,
                                Emit&& emit) {
    const int64_t n = T.n;
    int node = start;
    int steps = 0;
    while (node > 0 && node != stop && node < 2 * n && ++steps <= 2 * n) {
        if (node >= n) {
            const int li = node - (int)n;
            bucket(li, 1);
            node = uniform(T.leaf_esc[li]);
            continue;
        }
        const Node32 nd = T.nodes[node];
        if (!ballot(need(nd))) {
            node = uniform(nd.escape);
            continue;
        }
        const int lo = uniform(nd.lo), cnt = uniform(nd.cnt);
        if (cnt >= 1 && lo >= 0 && lo + cnt <= n && ballot(take(nd, cnt))) {
            bucket(lo, cnt);
            node = uniform(nd.escape);
            continue;
        }
        if (emit(node, cnt)) {
            node = uniform(nd.escape);
            continue;
        }
        node = uniform(nd.left);
    }
}

template <class Need, class Bucket, class Take>
__device__ inline void walk(const T32& T, bool any_valid, Need&& need, Bucket&& bucket, Take&& take) {
    if (!any_valid) return;
    if (T.root == 0) {                         // n == 1: the reference brute-forces the leaf
        bucket(0, (int)T.n);
        return;
    }
    walk_sub(T, T.root, 0, need, bucket, take, [](int, int) { return false; });
}

template <class Need, class Bucket>
__device__ inline void walk(const T32& T, bool any_valid, Need&& need, Bucket&& bucket) {
    walk(T, any_valid, need, bucket, [&](const Node32&, int cnt) { return cnt <= T.bucket; });
}

// Leaf-run capacities of the walks (their staged particle data in LDS): derivative point tiles
// stage runs of up to 32 (13 KB of LDS per tile, 3 wavefronts per SIMD: C2 63.5 -> 60.2 ms against
// runs of 64 at 2.5 wavefronts per SIMD); scalar point tiles and the line union walk, whose
// occupancy LDS does not limit, runs of up to 64 (the batch interpolation: 143 ms with runs of 32
// against 137 ms).
constexpr int kRunDeriv = 32;
constexpr int kRunScalar = 64;
constexpr int kMaxRunLines = 64;
constexpr int run_cap(bool deriv) { return deriv ? kRunDeriv : kRunScalar; }

// FP32 tree view of the context's current LBVH (prepare_f32 must have run); `bucket` is the
// largest leaf run a walk hands over in one piece, clamped to kMinRun..cap (the walk's staging
// capacity: run_cap(deriv) for point tiles, kMaxRunLines for the line union walk).
T32 tree32(pmx_ctx* c, int bucket, int cap);

}  // namespace pmx
