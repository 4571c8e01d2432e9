// pmx_build.hip — LBVH build on gfx950: bounding box, 63-bit Morton codes, stable radix
// sort, code de-duplication, Karras radix tree with escape links, atomic bottom-up refit,
// and the gather of every particle column into Morton order.
//
// Reference: LinearBVH!(input, Val(3)) (Partia.jl/src/julia/KernelInterpolation/
// interpolation_setup/InterpolationInput.jl:203-224) and the NeighborSearch files cited at
// each kernel.  Codes, the sort permutation and the tree arrays are bit-identical to the
// reference's: the normalisation and quantisation run in FP64 without contraction, the
// sort is stable (the reference's sortperm breaks ties by index), and the de-dup scan is
// the closed form of the reference's serial loop.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <type_traits>

#include "pmx_internal.h"

namespace pmx {
namespace {

constexpr int kBlock = 256;

// ---------------------------------------------------------------------------------
// Bounding box of x, y, z (Base.minimum / Base.maximum in _normalize_vector,
// MortonCode.jl:443-456) and a NaN count (the reference throws for NaN coordinates).
// ---------------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(kBlock) k_bbox_partial(const T* __restrict__ x, const T* __restrict__ y,
                                                         const T* __restrict__ z, int64_t n,
                                                         double* __restrict__ part, int* __restrict__ nan_count) {
    double mn[3] = {__builtin_inf(), __builtin_inf(), __builtin_inf()};
    double mx[3] = {-__builtin_inf(), -__builtin_inf(), -__builtin_inf()};
    int nans = 0;
    auto take = [&](const double (&v)[3]) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            if (v[k] != v[k]) { ++nans; continue; }
            mn[k] = v[k] < mn[k] ? v[k] : mn[k];
            mx[k] = v[k] > mx[k] ? v[k] : mx[k];
        }
    };
    // four elements per thread and trip: twelve loads in flight
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    for (; i + 3 * stride < n; i += 4 * stride) {
        double v[4][3];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            v[u][0] = (double)x[i + u * stride];
            v[u][1] = (double)y[i + u * stride];
            v[u][2] = (double)z[i + u * stride];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) take(v[u]);
    }
    for (; i < n; i += stride) {
        const double v[3] = {(double)x[i], (double)y[i], (double)z[i]};
        take(v);
    }
    __shared__ double s[6][kBlock];
#pragma unroll
    for (int k = 0; k < 3; ++k) { s[k][threadIdx.x] = mn[k]; s[3 + k][threadIdx.x] = mx[k]; }
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                double a = s[k][threadIdx.x + w];
                if (a < s[k][threadIdx.x]) s[k][threadIdx.x] = a;
                double b = s[3 + k][threadIdx.x + w];
                if (b > s[3 + k][threadIdx.x]) s[3 + k][threadIdx.x] = b;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x < 6) part[threadIdx.x * gridDim.x + blockIdx.x] = s[threadIdx.x][0];
    if (nans) atomicAdd(nan_count, nans);
}

__global__ void __launch_bounds__(kBlock) k_bbox_final(const double* __restrict__ part, int nparts,
                                                       double* __restrict__ box) {
    __shared__ double s[6][kBlock];
    double mn[3] = {__builtin_inf(), __builtin_inf(), __builtin_inf()};
    double mx[3] = {-__builtin_inf(), -__builtin_inf(), -__builtin_inf()};
    for (int i = threadIdx.x; i < nparts; i += kBlock) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            double a = part[k * nparts + i], b = part[(3 + k) * nparts + i];
            mn[k] = a < mn[k] ? a : mn[k];
            mx[k] = b > mx[k] ? b : mx[k];
        }
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) { s[k][threadIdx.x] = mn[k]; s[3 + k][threadIdx.x] = mx[k]; }
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                double a = s[k][threadIdx.x + w];
                if (a < s[k][threadIdx.x]) s[k][threadIdx.x] = a;
                double b = s[3 + k][threadIdx.x + w];
                if (b > s[3 + k][threadIdx.x]) s[3 + k][threadIdx.x] = b;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x < 6) box[threadIdx.x] = s[threadIdx.x][0];
}

// _expand_bits3D(::UInt64), MortonCode.jl:209-217.
__device__ inline uint64_t expand_bits3(uint64_t x) {
    x = (x | (x << 32)) & 0x1f00000000ffffULL;
    x = (x | (x << 16)) & 0x1f0000ff0000ffULL;
    x = (x | (x << 8)) & 0x100f00f00f00f00fULL;
    x = (x | (x << 4)) & 0x10c30c30c30c30c3ULL;
    x = (x | (x << 2)) & 0x1249249249249249ULL;
    return x;
}

// FP64 quantisation of one coordinate: f = (v - vmin) * inv(vmax - vmin), 0.5 on a flat
// axis; i = UInt64(floor((exp2(21) - 1) * f)) (MortonCode.jl:436-441, 473-498).
__device__ inline uint64_t quantize(double v, double vmin, double vmax) {
    const double scale = 2097151.0;
    double dv = vmax - vmin;
    double f;
    if (dv == 0.0) {
        f = 0.5;
    } else {
        double inv = 1.0 / dv;
        f = (v - vmin) * inv;
    }
    // f lies in [0, 1], so floor(scale * f) < 2^21: the 32-bit conversion is exact
    return (uint64_t)(uint32_t)floor(scale * f);
}

// ---------------------------------------------------------------------------------
// Stable LSD radix sort of (u64 key, i32 value) pairs, 8-bit digits.  Per pass:
//   k_radix_hist     per-tile digit counts, stored digit-major: hist[d * ntiles + tile]
//   k_radix_scan     one workgroup per digit: exclusive scan over tiles, digit totals
//   k_radix_scatter  tile loaded to registers (each wavefront one contiguous quarter),
//                    stable ranking with wavefront ballots and per-wave running digit
//                    counts, keys staged in LDS in digit order, then written out so that
//                    consecutive lanes store consecutive addresses of each digit run.
// Stability: items are ranked in (wave, round, lane) order = input order within a tile,
// and tiles are laid out in order inside each digit.
// ---------------------------------------------------------------------------------
constexpr int kRadixBits = 8;
constexpr int kBins = 1 << kRadixBits;
constexpr int kItems = 16;
constexpr int kTile = kBlock * kItems;   // 4096
constexpr int kWaves = kBlock / 64;

// Input columns of the build (x y z m h rho quant...), each of the input type.
struct ColPtrs {
    const void* p[24];
};

// _encode_morton_code3D (MortonCode.jl:275-305): code = ex<<2 | ey<<1 | ez (the sort generates the
// index values itself in its first pass).
// One workgroup per 4096-key tile of the sort, which also counts the tile's first digit
// (bits 0-7) as k_radix_hist would, so the sort's first pass needs no histogram pass, and writes
// each particle's interleaved input record (x y z m h rho quant...; the columns are then read
// once, and the record gather after the sort takes one random record per particle).  NC > 0:
// NC columns known at compile time, records stored with the widest aligned vector stores their
// size allows; NC = 0: ncols at run time.
template <typename T, int NC>
__global__ void __launch_bounds__(kBlock) k_encode(ColPtrs src, int ncols_rt, int64_t n,
                                                   const double* __restrict__ box, int ntiles,
                                                   uint64_t* __restrict__ codes,
                                                   uint32_t* __restrict__ hist, T* __restrict__ rec) {
    __shared__ uint32_t cnt[kBins];
    cnt[threadIdx.x] = 0;
    const int ncols = NC > 0 ? NC : ncols_rt;
    const T* __restrict__ x = static_cast<const T*>(src.p[0]);
    const T* __restrict__ y = static_cast<const T*>(src.p[1]);
    const T* __restrict__ z = static_cast<const T*>(src.p[2]);
    const double x0 = box[0], y0 = box[1], z0 = box[2], x1 = box[3], y1 = box[4], z1 = box[5];
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * kTile;
#pragma unroll 2
    for (int r = 0; r < kItems; ++r) {
        const int64_t i = base + r * kBlock + threadIdx.x;
        if (i < n) {
            const T xv = x[i], yv = y[i], zv = z[i];
            if constexpr (NC > 0) {
                // records start at multiples of NC * sizeof(T) bytes from a 256-byte aligned array
                constexpr int kBytes = NC * (int)sizeof(T);
                constexpr int kVec = kBytes % 16 == 0 ? 16 : (kBytes % 8 == 0 ? 8 : 4);
                using V = typename std::conditional<kVec == 16, int4, typename std::conditional<kVec == 8, int2, int>::type>::type;
                T rv[NC];
                rv[0] = xv;
                rv[1] = yv;
                rv[2] = zv;
#pragma unroll
                for (int c = 3; c < NC; ++c) rv[c] = static_cast<const T*>(src.p[c])[i];
                V* out = reinterpret_cast<V*>(rec + i * NC);
#pragma unroll
                for (int k = 0; k < kBytes / kVec; ++k) {
                    V v;
                    __builtin_memcpy(&v, reinterpret_cast<const char*>(rv) + k * kVec, kVec);
                    out[k] = v;
                }
            } else {
                T* out = rec + i * ncols;
                out[0] = xv;
                out[1] = yv;
                out[2] = zv;
                for (int c = 3; c < ncols; ++c) out[c] = static_cast<const T*>(src.p[c])[i];
            }
            const uint64_t ix = quantize((double)xv, x0, x1);
            const uint64_t iy = quantize((double)yv, y0, y1);
            const uint64_t iz = quantize((double)zv, z0, z1);
            const uint64_t code = (expand_bits3(ix) << 2) | (expand_bits3(iy) << 1) | expand_bits3(iz);
            codes[i] = code;
            atomicAdd(&cnt[code & (kBins - 1)], 1u);
        }
    }
    __syncthreads();
    hist[(int64_t)threadIdx.x * ntiles + blockIdx.x] = cnt[threadIdx.x];
}

// Each wavefront counts into its own LDS histogram (fewer lanes of one atomic instruction meet on
// one counter), reading two keys per lane and load.
__global__ void __launch_bounds__(kBlock) k_radix_hist(const uint64_t* __restrict__ keys, int64_t n, int shift,
                                                       int ntiles, uint32_t* __restrict__ hist) {
    __shared__ uint32_t cnt[kWaves][kBins];
#pragma unroll
    for (int w = 0; w < kWaves; ++w) cnt[w][threadIdx.x] = 0;
    __syncthreads();
    uint32_t* mine = cnt[threadIdx.x >> 6];
    const int64_t base = (int64_t)blockIdx.x * kTile;
    if (base + kTile <= n) {
        const ulonglong2* k2 = reinterpret_cast<const ulonglong2*>(keys + base);
#pragma unroll
        for (int r = 0; r < kItems / 2; ++r) {
            const ulonglong2 v = k2[r * kBlock + threadIdx.x];
            atomicAdd(&mine[(v.x >> shift) & (kBins - 1)], 1u);
            atomicAdd(&mine[(v.y >> shift) & (kBins - 1)], 1u);
        }
    } else {
#pragma unroll
        for (int r = 0; r < kItems; ++r) {
            const int64_t i = base + r * kBlock + threadIdx.x;
            if (i < n) atomicAdd(&mine[(keys[i] >> shift) & (kBins - 1)], 1u);
        }
    }
    __syncthreads();
    uint32_t tot = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) tot += cnt[w][threadIdx.x];
    hist[(int64_t)threadIdx.x * ntiles + blockIdx.x] = tot;
}

// Block-wide exclusive scan of one uint32 per thread; returns the block total.
__device__ inline uint32_t block_exclusive_scan(uint32_t v, uint32_t* s_tmp, uint32_t& excl) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t inc = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t t = __shfl_up(inc, o, 64);
        if (lane >= o) inc += t;
    }
    if (lane == 63) s_tmp[wave] = inc;
    __syncthreads();
    uint32_t wbase = 0, total = 0;
    for (int w = 0; w < kWaves; ++w) {
        uint32_t t = s_tmp[w];
        if (w < wave) wbase += t;
        total += t;
    }
    __syncthreads();
    excl = wbase + inc - v;
    return total;
}

__global__ void __launch_bounds__(kBlock) k_radix_scan(uint32_t* __restrict__ hist, int ntiles,
                                                       uint32_t* __restrict__ totals) {
    __shared__ uint32_t s_tmp[kWaves];
    uint32_t* row = hist + (int64_t)blockIdx.x * ntiles;
    uint32_t carry = 0;
    for (int base = 0; base < ntiles; base += kBlock * 4) {
        uint32_t v[4], sum = 0;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            int i = base + threadIdx.x * 4 + j;
            v[j] = i < ntiles ? row[i] : 0u;
            sum += v[j];
        }
        uint32_t excl;
        uint32_t total = block_exclusive_scan(sum, s_tmp, excl);
        uint32_t run = carry + excl;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            int i = base + threadIdx.x * 4 + j;
            if (i < ntiles) row[i] = run;
            run += v[j];
        }
        carry += total;
    }
    if (threadIdx.x == 0) totals[blockIdx.x] = carry;
}

template <bool IDX>
__global__ void __launch_bounds__(kBlock, 4) k_radix_scatter(const uint64_t* __restrict__ keys_in,
                                                          const int32_t* __restrict__ vals_in, int64_t n,
                                                          int shift, int ntiles,
                                                          const uint32_t* __restrict__ hist,
                                                          const uint32_t* __restrict__ totals,
                                                          uint64_t* __restrict__ keys_out,
                                                          int32_t* __restrict__ vals_out) {
    constexpr int kPerWave = kTile / kWaves;   // each wavefront ranks one contiguous quarter
    // keys staged in digit order, then (the same bytes) the values: 38 KB of LDS, 4 workgroups
    // per CU
    __shared__ uint64_t s_stage[kTile];
    __shared__ uint32_t s_gofs[kBins];     // global destination of this tile's digit run
    __shared__ uint32_t s_lstart[kBins];   // start of the digit run inside the tile
    __shared__ uint32_t s_wc[kWaves][kBins];   // per-wave digit counts, then per-wave run bases
    __shared__ uint32_t s_tmp[kWaves];

    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int64_t base = (int64_t)blockIdx.x * kTile;
    const int64_t wbase = base + (int64_t)wave * kPerWave;

    // global offsets: digit base (scan of totals) + this tile's prefix within the digit
    uint32_t tot = totals[t], dbase;
    block_exclusive_scan(tot, s_tmp, dbase);
    s_gofs[t] = dbase + hist[(int64_t)t * ntiles + blockIdx.x];
#pragma unroll
    for (int w = 0; w < kWaves; ++w) s_wc[w][t] = 0;
    __syncthreads();

    uint64_t k[kItems];
    uint32_t dr[kItems];     // digit << 16 | rank inside the wave's digit run; 0xffffffff: no item
#pragma unroll
    for (int r = 0; r < kItems; ++r) {
        const int64_t i = wbase + r * 64 + lane;
        k[r] = i < n ? keys_in[i] : 0ull;
    }
    // rank inside the wavefront's quarter: lanes of equal digit found by ballots, a running
    // count per digit kept by the wavefront (its own LDS row, no block barriers)
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
#pragma unroll
    for (int r = 0; r < kItems; ++r) {
        const bool valid = wbase + r * 64 + lane < n;
        const int d = (int)((k[r] >> shift) & (kBins - 1));
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < kRadixBits; ++b) {
            const uint64_t bal = __ballot(valid && ((d >> b) & 1));
            peers &= ((d >> b) & 1) ? bal : ~bal;
        }
        const uint32_t before = (uint32_t)__popcll(peers & lt_mask);
        uint32_t run = 0;
        if (valid) run = s_wc[wave][d];
        __builtin_amdgcn_wave_barrier();
        if (valid && before == 0) s_wc[wave][d] = run + (uint32_t)__popcll(peers);
        __builtin_amdgcn_wave_barrier();
        dr[r] = valid ? ((uint32_t)d << 16 | (run + before)) : 0xffffffffu;
    }
    __syncthreads();
    // digit t: bases of the four quarters inside the digit's run, run starts inside the tile
    uint32_t cnt = 0, wb[kWaves];
#pragma unroll
    for (int w = 0; w < kWaves; ++w) { wb[w] = cnt; cnt += s_wc[w][t]; }
    uint32_t lst;
    block_exclusive_scan(cnt, s_tmp, lst);
    s_lstart[t] = lst;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) s_wc[w][t] = lst + wb[w];
    __syncthreads();
    // slots of this thread's items; keys into the staging buffer
#pragma unroll
    for (int r = 0; r < kItems; ++r) {
        if (dr[r] != 0xffffffffu) {
            const uint32_t slot = s_wc[wave][dr[r] >> 16] + (dr[r] & 0xffffu);
            dr[r] = slot < (uint32_t)kTile ? slot : 0xffffffffu;
            if (dr[r] != 0xffffffffu) s_stage[slot] = k[r];
        }
    }
    __syncthreads();
    // keys out (consecutive lanes store consecutive addresses of a digit run), keeping each
    // staged slot's destination for the values
    const int count = (int)((n - base) < kTile ? (n - base) : kTile);
    uint32_t dst[kItems];
#pragma unroll
    for (int q = 0; q < kItems; ++q) {
        const int j = t + q * kBlock;
        dst[q] = 0xffffffffu;
        if (j < count) {
            const uint64_t key = s_stage[j];
            const int d = (int)((key >> shift) & (kBins - 1));
            dst[q] = s_gofs[d] + (uint32_t)(j - s_lstart[d]);
            if (dst[q] < (uint64_t)n) keys_out[dst[q]] = key;   // always true for a consistent histogram
        }
    }
    __syncthreads();
    // the values (read only now: fewer registers live through the ranking)
    int32_t* s_v = reinterpret_cast<int32_t*>(s_stage);
#pragma unroll
    for (int r = 0; r < kItems; ++r)
        if (dr[r] != 0xffffffffu) {
            if constexpr (IDX) s_v[dr[r]] = (int32_t)(wbase + r * 64 + lane);   // the values are the indices
            else s_v[dr[r]] = vals_in[wbase + r * 64 + lane];
        }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kItems; ++q) {
        const int j = t + q * kBlock;
        if (j < count && dst[q] < (uint64_t)n) vals_out[dst[q]] = s_v[j];
    }
}

// ---------------------------------------------------------------------------------
// De-duplication (sort_by_morton!, MortonCode.jl:189-193): c'_i = max(c_i, c'_{i-1} + 1)
// has the closed form c'_i = i + max_{j <= i}(c_j - j), computed as a max-scan.  Sorted codes
// without two equal neighbours increase by at least 1 per index, so c_j - j never decreases and
// c' = c: the tile pass flags equal neighbours (dup, zeroed with the NaN counter) and the scan
// and apply passes return at once when there are none.
// ---------------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_dedup_tilemax(const uint64_t* __restrict__ c, int64_t n,
                                                          int64_t* __restrict__ tmax, int* __restrict__ dup) {
    __shared__ int64_t s[kBlock];
    const int64_t base = (int64_t)blockIdx.x * kTile;
    int64_t m = INT64_MIN;
    int eq = 0;
#pragma unroll
    for (int r = 0; r < kItems; ++r) {
        int64_t i = base + r * kBlock + threadIdx.x;
        if (i < n) {
            const uint64_t ci = c[i];
            int64_t d = (int64_t)ci - i;
            m = d > m ? d : m;
            if (i > 0 && c[i - 1] == ci) eq = 1;
        }
    }
    if (__syncthreads_or(eq) && threadIdx.x == 0) *dup = 1;
    s[threadIdx.x] = m;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w && s[threadIdx.x + w] > s[threadIdx.x]) s[threadIdx.x] = s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) tmax[blockIdx.x] = s[0];
}

// exclusive max-scan of tile maxima (single workgroup)
__global__ void __launch_bounds__(kBlock) k_dedup_scan(int64_t* __restrict__ tmax, int ntiles,
                                                       const int* __restrict__ dup) {
    if (*dup == 0) return;
    __shared__ int64_t s[kBlock];
    int64_t carry = INT64_MIN;
    for (int base = 0; base < ntiles; base += kBlock) {
        int i = base + threadIdx.x;
        int64_t v = i < ntiles ? tmax[i] : INT64_MIN;
        s[threadIdx.x] = v;
        __syncthreads();
        for (int o = 1; o < kBlock; o <<= 1) {
            int64_t t = (int)threadIdx.x >= o ? s[threadIdx.x - o] : INT64_MIN;
            __syncthreads();
            if (t > s[threadIdx.x]) s[threadIdx.x] = t;
            __syncthreads();
        }
        int64_t excl = threadIdx.x == 0 ? INT64_MIN : s[threadIdx.x - 1];
        excl = carry > excl ? carry : excl;
        if (i < ntiles) tmax[i] = excl;
        int64_t last = s[kBlock - 1];
        __syncthreads();
        carry = last > carry ? last : carry;
    }
}

__global__ void __launch_bounds__(kBlock) k_dedup_apply(uint64_t* __restrict__ c, int64_t n,
                                                        const int64_t* __restrict__ tpre, const int* __restrict__ dup) {
    if (*dup == 0) return;
    // Four rounds of kBlock x 4 consecutive codes (coalesced): per-thread running max, a
    // wavefront max-scan, then the wavefronts combined with the carry of earlier rounds.
    constexpr int kVec = 4;
    __shared__ int64_t s_w[kWaves];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    int64_t carry = tpre[blockIdx.x];
    for (int r = 0; r < kItems / kVec; ++r) {
        const int64_t i0 = (int64_t)blockIdx.x * kTile + (int64_t)r * kBlock * kVec + (int64_t)t * kVec;
        int64_t d[kVec], loc = INT64_MIN;
#pragma unroll
        for (int q = 0; q < kVec; ++q) {
            d[q] = i0 + q < n ? (int64_t)c[i0 + q] - (i0 + q) : INT64_MIN;
            loc = d[q] > loc ? d[q] : loc;
        }
        int64_t inc = loc;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int64_t u = __shfl_up(inc, o, 64);
            if (lane >= o && u > inc) inc = u;
        }
        int64_t excl = __shfl_up(inc, 1, 64);
        if (lane == 0) excl = INT64_MIN;
        if (lane == 63) s_w[wave] = inc;
        __syncthreads();
        int64_t run = carry, all = carry;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) {
            const int64_t v = s_w[w];
            if (w < wave && v > run) run = v;
            if (v > all) all = v;
        }
        __syncthreads();
        run = excl > run ? excl : run;
#pragma unroll
        for (int q = 0; q < kVec; ++q) {
            run = d[q] > run ? d[q] : run;
            if (i0 + q < n) c[i0 + q] = (uint64_t)(run + i0 + q);
        }
        carry = all;
    }
}

// ---------------------------------------------------------------------------------
// Column gather into Morton order (the permute! calls of InterpolationInput.jl:211-219),
// converting to FP64.  Column c of the output lives at cols64 + c * n.
// ---------------------------------------------------------------------------------
// The input columns were interleaved into one record per particle by k_encode (coalesced
// reads and writes), so the permutation reads one record per particle instead of one element
// per column from unrelated places.

// Also writes the FP32 leaf data of the compensated path: {x_hi, y_hi, z_hi, h} and
// {x_lo, y_lo, z_lo, 1/h} (hi + lo = the FP64 coordinate to 2^-48), and the payload columns
// m/rho | quant... | rho in FP32 followed by m/rho in FP64.  An FP32 input's FP64 columns are
// not written: the leaf (residual zero) and payload arrays hold its values exactly, and
// k_gather_cols64 writes them when a caller needs the FP64 layout.
struct Leaf32Out {
    float4* leaf;
    float4* leaflo;
    float* pay;
    double* w64;
    double* ih64;
};

// NC > 0: records of NC columns known at compile time, read with the widest aligned vector loads
// the record size allows (a random record then costs one or two load instructions instead of one
// per column); NC = 0: ncols at run time, one load per column.
template <typename T, int NC>
__global__ void __launch_bounds__(kBlock) k_gather_records(const T* __restrict__ rec, int ncols_rt,
                                                           const int32_t* __restrict__ order, int64_t n,
                                                           double* __restrict__ cols64, Leaf32Out o32) {
    const int ncols = NC > 0 ? NC : ncols_rt;
    const int ncol = ncols - 6;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
        int32_t o = order[i];
        o = ((uint64_t)(uint32_t)o < (uint64_t)n) ? o : 0;   // in-bounds guard (order is a permutation)
        const T* r = rec + (int64_t)o * ncols;
        constexpr bool kCols = sizeof(T) == 8;
        double v[6];
        if constexpr (NC > 0) {
            // records start at multiples of NC * sizeof(T) bytes from a 256-byte aligned array
            constexpr int kBytes = NC * (int)sizeof(T);
            constexpr int kVec = kBytes % 16 == 0 ? 16 : (kBytes % 8 == 0 ? 8 : 4);
            using V = typename std::conditional<kVec == 16, int4, typename std::conditional<kVec == 8, int2, int>::type>::type;
            T rv[NC];
            const V* rvp = reinterpret_cast<const V*>(r);
#pragma unroll
            for (int k = 0; k < kBytes / kVec; ++k) {
                const V x = rvp[k];
                __builtin_memcpy(reinterpret_cast<char*>(rv) + k * kVec, &x, kVec);
            }
#pragma unroll
            for (int c = 0; c < 6; ++c) {
                v[c] = (double)rv[c];
                if (kCols) cols64[(int64_t)c * n + i] = v[c];
            }
#pragma unroll
            for (int c = 6; c < NC; ++c) {
                const double q = (double)rv[c];
                if (kCols) cols64[(int64_t)c * n + i] = q;
                o32.pay[(int64_t)(c - 5) * n + i] = (float)q;
            }
        } else {
#pragma unroll
            for (int c = 0; c < 6; ++c) {
                v[c] = (double)r[c];
                if (kCols) cols64[(int64_t)c * n + i] = v[c];
            }
            for (int c = 6; c < ncols; ++c) {
                const double q = (double)r[c];
                if (kCols) cols64[(int64_t)c * n + i] = q;
                o32.pay[(int64_t)(c - 5) * n + i] = (float)q;
            }
        }
        const float xh = (float)v[0], yh = (float)v[1], zh = (float)v[2];
        o32.leaf[i] = make_float4(xh, yh, zh, (float)v[4]);
        o32.leaflo[i] = make_float4((float)(v[0] - (double)xh), (float)(v[1] - (double)yh),
                                    (float)(v[2] - (double)zh), (float)(1.0 / v[4]));
        const double w = v[3] / v[5];
        o32.pay[i] = (float)w;
        o32.pay[(int64_t)(1 + ncol) * n + i] = (float)v[5];
        o32.w64[i] = w;
        o32.ih64[i] = 1.0 / v[4];
    }
}

// Record gather with the record width known at compile time for 0 to 4 quant columns.
template <typename T>
void launch_gather(pmx_ctx* c, const T* rec, int ncols, const int32_t* order, int64_t n, double* cols64,
                   const Leaf32Out& o32, unsigned g) {
    hipStream_t s = c->stream;
    switch (ncols) {
    case 6: hipLaunchKernelGGL((k_gather_records<T, 6>), dim3(g), dim3(kBlock), 0, s, rec, ncols, order, n, cols64, o32); break;
    case 7: hipLaunchKernelGGL((k_gather_records<T, 7>), dim3(g), dim3(kBlock), 0, s, rec, ncols, order, n, cols64, o32); break;
    case 8: hipLaunchKernelGGL((k_gather_records<T, 8>), dim3(g), dim3(kBlock), 0, s, rec, ncols, order, n, cols64, o32); break;
    case 9: hipLaunchKernelGGL((k_gather_records<T, 9>), dim3(g), dim3(kBlock), 0, s, rec, ncols, order, n, cols64, o32); break;
    case 10: hipLaunchKernelGGL((k_gather_records<T, 10>), dim3(g), dim3(kBlock), 0, s, rec, ncols, order, n, cols64, o32); break;
    default: hipLaunchKernelGGL((k_gather_records<T, 0>), dim3(g), dim3(kBlock), 0, s, rec, ncols, order, n, cols64, o32); break;
    }
}

// The reference's in-place permute! of the input columns (InterpolationInput.jl:211-219) for
// columns in device memory: column c of the caller becomes rec[order[i]][c].  The records are a
// copy of the columns, so the columns can be overwritten in place.
struct ColOut {
    void* p[24];
};

template <typename T>
__global__ void __launch_bounds__(kBlock) k_writeback(const T* __restrict__ rec, int ncols,
                                                      const int32_t* __restrict__ order, int64_t n, ColOut dst) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
        int32_t o = order[i];
        o = ((uint64_t)(uint32_t)o < (uint64_t)n) ? o : 0;
        const T* r = rec + (int64_t)o * ncols;
        for (int c = 0; c < ncols; ++c) static_cast<T*>(dst.p[c])[i] = r[c];
    }
}

// FP64 columns of an FP32 input in Morton order, from the packed records.
__global__ void __launch_bounds__(kBlock) k_gather_cols64(const float* __restrict__ rec, int ncols,
                                                          const int32_t* __restrict__ order, int64_t n,
                                                          double* __restrict__ cols64) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
        int32_t o = order[i];
        o = ((uint64_t)(uint32_t)o < (uint64_t)n) ? o : 0;
        const float* r = rec + (int64_t)o * ncols;
        for (int c = 0; c < ncols; ++c) cols64[(int64_t)c * n + i] = (double)r[c];
    }
}

__global__ void __launch_bounds__(kBlock) k_pack_leaf(const double* __restrict__ cols64, int64_t n,
                                                      Leaf64* __restrict__ leaf) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
        Leaf64 L;
        L.x = cols64[i];
        L.y = cols64[n + i];
        L.z = cols64[2 * n + i];
        L.h = cols64[4 * n + i];
        leaf[i] = L;
    }
}

// ---------------------------------------------------------------------------------
// Karras radix tree (BinaryRadixTree.jl:81-281).  Unified ids: internal 1..n-1, leaf k
// (1-based) -> n-1+k.  All index arithmetic below is 1-based as in the reference.
// ---------------------------------------------------------------------------------
// _build_child! (BinaryRadixTree.jl:209-243) with _find_range (:148-179) and
// _split_position (:181-202), and the internal escape of Algorithm 2 of Prokopenko &
// Lebrun-Grandie (_build_adjacent! + escapes, BinaryRadixTree.jl:204-260): with
// adj[k] = delta(k, k + 1) (typemax(Int32) at k = n), node i's escape follows from the two
// adjacency values around the end of its range, computed here from the codes directly.
// lr holds left at 2*id and right at 2*id+1.
// A workgroup's nodes read codes near their own ids (ranges of small nodes, the first steps of
// the searches): a window of codes around the workgroup's ids is staged in LDS, and reads
// outside it go to global memory.
constexpr int kBrtPad = 128;

// unread (min_run > 0): per internal node, 1 when its parent covers at most min_run leaves, so
// that no FP32 walk reads it (k_refit_small writes no record for it).  Each node writes the flag
// of its internal children (every internal node but the root has one parent; the root's is 0).
__global__ void __launch_bounds__(kBlock) k_brt(const uint64_t* __restrict__ c, int64_t n,
                                                int32_t* __restrict__ lr, int32_t* __restrict__ parent,
                                                int32_t* __restrict__ range_far,
                                                int32_t* __restrict__ escape, int32_t* __restrict__ leaf_esc,
                                                uint8_t* __restrict__ unread, int min_run) {
    __shared__ uint64_t win[kBlock + 2 * kBrtPad];
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    const int64_t m = n - 1;                                  // internal nodes 1 .. n-1
    const int64_t iters = (m + stride - 1) / stride;
    for (int64_t itr = 0; itr < iters; ++itr) {
        const int64_t i0 = 1 + itr * stride + (int64_t)blockIdx.x * kBlock;   // first id of the block
        // window: codes of 1-based indices w0 .. w1-1
        const int64_t w0 = i0 - kBrtPad > 1 ? i0 - kBrtPad : 1;
        const int64_t w1 = i0 + kBlock + kBrtPad <= n + 1 ? i0 + kBlock + kBrtPad : n + 1;
        __syncthreads();
        for (int64_t k = w0 + threadIdx.x; k < w1; k += kBlock) win[k - w0] = c[k - 1];
        __syncthreads();
        auto code = [&](int64_t j) -> uint64_t { return (j >= w0 && j < w1) ? win[j - w0] : c[j - 1]; };
        // _longest_common_prefix_length(codes, i, j), MortonCode.jl:380-387 (1-based)
        auto delta = [&](int64_t a, int64_t b) -> int {
            const uint64_t ca = code(a), cb = code(b);
            if (ca == cb) return __clzll((uint64_t)a ^ (uint64_t)b);
            return __clzll(ca ^ cb);
        };
        const int64_t i = i0 + threadIdx.x;
        if (i > n - 1) continue;
        int dl = (i > 1) ? delta(i, i - 1) : -1;
        int dr = (i < n) ? delta(i, i + 1) : -1;
        int64_t d = (dr - dl > 0) ? 1 : ((dr - dl < 0) ? -1 : 1);
        int dmin = (1 <= i - d && i - d <= n) ? delta(i, i - d) : -1;
        int64_t lmax = 2;
        for (;;) {
            int64_t j = i + lmax * d;
            if (j < 1 || j > n) break;
            if (delta(i, j) <= dmin) break;
            lmax *= 2;
        }
        int64_t l = 0;
        for (int64_t t = lmax >> 1; t > 0; t >>= 1) {
            int64_t j = i + (l + t) * d;
            if (j >= 1 && j <= n && delta(i, j) > dmin) l += t;
        }
        int64_t j = i + l * d;
        int64_t lo = i < j ? i : j, hi = i < j ? j : i;
        range_far[i] = (int32_t)j;     // the range's other end (one end is i)
        int64_t split;
        if (code(lo) == code(hi)) {
            split = (lo + hi) >> 1;
            split = split < lo ? lo : (split > hi - 1 ? hi - 1 : split);
        } else {
            int prefix = delta(lo, hi);
            split = lo;
            int64_t step = hi - lo;
            while (step > 1) {
                step = (step + 1) >> 1;
                int64_t ns = split + step;
                if (ns < hi && delta(lo, ns) > prefix) split = ns;
            }
        }
        const int64_t off = n - 1;
        int64_t left = (split == lo) ? lo + off : split;
        int64_t right = (split + 1 == hi) ? hi + off : split + 1;
        lr[2 * i] = (int32_t)left;
        lr[2 * i + 1] = (int32_t)right;
        if (min_run > 0) {
            const uint8_t u = hi - lo + 1 <= min_run ? 1 : 0;
            if (split != lo) unread[split] = u;
            if (split + 1 != hi) unread[split + 1] = u;
            if (i == 1) unread[1] = 0;
        }
        parent[left] = (int32_t)i;
        parent[right] = (int32_t)i;
        int32_t e;
        if (hi == n) e = 0;
        else if (hi == n - 1) e = (int32_t)(off + n);
        else {
            const int64_t next = hi + 1;     // adj[next] = delta(next, next + 1), next + 1 <= n
            e = (delta(next, next + 1) < delta(next - 1, next)) ? (int32_t)(off + next) : (int32_t)next;
        }
        escape[i] = e;
        // leaf escapes (BinaryRadixTree.jl:262-281) by the rule of the internal ones: the
        // top-most node starting right after the leaf (a left child's right sibling and a right
        // child's parent's escape both start there).  Only leaf_esc holds them: no kernel reads
        // the leaves' entries of escape / lr (the tree export fills them from leaf_esc and 0).
        auto escape_after = [&](int64_t h) -> int32_t {
            if (h == n) return 0;
            if (h == n - 1) return (int32_t)(off + n);
            const int64_t nx = h + 1;
            return (delta(nx, nx + 1) < delta(nx - 1, nx)) ? (int32_t)(off + nx) : (int32_t)nx;
        };
        auto leaf_out = [&](int64_t li) { leaf_esc[li] = escape_after(li + 1); };
        leaf_out(i - 1);
        if (i == n - 1) leaf_out(n - 1);
    }
}

// Leaf escapes (BinaryRadixTree.jl:262-281): left child -> right sibling, right child ->
// escape of the parent, from the parent links; also the leaves' (zero) child links.  Used for
// the one-leaf tree only (k_brt computes every other tree's leaf escapes).
__global__ void __launch_bounds__(kBlock) k_escape_leaf(int32_t* __restrict__ lr,
                                                        const int32_t* __restrict__ parent, int64_t n,
                                                        int32_t* __restrict__ escape, int32_t* __restrict__ leaf_esc) {
    for (int64_t li = (int64_t)blockIdx.x * kBlock + threadIdx.x; li < n; li += (int64_t)gridDim.x * kBlock) {
        int64_t id = n + li;   // = (n - 1) + (li + 1)
        int32_t p = parent[id];
        int32_t e;
        if (p == 0) e = 0;
        else e = (lr[2 * p] == (int32_t)id) ? lr[2 * p + 1] : escape[p];
        escape[id] = e;
        leaf_esc[li] = e;
        lr[2 * id] = 0;
        lr[2 * id + 1] = 0;
    }
}

// The tree entries no construction kernel writes: the child links of id 0 and the parent of the
// root (id 1; also the only leaf when n = 1), and escape[0].
__global__ void k_tree_heads(int32_t* __restrict__ lr, int32_t* __restrict__ parent, int32_t* __restrict__ escape) {
    if (threadIdx.x == 0) {
        lr[0] = 0;
        lr[1] = 0;
        parent[0] = 0;
        parent[1] = 0;
        escape[0] = 0;
    }
}

__global__ void __launch_bounds__(kBlock) k_fill_u32(uint32_t* __restrict__ p, int64_t n, uint32_t v) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) p[i] = v;
}

// Bottom-up refit (LinearBVH.jl:78-80, _build_internal_aabb! :98-156), in rounds.
// A climber sits on a completed node; it moves to the parent and the second arrival at
// a parent combines both children with strict comparisons (ifelse(<) / ifelse(>)),
// writes the parent and continues in the next round.  Round r only reads nodes completed
// in rounds < r (or leaves), and consecutive rounds are separate kernel launches on one
// stream, so node data needs no fences; the arrival counter is the only shared atomic.
// Round 0 starts from every leaf; later rounds from the list the previous round appended.
// The FP32 copy of a node used by the FP32-compensated traversals: box rounded outward,
// hmax rounded up (every FP32 node test stays conservative), plus its 0-based leaf run.
// `nodes` is null when only the FP32 tree is built (the FP64 nodes are refitted on first use).
// Leaf range [lo, hi] (1-based) of internal node i: one end is i itself, the other range_far[i].
__device__ inline void node_range(const int32_t* __restrict__ far, int64_t i, int32_t& lo, int32_t& hi) {
    const int32_t j = far[i], k = (int32_t)i;
    lo = j < k ? j : k;
    hi = j < k ? k : j;
}

__device__ inline void store_node(const Node64& a, int64_t i, const int32_t* __restrict__ far, Node64* nodes,
                                  Node32* __restrict__ n32) {
    int32_t lo, hi;
    node_range(far, i, lo, hi);
    if (nodes) nodes[i] = a;
    Node32 b;
    b.mnx = f32_down(a.mn[0]); b.mny = f32_down(a.mn[1]); b.mnz = f32_down(a.mn[2]);
    b.mxx = f32_up(a.mx[0]); b.mxy = f32_up(a.mx[1]); b.mxz = f32_up(a.mx[2]);
    b.hmax = f32_up(a.hmax);
    b.left = a.left;
    b.escape = a.escape;
    b.lo = lo - 1;
    b.cnt = hi - lo + 1;
    b.pad = 0;
    n32[i] = b;
}

// Box and h of leaf li: from cols64 (FP64 input) or exactly from the FP32 leaf (FP32 input).
__device__ inline Leaf64 read_leaf(const double* __restrict__ cols64, const float4* __restrict__ leaf32, int64_t n,
                                   int64_t li) {
    Leaf64 L;
    if (leaf32) {
        const float4 H = leaf32[li];
        L.x = (double)H.x;
        L.y = (double)H.y;
        L.z = (double)H.z;
        L.h = (double)H.w;
    } else {
        L.x = cols64[li];
        L.y = cols64[n + li];
        L.z = cols64[2 * n + li];
        L.h = cols64[4 * n + li];
    }
    return L;
}

__device__ inline void combine_node(const double* __restrict__ cols64, const float4* __restrict__ leaf32,
                                    const int32_t* __restrict__ lr,
                                    const int32_t* __restrict__ escape, int64_t n, Node64* nodes, int32_t p,
                                    const int32_t* __restrict__ far, Node32* __restrict__ n32) {
    const int32_t l = lr[2 * p], r = lr[2 * p + 1];
    double lmn[3], lmx[3], rmn[3], rmx[3], hl, hr;
    if (l >= n) {
        const Leaf64 L = read_leaf(cols64, leaf32, n, l - n);
        lmn[0] = lmx[0] = L.x; lmn[1] = lmx[1] = L.y; lmn[2] = lmx[2] = L.z;
        hl = L.h;
    } else if (nodes) {
        const Node64 N = nodes[l];
        for (int k = 0; k < 3; ++k) { lmn[k] = N.mn[k]; lmx[k] = N.mx[k]; }
        hl = N.hmax;
    } else {
        // FP32 tree only: the child's outward-rounded box.  Rounding is monotone, so the
        // rounded min / max of the rounded children equal the rounding of the FP64 min / max
        const Node32 N = n32[l];
        lmn[0] = N.mnx; lmn[1] = N.mny; lmn[2] = N.mnz;
        lmx[0] = N.mxx; lmx[1] = N.mxy; lmx[2] = N.mxz;
        hl = N.hmax;
    }
    if (r >= n) {
        const Leaf64 L = read_leaf(cols64, leaf32, n, r - n);
        rmn[0] = rmx[0] = L.x; rmn[1] = rmx[1] = L.y; rmn[2] = rmx[2] = L.z;
        hr = L.h;
    } else if (nodes) {
        const Node64 N = nodes[r];
        for (int k = 0; k < 3; ++k) { rmn[k] = N.mn[k]; rmx[k] = N.mx[k]; }
        hr = N.hmax;
    } else {
        const Node32 N = n32[r];
        rmn[0] = N.mnx; rmn[1] = N.mny; rmn[2] = N.mnz;
        rmx[0] = N.mxx; rmx[1] = N.mxy; rmx[2] = N.mxz;
        hr = N.hmax;
    }
    Node64 out;
    for (int k = 0; k < 3; ++k) {
        out.mn[k] = lmn[k] < rmn[k] ? lmn[k] : rmn[k];
        out.mx[k] = lmx[k] > rmx[k] ? lmx[k] : rmx[k];
    }
    out.hmax = hl > hr ? hl : hr;
    out.left = l;
    out.escape = escape[p];
    store_node(out, p, far, nodes, n32);
}

// Append `p` (when `take`) to the next round's list with one atomic per wavefront.
__device__ inline void wave_append(bool take, int32_t p, int32_t* __restrict__ list, uint32_t* __restrict__ count) {
    const uint64_t m = __ballot(take);
    if (!m) return;
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll((unsigned long long)m) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(count, (uint32_t)__popcll(m));
    base = __shfl(base, leader, 64);
    if (take) list[base + __popcll(m & ((lane == 0) ? 0ull : (~0ull >> (64 - lane))))] = p;
}

// Nodes whose leaf run is short are computed directly from their leaves (min / max are
// exact, so the result does not depend on the order).  Every other node starts its counter
// at the number of children that are already complete; nodes with both children complete
// form the first list of the bottom-up rounds.

// Block-wide appends buffered in LDS: the ids a workgroup appends over its grid-stride loop are
// compacted into an LDS buffer and flushed to the global list with one atomic per flush.  One
// device-scope atomic on a single address costs about 11 ns serialised (measured: 62500
// workgroups with one atomicAdd each on one counter take 0.71 ms, spread over 64 cache lines
// 0.02 ms), so an atomic per workgroup and iteration would dominate the refit.
constexpr int kAppendCap = 2048;
struct BlockAppender {
    int32_t* buf;        // LDS, kAppendCap ids
    uint32_t* n;         // LDS, ids held
    uint32_t* wave_cnt;  // LDS, kBlock / 64 counts
    uint32_t* base;      // LDS, flush destination
};

__device__ inline void appender_flush(BlockAppender& A, int32_t* __restrict__ list, uint32_t* __restrict__ count) {
    __syncthreads();
    const uint32_t held = *A.n;
    if (held == 0) return;
    if (threadIdx.x == 0) *A.base = atomicAdd(count, held);
    __syncthreads();
    const uint32_t b = *A.base;
    for (uint32_t k = threadIdx.x; k < held; k += kBlock) list[b + k] = A.buf[k];
    __syncthreads();
    if (threadIdx.x == 0) *A.n = 0;
    __syncthreads();
}

// Call from every thread of the workgroup.
__device__ inline void appender_push(BlockAppender& A, bool take, int32_t p, int32_t* __restrict__ list,
                                     uint32_t* __restrict__ count) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint64_t m = __ballot(take);
    if (lane == 0) A.wave_cnt[wave] = (uint32_t)__popcll(m);
    __syncthreads();
    uint32_t off = *A.n, tot = 0;
    for (int w = 0; w < kBlock / 64; ++w) {
        const uint32_t c = A.wave_cnt[w];
        if (w < wave) off += c;
        tot += c;
    }
    if (take) A.buf[off + __popcll(m & ((lane == 0) ? 0ull : (~0ull >> (64 - lane))))] = p;
    __syncthreads();
    if (threadIdx.x == 0) *A.n += tot;
    __syncthreads();
    if (*A.n > (uint32_t)(kAppendCap - kBlock)) appender_flush(A, list, count);
}

#define PMX_BLOCK_APPENDER(A)                                 \
    __shared__ int32_t A##_buf[kAppendCap];                   \
    __shared__ uint32_t A##_n, A##_wc[kBlock / 64], A##_base; \
    if (threadIdx.x == 0) A##_n = 0;                          \
    __syncthreads();                                          \
    BlockAppender A{A##_buf, &A##_n, A##_wc, &A##_base}

constexpr int kRefitPad = 128;   // largest leaf range of a node refit directly (staged window margin)

// min_run > 0 (FP32 tree only): a node whose parent covers at most min_run leaves gets no
// record.  Every FP32 walk hands over a node of at most min_run leaves as one leaf run, so it
// never reads the children of such a node; the rounds above read only children of nodes larger
// than small_range >= min_run.
// F32: FP32 input and FP32 tree only; the leaf window is the exact FP32 leaves (16 bytes each) and
// the boxes are FP32 min / max, which equal the outward rounding of the FP64 min / max.
template <bool F32>
__global__ void __launch_bounds__(kBlock) k_refit_small(const double* __restrict__ cols64,
                                                        const float4* __restrict__ leaf32, const int32_t* __restrict__ lr,
                                                        const int32_t* __restrict__ escape,
                                                        const uint8_t* __restrict__ unread_in,
                                                        const int32_t* __restrict__ range_far, int64_t n,
                                                        uint32_t* __restrict__ visited, Node64* nodes,
                                                        Node32* __restrict__ n32, int small_range, int min_run,
                                                        int32_t* __restrict__ next, uint32_t* __restrict__ next_count) {
    // Internal node i covers leaves with one end at i - 1 or i, so the block's nodes read a
    // window of leaves around its own ids: staged once in LDS, then read by every node.
    constexpr int kPad = kRefitPad;
    constexpr int kWin = kBlock + 2 * kPad + 1;      // leaves i0 - 1 - kPad .. i0 + kBlock + kPad - 1
    using WinT = typename std::conditional<F32, float4, Leaf64>::type;
    __shared__ WinT win[kWin];
    PMX_BLOCK_APPENDER(app);
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    const int64_t m = n - 1;                            // internal nodes 1 .. n-1
    const int64_t iters = (m + stride - 1) / stride;    // uniform trip count keeps ballots convergent
    for (int64_t it = 0; it < iters; ++it) {
        const int64_t i0 = 1 + it * stride + (int64_t)blockIdx.x * kBlock;
        const int64_t i = i0 + threadIdx.x;
        // the node's own entries first (independent loads, in flight during the staging)
        int32_t lo = 0, hi = -1, lch = 0, rch = 0, esc = 0;
        uint8_t un = 0;
        if (i < n) {
            node_range(range_far, i, lo, hi);
            lch = lr[2 * i];
            rch = lr[2 * i + 1];
            esc = escape[i];
            if (min_run > 0) un = unread_in[i];
        }
        const int64_t w0 = i0 - 1 - kPad > 0 ? i0 - 1 - kPad : 0;
        const int64_t w1 = i0 + kBlock + kPad < n ? i0 + kBlock + kPad : n;
        __syncthreads();
        for (int64_t k = w0 + threadIdx.x; k < w1; k += kBlock) {
            if constexpr (F32) win[k - w0] = leaf32[k];
            else win[k - w0] = read_leaf(cols64, leaf32, n, k);
        }
        __syncthreads();
        bool take = false;
        if (i < n) {
            if (min_run > 0 && un) {
                // no record: no FP32 walk reaches this node
            } else if (lo >= 1 && hi <= n && hi - lo + 1 <= small_range) {
                // small_range <= kPad: the leaf range lies inside the staged window
                if constexpr (F32) {
                    float4 a = win[lo - 1 - w0];
                    float mn0 = a.x, mn1 = a.y, mn2 = a.z, mx0 = a.x, mx1 = a.y, mx2 = a.z, hm = a.w;
                    for (int32_t k = lo; k < hi; ++k) {
                        const float4 L = win[k - w0];
                        mn0 = fminf(mn0, L.x); mx0 = fmaxf(mx0, L.x);
                        mn1 = fminf(mn1, L.y); mx1 = fmaxf(mx1, L.y);
                        mn2 = fminf(mn2, L.z); mx2 = fmaxf(mx2, L.z);
                        hm = fmaxf(hm, L.w);
                    }
                    Node32 b;
                    b.mnx = mn0; b.mny = mn1; b.mnz = mn2; b.hmax = hm;
                    b.mxx = mx0; b.mxy = mx1; b.mxz = mx2;
                    b.left = lch;
                    b.escape = esc;
                    b.lo = lo - 1;
                    b.cnt = hi - lo + 1;
                    b.pad = 0;
                    n32[i] = b;
                } else {
                    double mn[3] = {__builtin_inf(), __builtin_inf(), __builtin_inf()};
                    double mx[3] = {-__builtin_inf(), -__builtin_inf(), -__builtin_inf()};
                    double hm = -__builtin_inf();
                    for (int32_t k = lo - 1; k < hi; ++k) {
                        const Leaf64 L = win[k - w0];
                        mn[0] = L.x < mn[0] ? L.x : mn[0]; mx[0] = L.x > mx[0] ? L.x : mx[0];
                        mn[1] = L.y < mn[1] ? L.y : mn[1]; mx[1] = L.y > mx[1] ? L.y : mx[1];
                        mn[2] = L.z < mn[2] ? L.z : mn[2]; mx[2] = L.z > mx[2] ? L.z : mx[2];
                        hm = L.h > hm ? L.h : hm;
                    }
                    Node64 out;
                    for (int k = 0; k < 3; ++k) { out.mn[k] = mn[k]; out.mx[k] = mx[k]; }
                    out.hmax = hm;
                    out.left = lch;
                    out.escape = esc;
                    store_node(out, i, range_far, nodes, n32);
                }
            } else {
                uint32_t done = 0;
                const int32_t chs[2] = {lch, rch};
                for (int side = 0; side < 2; ++side) {
                    const int32_t ch = chs[side];
                    if (ch >= n) done += 1;
                    else if (ch > 0 && abs(range_far[ch] - ch) + 1 <= small_range) done += 1;
                }
                visited[i] = done;
                take = done == 2;
            }
        }
        appender_push(app, take, (int32_t)i, next, next_count);
    }
    appender_flush(app, next, next_count);
}

// One bottom-up round: complete every node of the list from its two complete children, then
// count it at its parent; a parent counted twice is complete-able in the next round.
struct RefitTree {
    const double* cols64;       // x | y | z | m | h ... in Morton order (FP64 input)
    const float4* leaf32;       // x y z h (FP32 input: exact; null for an FP64 input)
    const int32_t* lr;
    const int32_t* parent;
    const int32_t* escape;
    const int32_t* far;         // other end of each internal node's leaf range
    int64_t n;
    uint32_t* visited;
    Node64* nodes;
    Node32* n32;
};

template <bool BLOCK>
__device__ inline void refit_step(const RefitTree& R, int32_t x, bool valid, int32_t* __restrict__ next,
                                  uint32_t* __restrict__ next_count, BlockAppender* app = nullptr) {
    const int64_t n = R.n;
    bool take = false;
    int32_t p = 0;
    if (valid && x > 0 && x < n) {
        combine_node(R.cols64, R.leaf32, R.lr, R.escape, n, R.nodes, x, R.far, R.n32);
        const int32_t* parent = R.parent;
        uint32_t* visited = R.visited;
        p = parent[x];
        if (p > 0 && p < n) take = atomicAdd(&visited[p], 1u) == 1u;
    }
    if (BLOCK) appender_push(*app, take, p, next, next_count);
    else wave_append(take, p, next, next_count);
}

__global__ void __launch_bounds__(kBlock) k_refit_round(RefitTree R, const int32_t* __restrict__ cur,
                                                        const uint32_t* __restrict__ cur_count,
                                                        int32_t* __restrict__ next, uint32_t* __restrict__ next_count) {
    PMX_BLOCK_APPENDER(app);
    const int64_t count = *cur_count;
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    const int64_t iters = (count + stride - 1) / stride;
    for (int64_t it = 0; it < iters; ++it) {
        const int64_t k = it * stride + (int64_t)blockIdx.x * kBlock + threadIdx.x;
        refit_step<true>(R, k < count ? cur[k] : 0, k < count, next, next_count, &app);
    }
    appender_flush(app, next, next_count);
}

// The remaining rounds (short lists near the root) in one workgroup: rounds are separated by
// workgroup barriers instead of kernel launches.  Ends when a round appends nothing.
constexpr int kTailBlock = 1024;
constexpr int kMaxRounds = 4000;

__global__ void __launch_bounds__(kTailBlock) k_refit_tail(RefitTree R, int32_t* listA, int32_t* listB,
                                                           uint32_t* counts, int first_round) {
    for (int r = first_round; r < kMaxRounds - 1; ++r) {
        __syncthreads();
        const int64_t count = __hip_atomic_load(&counts[r], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (count == 0) break;
        const int32_t* cur = (r & 1) ? listB : listA;
        int32_t* nxt = (r & 1) ? listA : listB;
        const int64_t iters = (count + kTailBlock - 1) / kTailBlock;
        for (int64_t it = 0; it < iters; ++it) {
            const int64_t k = it * kTailBlock + threadIdx.x;
            refit_step<false>(R, k < count ? cur[k] : 0, k < count, nxt, &counts[r + 1]);
        }
    }
}

// Root (internal node 1) of the FP64 tree before its refit: the exact box is the particle
// bounding box; hmax and the links are those of the FP32 root (hmax rounded up, replaced by
// the FP64 refit before any FP64 traversal reads it).
__global__ void k_root_box(const double* __restrict__ box, const Node32* __restrict__ n32, Node64* __restrict__ nodes) {
    Node64 r;
    for (int k = 0; k < 3; ++k) { r.mn[k] = box[k]; r.mx[k] = box[3 + k]; }
    r.hmax = n32[1].hmax;
    r.left = n32[1].left;
    r.escape = n32[1].escape;
    nodes[1] = r;
}

template <typename T>
int launch_front(pmx_ctx* c, const ColPtrs& cp, int ncols, int64_t n, double* part, int nparts, int* nan_count,
                 double* box, uint64_t* codes, uint32_t* hist, T* rec) {
    hipLaunchKernelGGL(k_bbox_partial<T>, dim3(nparts), dim3(kBlock), 0, c->stream, (const T*)cp.p[0],
                       (const T*)cp.p[1], (const T*)cp.p[2], n, part, nan_count);
    hipLaunchKernelGGL(k_bbox_final, dim3(1), dim3(kBlock), 0, c->stream, part, nparts, box);
    const int ntiles = (int)((n + kTile - 1) / kTile);
    const dim3 g(ntiles), b(kBlock);
    hipStream_t s = c->stream;
    switch (ncols) {
    case 6: hipLaunchKernelGGL((k_encode<T, 6>), g, b, 0, s, cp, ncols, n, box, ntiles, codes, hist, rec); break;
    case 7: hipLaunchKernelGGL((k_encode<T, 7>), g, b, 0, s, cp, ncols, n, box, ntiles, codes, hist, rec); break;
    case 8: hipLaunchKernelGGL((k_encode<T, 8>), g, b, 0, s, cp, ncols, n, box, ntiles, codes, hist, rec); break;
    case 9: hipLaunchKernelGGL((k_encode<T, 9>), g, b, 0, s, cp, ncols, n, box, ntiles, codes, hist, rec); break;
    case 10: hipLaunchKernelGGL((k_encode<T, 10>), g, b, 0, s, cp, ncols, n, box, ntiles, codes, hist, rec); break;
    default: hipLaunchKernelGGL((k_encode<T, 0>), g, b, 0, s, cp, ncols, n, box, ntiles, codes, hist, rec); break;
    }
    return PMX_OK;
}

}  // namespace

int radix_sort_pairs(pmx_ctx* c, uint64_t* keys, int32_t* vals, int64_t n, uint64_t* keys_alt,
                     int32_t* vals_alt, int begin_bit, int end_bit, bool* result_in_alt, uint32_t* hist,
                     bool first_hist_ready, bool vals_are_indices) {
    static_assert(kTile == 4096 && kBins == 256, "radix_hist_bytes assumes 4096-key tiles of 8-bit digits");
    const int ntiles = (int)((n + kTile - 1) / kTile);
    if (!hist) {
        PMX_TRY(ensure(c, c->hist, radix_hist_bytes(n)));
        hist = c->hist.as<uint32_t>();
    }
    uint32_t* totals = hist + (size_t)kBins * ntiles;
    uint64_t *kin = keys, *kout = keys_alt;
    int32_t *vin = vals, *vout = vals_alt;
    bool alt = false;
    for (int shift = begin_bit; shift < end_bit; shift += kRadixBits) {
        if (!(first_hist_ready && shift == begin_bit))
            hipLaunchKernelGGL(k_radix_hist, dim3(ntiles), dim3(kBlock), 0, c->stream, kin, n, shift, ntiles, hist);
        hipLaunchKernelGGL(k_radix_scan, dim3(kBins), dim3(kBlock), 0, c->stream, hist, ntiles, totals);
        if (vals_are_indices && shift == begin_bit)
            hipLaunchKernelGGL(k_radix_scatter<true>, dim3(ntiles), dim3(kBlock), 0, c->stream, kin, vin, n, shift,
                               ntiles, hist, totals, kout, vout);
        else
            hipLaunchKernelGGL(k_radix_scatter<false>, dim3(ntiles), dim3(kBlock), 0, c->stream, kin, vin, n, shift,
                               ntiles, hist, totals, kout, vout);
        std::swap(kin, kout);
        std::swap(vin, vout);
        alt = !alt;
    }
    PMX_HIP(c, hipGetLastError());
    *result_in_alt = alt;
    return PMX_OK;
}

// Refit: short-range nodes directly, then kLaunchRounds grid-wide rounds (the long lists
// near the leaves), then the tail kernel for the rest; no host synchronisation.

// fp64: also write the FP64 nodes (otherwise only the FP32 tree of the compensated path).
int refit_rounds(pmx_ctx* c, int64_t n, unsigned g, bool fp64) {
    hipStream_t s = c->stream;
    constexpr int kLaunchRounds = 12;
    PMX_TRY(ensure(c, c->scratch_c, 4 * (size_t)kMaxRounds + 64));
    int32_t* listA = c->keys_tmp.as<int32_t>();              // free after the sort (8n bytes)
    int32_t* listB = listA + (n / 2 + 64);
    uint32_t* counts = reinterpret_cast<uint32_t*>(c->scratch_c.p);
    hipLaunchKernelGGL(k_fill_u32, dim3(16), dim3(kBlock), 0, s, counts, kMaxRounds, 0u);
    RefitTree R;
    R.cols64 = c->tree.f32in ? nullptr : c->cols64.as<double>();
    R.leaf32 = c->tree.f32in ? c->leaf32.as<float4>() : nullptr;
    R.lr = c->lr.as<int32_t>();
    R.parent = c->parent.as<int32_t>();
    R.escape = c->escape.as<int32_t>();
    R.far = c->range_far.as<int32_t>();
    R.n = n;
    R.visited = c->visited.as<uint32_t>();
    R.nodes = fp64 ? c->nodes64.as<Node64>() : nullptr;
    R.n32 = c->nodes32.as<Node32>();
    const unsigned gs = grid_for(n, kBlock, 8 * (int64_t)c->ncu);   // grid-stride, appends buffered per workgroup
    const int small = std::max(kMinRun, std::min(c->tune.refit_small, kRefitPad));
    if (!fp64 && R.leaf32)
        hipLaunchKernelGGL(k_refit_small<true>, dim3(gs), dim3(kBlock), 0, s, R.cols64, R.leaf32, R.lr, R.escape,
                           c->unread.as<uint8_t>(), R.far, n, R.visited, R.nodes, R.n32, small, kMinRun, listA,
                           &counts[0]);
    else
        hipLaunchKernelGGL(k_refit_small<false>, dim3(gs), dim3(kBlock), 0, s, R.cols64, R.leaf32, R.lr, R.escape,
                           c->unread.as<uint8_t>(), R.far, n, R.visited, R.nodes, R.n32, small,
                           fp64 ? 0 : kMinRun, listA, &counts[0]);
    const unsigned grr = std::max(1u, std::min(g / 4, 1024u));
    for (int round = 0; round < kLaunchRounds; ++round) {
        int32_t* cur = (round & 1) ? listB : listA;
        int32_t* nxt = (round & 1) ? listA : listB;
        hipLaunchKernelGGL(k_refit_round, dim3(grr), dim3(kBlock), 0, s, R, cur, &counts[round], nxt,
                           &counts[round + 1]);
    }
    hipLaunchKernelGGL(k_refit_tail, dim3(1), dim3(kTailBlock), 0, s, R, listA, listB, counts, kLaunchRounds);
    PMX_HIP(c, hipGetLastError());
    return PMX_OK;
}

int build_lbvh(pmx_ctx* c, const pmx_particles* p, int flags, uint64_t* codes_out, int64_t* order_out) {
    if (!p) return set_error(c, PMX_ERR_ARGUMENT, "particles: null pointer");
    const int64_t n = p->n;
    if (n < 1)
        return set_error(c, PMX_ERR_ARGUMENT, "BinaryRadixTree: enc.codes must be non-empty (got n=0).");
    if (n > 1073741824LL)
        return set_error(c, PMX_ERR_ARGUMENT, "BinaryRadixTree: n exceeds the Int32 node-id capacity (n <= 1073741824).");
    if (p->dtype != PMX_F32 && p->dtype != PMX_F64) return set_error(c, PMX_ERR_ARGUMENT, "particles: dtype");
    if (p->ncol < 0 || p->ncol > 18) return set_error(c, PMX_ERR_ARGUMENT, "particles: ncol must be 0..18");
    if (p->kernel < 0 || p->kernel > 5) return set_error(c, PMX_ERR_ARGUMENT, "particles: unknown kernel");
    if (!p->x || !p->y || !p->z || !p->m || !p->h || !p->rho || (p->ncol > 0 && !p->quant))
        return set_error(c, PMX_ERR_ARGUMENT, "particles: null column pointer");
    for (int j = 0; j < p->ncol; ++j)
        if (!p->quant[j]) return set_error(c, PMX_ERR_ARGUMENT, "particles: null quant column");
    const int ncols = 6 + p->ncol;
    const size_t esz = p->dtype == PMX_F64 ? 8 : 4;
    hipStream_t s = c->stream;
    c->tree.built = false;
    // kBuildAsync: no events, no device-to-host copies, no host synchronisation (the build can
    // be captured into a graph; the NaN count stays on the device, see nan_count_device)
    const bool async = (flags & kBuildAsync) != 0;
    if (async && (p->mem == PMX_MEM_HOST || codes_out || order_out || (flags & PMX_BUILD_WRITEBACK)))
        return set_error(c, PMX_ERR_ARGUMENT, "asynchronous build needs device columns and no host outputs");

    if (!async) PMX_HIP(c, hipEventRecord(c->ev[0], s));
    // ---- inputs on the device (host columns are staged) ----
    const void* src[24];
    const void* in[24] = {p->x, p->y, p->z, p->m, p->h, p->rho};
    for (int j = 0; j < p->ncol; ++j) in[6 + j] = p->quant[j];
    if (p->mem == PMX_MEM_HOST) {
        PMX_TRY(ensure(c, c->staging, esz * (size_t)n * ncols));
        for (int j = 0; j < ncols; ++j) {
            void* dst = (char*)c->staging.p + esz * (size_t)n * j;
            PMX_HIP(c, hipMemcpyAsync(dst, in[j], esz * (size_t)n, hipMemcpyHostToDevice, s));
            src[j] = dst;
        }
    } else {
        for (int j = 0; j < ncols; ++j) src[j] = in[j];
    }
    // ---- bbox + encode ----
    const int nparts = (int)std::min<int64_t>(1024, (n + kBlock - 1) / kBlock);
    PMX_TRY(ensure(c, c->misc, sizeof(double) * (6 * 1024 + 8) + 64));
    double* part = c->misc.as<double>();
    double* box = part + 6 * 1024;
    int* nan_count = reinterpret_cast<int*>(box + 6);
    PMX_TRY(ensure(c, c->codes, 8 * (size_t)n + 2048));
    PMX_TRY(ensure(c, c->keys_tmp, 8 * (size_t)n + 2048));   // also holds the refit lists (n + 128 ints)
    PMX_TRY(ensure(c, c->order, 4 * (size_t)n));
    PMX_TRY(ensure(c, c->vals_tmp, 4 * (size_t)n));
    uint64_t* codes = c->codes.as<uint64_t>();
    int32_t* order = c->order.as<int32_t>();
    PMX_TRY(ensure(c, c->hist, radix_hist_bytes(n)));
    uint32_t* hist = c->hist.as<uint32_t>();
    PMX_TRY(ensure(c, c->records, esz * (size_t)n * ncols));
    ColPtrs cp;
    for (int j = 0; j < ncols; ++j) cp.p[j] = src[j];
    int* dup = nan_count + 1;     // equal neighbouring codes after the sort (de-dup needed)
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(kBlock), 0, s, (uint32_t*)nan_count, 2, 0u);
    if (p->dtype == PMX_F64)
        launch_front<double>(c, cp, ncols, n, part, nparts, nan_count, box, codes, hist, c->records.as<double>());
    else
        launch_front<float>(c, cp, ncols, n, part, nparts, nan_count, box, codes, hist, c->records.as<float>());
    // The encode is the last kernel that reads the caller's columns: the NaN check waits for it,
    // and the rest of the build runs on after this call returns (every later call on the
    // context is ordered after it on the context's stream; host outputs wait for it below).
    if (!async) {
        int nan_host = 0;
        PMX_HIP(c, hipMemcpyAsync(&nan_host, nan_count, sizeof(int), hipMemcpyDeviceToHost, s));
        PMX_HIP(c, hipStreamSynchronize(s));
        if (nan_host > 0)
            return set_error(c, PMX_ERR_ARGUMENT, "InexactError: NaN particle coordinate (UInt64(floor(NaN)))");
    }
    // ---- stable sort + de-dup (the encode counted the first digit) ----
    bool in_alt = false;
    PMX_TRY(radix_sort_pairs(c, codes, order, n, c->keys_tmp.as<uint64_t>(), c->vals_tmp.as<int32_t>(), 0, 64,
                             &in_alt, hist, true, true));
    if (in_alt) {
        std::swap(c->codes, c->keys_tmp);
        std::swap(c->order, c->vals_tmp);
        codes = c->codes.as<uint64_t>();
        order = c->order.as<int32_t>();
    }
    const int ntiles = (int)((n + kTile - 1) / kTile);
    int64_t* tmax = reinterpret_cast<int64_t*>(c->keys_tmp.p);   // free after the sort
    hipLaunchKernelGGL(k_dedup_tilemax, dim3(ntiles), dim3(kBlock), 0, s, codes, n, tmax, dup);
    hipLaunchKernelGGL(k_dedup_scan, dim3(1), dim3(kBlock), 0, s, tmax, ntiles, dup);
    hipLaunchKernelGGL(k_dedup_apply, dim3(ntiles), dim3(kBlock), 0, s, codes, n, tmax, dup);
    if (!async) PMX_HIP(c, hipEventRecord(c->ev[1], s));
    // ---- columns into Morton order ----
    const bool f32in = p->dtype == PMX_F32;
    if (!f32in) PMX_TRY(ensure(c, c->cols64, 8 * (size_t)n * ncols));
    double* cols64 = f32in ? nullptr : c->cols64.as<double>();
    const unsigned g = grid_for(n, kBlock, 16384);
    // FP32 data of the compensated path (layout shared with tree32 / prepare_f32)
    const size_t payf = (size_t)n * (2 + p->ncol);
    PMX_TRY(ensure(c, c->nodes32, sizeof(Node32) * (size_t)(n + 1)));
    PMX_TRY(ensure(c, c->leaf32, 2 * sizeof(float4) * (size_t)n));
    PMX_TRY(ensure(c, c->pay32, 4 * (payf + (payf & 1)) + 16 * (size_t)n));
    Leaf32Out o32;
    o32.leaf = c->leaf32.as<float4>();
    o32.leaflo = c->leaf32.as<float4>() + n;
    o32.pay = c->pay32.as<float>();
    o32.w64 = reinterpret_cast<double*>(c->pay32.as<float>() + payf + (payf & 1));
    o32.ih64 = o32.w64 + n;
    if (p->dtype == PMX_F64)
        launch_gather<double>(c, c->records.as<double>(), ncols, order, n, cols64, o32, g);
    else
        launch_gather<float>(c, c->records.as<float>(), ncols, order, n, cols64, o32, g);
    c->tree.f32in = f32in;
    c->tree.cols64_ready = !f32in;
    if (!async) PMX_HIP(c, hipEventRecord(c->ev[2], s));
    // ---- radix tree, escapes, refit ----
    PMX_TRY(ensure(c, c->lr, 8 * (size_t)(2 * n + 2)));
    PMX_TRY(ensure(c, c->parent, 4 * (size_t)(2 * n + 2)));
    PMX_TRY(ensure(c, c->escape, 4 * (size_t)(2 * n + 2)));
    PMX_TRY(ensure(c, c->range_far, 4 * (size_t)(n + 2)));
    PMX_TRY(ensure(c, c->visited, 4 * (size_t)(n + 2)));
    PMX_TRY(ensure(c, c->unread, (size_t)(n + 2)));
    PMX_TRY(ensure(c, c->leaf_esc, 4 * (size_t)n));
    PMX_TRY(ensure(c, c->nodes64, sizeof(Node64) * (size_t)(n + 1)));
    int32_t* lr = c->lr.as<int32_t>();
    int32_t* parent = c->parent.as<int32_t>();
    int32_t* escape = c->escape.as<int32_t>();
    // No fills: k_brt writes the links, parents, ranges and escapes of every node and the leaves'
    // (zero) links, k_refit_small the arrival counter of every node the bottom-up rounds count at
    hipLaunchKernelGGL(k_tree_heads, dim3(1), dim3(64), 0, s, lr, parent, escape);
    if (n >= 2) {
        hipLaunchKernelGGL(k_brt, dim3(grid_for(n, kBlock, 1 << 20)), dim3(kBlock), 0, s, codes, n, lr, parent,
                           c->range_far.as<int32_t>(), escape, c->leaf_esc.as<int32_t>(),
                           c->unread.as<uint8_t>(), kMinRun);
    } else {
        hipLaunchKernelGGL(k_escape_leaf, dim3(g), dim3(kBlock), 0, s, lr, parent, n, escape, c->leaf_esc.as<int32_t>());
    }
    if (n >= 2) {
        // the FP32 tree only; the FP64 nodes follow on first use (ensure_leaf64), except the
        // root's box, which is the particle bounding box of the Morton normalisation
        PMX_TRY(refit_rounds(c, n, g, false));
        hipLaunchKernelGGL(k_root_box, dim3(1), dim3(1), 0, s, box, c->nodes32.as<Node32>(), c->nodes64.as<Node64>());
    }
    PMX_HIP(c, hipGetLastError());
    c->tree.n = n;
    c->tree.leaf64_ready = false;
    c->tree.root = n >= 2 ? 1 : 0;
    c->tree.ncol = p->ncol;
    c->tree.kernel = p->kernel;
    if (async) {
        c->tree.built = true;
        c->tree.f32_ready = true;
        return PMX_OK;
    }
    PMX_HIP(c, hipEventRecord(c->ev[3], s));
    c->build_timings_pending = true;     // read from the events by the first pmx_last_timings
    // ---- host outputs (each waits for the build) ----
    if (codes_out) {
        PMX_HIP(c, hipMemcpyAsync(codes_out, codes, 8 * (size_t)n, hipMemcpyDeviceToHost, s));
        PMX_HIP(c, hipStreamSynchronize(s));
    }
    if (order_out) {
        PMX_HIP(c, hipStreamSynchronize(s));
        std::vector<int32_t> tmp((size_t)n);
        PMX_HIP(c, hipMemcpy(tmp.data(), order, 4 * (size_t)n, hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < n; ++i) order_out[i] = tmp[(size_t)i];
    }
    if ((flags & PMX_BUILD_WRITEBACK) && p->mem == PMX_MEM_DEVICE) {
        ColOut dst;
        for (int j = 0; j < ncols; ++j) dst.p[j] = const_cast<void*>(in[j]);
        if (p->dtype == PMX_F64)
            hipLaunchKernelGGL(k_writeback<double>, dim3(g), dim3(kBlock), 0, s, c->records.as<double>(), ncols,
                               order, n, dst);
        else
            hipLaunchKernelGGL(k_writeback<float>, dim3(g), dim3(kBlock), 0, s, c->records.as<float>(), ncols,
                               order, n, dst);
        PMX_HIP(c, hipGetLastError());
        PMX_HIP(c, hipStreamSynchronize(s));
    }
    if ((flags & PMX_BUILD_WRITEBACK) && p->mem == PMX_MEM_HOST) {
        PMX_TRY(ensure_cols64(c));
        cols64 = c->cols64.as<double>();
        PMX_HIP(c, hipStreamSynchronize(s));
        void* outp[24] = {p->x, p->y, p->z, p->m, p->h, p->rho};
        for (int j = 0; j < p->ncol; ++j) outp[6 + j] = p->quant[j];
        if (p->dtype == PMX_F64) {
            for (int j = 0; j < ncols; ++j)
                PMX_HIP(c, hipMemcpy(outp[j], cols64 + (size_t)j * n, 8 * (size_t)n, hipMemcpyDeviceToHost));
        } else {
            std::vector<double> tmp((size_t)n);
            for (int j = 0; j < ncols; ++j) {
                PMX_HIP(c, hipMemcpy(tmp.data(), cols64 + (size_t)j * n, 8 * (size_t)n, hipMemcpyDeviceToHost));
                float* dst = static_cast<float*>(outp[j]);
                for (int64_t i = 0; i < n; ++i) dst[i] = (float)tmp[(size_t)i];
            }
        }
    }
    c->tree.n = n;
    c->tree.leaf64_ready = false;
    c->tree.root = n >= 2 ? 1 : 0;
    c->tree.ncol = p->ncol;
    c->tree.kernel = p->kernel;
    c->tree.built = true;
    c->tree.f32_ready = true;    // the refit and the record gather also wrote the FP32 tree
    return PMX_OK;
}

// Phase times of the last synchronous build, from its events (waits for the build to finish).
int settle_build_timings(pmx_ctx* c) {
    if (!c->build_timings_pending) return PMX_OK;
    PMX_HIP(c, hipEventSynchronize(c->ev[3]));
    float ms[3];
    PMX_HIP(c, hipEventElapsedTime(&ms[0], c->ev[0], c->ev[1]));
    PMX_HIP(c, hipEventElapsedTime(&ms[1], c->ev[1], c->ev[2]));
    PMX_HIP(c, hipEventElapsedTime(&ms[2], c->ev[2], c->ev[3]));
    c->timings[0] = ms[0] + ms[1] + ms[2];
    c->timings[1] = ms[0];
    c->timings[2] = ms[2];
    c->timings[3] = ms[1];
    c->build_timings_pending = false;
    return PMX_OK;
}

}  // namespace pmx

namespace pmx {
// Device address of the NaN-coordinate counter of the context's last build.
// The plain FP64 path's structures, built on first use after a build: Leaf64 records and the
// FP64 node boxes (a second bottom-up refit; the build refits the FP32 tree only).
int ensure_cols64(pmx_ctx* c) {
    if (!c->tree.f32in || c->tree.cols64_ready) return PMX_OK;
    const int64_t n = c->tree.n;
    const int ncols = 6 + c->tree.ncol;
    PMX_TRY(ensure(c, c->cols64, 8 * (size_t)n * ncols));
    hipLaunchKernelGGL(k_gather_cols64, dim3(grid_for(n, kBlock, 16384)), dim3(kBlock), 0, c->stream,
                       c->records.as<float>(), ncols, c->order.as<int32_t>(), n, c->cols64.as<double>());
    PMX_HIP(c, hipGetLastError());
    c->tree.cols64_ready = true;
    return PMX_OK;
}

int ensure_leaf64(pmx_ctx* c) {
    if (c->tree.leaf64_ready) return PMX_OK;
    PMX_TRY(ensure_cols64(c));
    const int64_t n = c->tree.n;
    if (n >= 2) {
        const unsigned g = grid_for(n, kBlock, 16384);
        PMX_TRY(refit_rounds(c, n, g, true));
    }
    PMX_TRY(ensure(c, c->leaf64, sizeof(Leaf64) * (size_t)std::max<int64_t>(n, 1)));
    hipLaunchKernelGGL(k_pack_leaf, dim3(grid_for(n, kBlock, 16384)), dim3(kBlock), 0, c->stream,
                       c->cols64.as<double>(), n, c->leaf64.as<Leaf64>());
    PMX_HIP(c, hipGetLastError());
    c->tree.leaf64_ready = true;
    return PMX_OK;
}

int* nan_count_device(pmx_ctx* c) { return reinterpret_cast<int*>(c->misc.as<double>() + 6 * 1024 + 6); }
}  // namespace pmx
