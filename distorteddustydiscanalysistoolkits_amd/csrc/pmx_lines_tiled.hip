// pmx_lines_tiled.hip — line maps of parallel lines (face-on column maps) on gfx950.
//
// LineSamples_interpolation (line_integrated_scalar_interpolation.jl:123-166) accepts a
// (line, particle) pair when the reference's FP64 perpendicular distance^2
// _squared_distance_line_coords (|Delta|^2 - (Delta.d)^2) is <= (K h_b)^2 (itpScatter).  When
// every line has the same unit direction d, a line is a point (s, t) in the plane normal to d,
// and a particle's accepted lines are the points inside its disc of radius K h_b there.  This
// path works particle-centric:
//   * lines are binned into a grid of cells in the (s, t) plane; 8 x 8 cells form a tile, and
//     lines are sorted by (tile, cell) so a tile's lines are one contiguous run;
//   * one wavefront per tile (8 per workgroup, sharing an LDS copy of the Inorm table) walks
//     the LBVH once (wave-uniform node test: the node box projected on the plane against the
//     tile rectangle, radius K hmax), and each bucket of
//     up to 64 particles (contiguous in Morton order) is taken one particle per lane;
//   * each lane visits only the cells its disc covers, tests those lines in FP64 on the
//     projected coordinates with a guard band, re-checks band cases with the reference's
//     FP64 formula on the FP64 columns, and adds m/rho * W_line and the field terms to FP64
//     accumulators of the tile's lines in LDS (ds_add_f64);
//   * the wave then applies the Shepard division and writes the tile's outputs.
// The accepted pair set is the reference's FP64 predicate set; only the order of the FP64
// additions into one line differs from the per-line traversal (not bit-reproducible across
// runs, agreement ~1e-15 relative).
// Nearly parallel sets (unit directions within kTiltMinCos of their mean direction d, such as a
// perspective camera's rays) take the same kernels with TILT: lines are binned where they cross a
// reference plane normal to d through the particles' mid-depth, and a particle is placed on that
// plane per tile through the tile's range of slopes (lmap::lines_tile).  Sets whose slopes spread
// too much inside a tile (more than kTiltMaxSpread cells of image width on average) and lines
// that are not unit-length use the per-tile union walk of pmx_interp32.hip instead.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <vector>

#include "pmx_internal.h"
#include "pmx_tree32.h"
#include "pmx_lines_tile.h"

namespace pmx {

namespace {

using namespace lmap;


// (s, t) of each line origin, parallel check against line 0, per-block bounds.
__global__ void __launch_bounds__(256) k_line_project(const double* __restrict__ o, const double* __restrict__ d,
                                                      int64_t nl, Frame f, double* __restrict__ st,
                                                      double* __restrict__ part, int* __restrict__ flag) {
    __shared__ double red[5][256];
    double smin = __builtin_inf(), smax = -__builtin_inf(), tmin = __builtin_inf(), tmax = -__builtin_inf(),
           amax = 0.0;
    bool bad = false;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < nl; i += (int64_t)gridDim.x * 256) {
        const double ox = o[i], oy = o[nl + i], oz = o[2 * nl + i];
        bad |= !(d[i] == f.d[0] && d[nl + i] == f.d[1] && d[2 * nl + i] == f.d[2]);
        const double s = (ox * f.u[0] + oy * f.u[1]) + oz * f.u[2];
        const double t = (ox * f.v[0] + oy * f.v[1]) + oz * f.v[2];
        bad |= !(fabs(s) < 1e300 && fabs(t) < 1e300);
        st[i] = s;
        st[nl + i] = t;
        smin = fmin(smin, s); smax = fmax(smax, s);
        tmin = fmin(tmin, t); tmax = fmax(tmax, t);
        amax = fmax(amax, fmax(fabs(ox), fmax(fabs(oy), fabs(oz))));
    }
    // one atomic per workgroup: a device-scope atomic on one address costs ~11 ns serialised
    if (__syncthreads_or(bad ? 1 : 0) && threadIdx.x == 0) atomicOr(flag, 1);
    red[0][threadIdx.x] = smin; red[1][threadIdx.x] = -smax;
    red[2][threadIdx.x] = tmin; red[3][threadIdx.x] = -tmax;
    red[4][threadIdx.x] = -amax;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int k = 0; k < 5; ++k) red[k][threadIdx.x] = fmin(red[k][threadIdx.x], red[k][threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x < 5) part[blockIdx.x * 5 + threadIdx.x] = red[threadIdx.x][0];
}

// Nearly parallel sets.  Sum of the directions (per-workgroup partials) and a flag for a line that
// is not unit-length or not finite.
__global__ void __launch_bounds__(256) k_line_dirsum(const double* __restrict__ d, int64_t nl,
                                                     double* __restrict__ part, int* __restrict__ flag) {
    __shared__ double red[3][256];
    double a[3] = {0.0, 0.0, 0.0};
    bool bad = false;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < nl; i += (int64_t)gridDim.x * 256) {
        const double x = d[i], y = d[nl + i], z = d[2 * nl + i];
        const double l2 = x * x + y * y + z * z;
        bad |= !(fabs(l2 - 1.0) <= 1e-12);
        a[0] += x; a[1] += y; a[2] += z;
    }
    if (__syncthreads_or(bad ? 1 : 0) && threadIdx.x == 0) atomicOr(flag, 2);
    for (int k = 0; k < 3; ++k) red[k][threadIdx.x] = a[k];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int k = 0; k < 3; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < 3) part[blockIdx.x * 3 + threadIdx.x] = red[threadIdx.x][0];
}

// Per line of a nearly parallel set: the slopes against the mean direction f.d, the direction's
// components on u and v, and the point where the line crosses the reference plane (depth f.w0),
// projected as (s, t); per-workgroup bounds as in k_line_project.  Flag 4: a line more than
// acos(kTiltMinCos) from f.d, or a non-finite value.
__global__ void __launch_bounds__(256) k_line_project_tilt(const double* __restrict__ o, const double* __restrict__ d,
                                                           int64_t nl, Frame f, double min_cos, double* __restrict__ st,
                                                           double4* __restrict__ sl, double* __restrict__ part,
                                                           int* __restrict__ flag) {
    __shared__ double red[5][256];
    double smin = __builtin_inf(), smax = -__builtin_inf(), tmin = __builtin_inf(), tmax = -__builtin_inf(),
           amax = 0.0;
    bool bad = false;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < nl; i += (int64_t)gridDim.x * 256) {
        const double ox = o[i], oy = o[nl + i], oz = o[2 * nl + i];
        const double dx = d[i], dy = d[nl + i], dz = d[2 * nl + i];
        const double c = (dx * f.d[0] + dy * f.d[1]) + dz * f.d[2];
        const double du = (dx * f.u[0] + dy * f.u[1]) + dz * f.u[2];
        const double dv = (dx * f.v[0] + dy * f.v[1]) + dz * f.v[2];
        bad |= !(c >= min_cos);
        const double cc = c >= min_cos ? c : 1.0;
        const double ss = du / cc, stl = dv / cc;
        const double ow = (ox * f.d[0] + oy * f.d[1]) + oz * f.d[2];
        const double s = fma(ss, f.w0 - ow, (ox * f.u[0] + oy * f.u[1]) + oz * f.u[2]);
        const double t = fma(stl, f.w0 - ow, (ox * f.v[0] + oy * f.v[1]) + oz * f.v[2]);
        bad |= !(fabs(s) < 1e300 && fabs(t) < 1e300);
        st[i] = s;
        st[nl + i] = t;
        sl[i] = make_double4(ss, stl, du, dv);
        smin = fmin(smin, s); smax = fmax(smax, s);
        tmin = fmin(tmin, t); tmax = fmax(tmax, t);
        amax = fmax(amax, fmax(fabs(ox), fmax(fabs(oy), fabs(oz))));
        amax = fmax(amax, fmax(fabs(s), fabs(t)));
    }
    if (__syncthreads_or(bad ? 1 : 0) && threadIdx.x == 0) atomicOr(flag, 4);
    red[0][threadIdx.x] = smin; red[1][threadIdx.x] = -smax;
    red[2][threadIdx.x] = tmin; red[3][threadIdx.x] = -tmax;
    red[4][threadIdx.x] = -amax;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int k = 0; k < 5; ++k) red[k][threadIdx.x] = fmin(red[k][threadIdx.x], red[k][threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x < 5) part[blockIdx.x * 5 + threadIdx.x] = red[threadIdx.x][0];
}

// Per tile of a nearly parallel set: bounds of its lines' slopes, 1 / the smallest d_i.d (from
// the unit direction's u and v components), 1 + the largest |slope|; and the image width the
// slope spread adds over the particles' depth range, in cells, summed per workgroup (one atomic
// each) with the count of non-empty tiles.
__global__ void __launch_bounds__(256) k_tile_slopes(const int32_t* __restrict__ cstart, const double4* __restrict__ sl,
                                                     int64_t ntiles, Frame f, double* __restrict__ tsl,
                                                     double* __restrict__ spread) {
    __shared__ double red[2][256];
    double sp = 0.0, ne = 0.0;
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t < ntiles) {
        const int a = cstart[t * kCells], b = cstart[(t + 1) * kCells];
        double q[kTileSl] = {0.0, 0.0, 0.0, 0.0, 1.0, 1.0};
        if (a < b) {
            double slo = __builtin_inf(), shi = -__builtin_inf(), tlo = __builtin_inf(), thi = -__builtin_inf();
            double cmin = 1.0, amax = 0.0;
            for (int k = a; k < b; ++k) {
                const double4 L = sl[k];
                slo = fmin(slo, L.x); shi = fmax(shi, L.x);
                tlo = fmin(tlo, L.y); thi = fmax(thi, L.y);
                cmin = fmin(cmin, sqrt(fmax(1.0 - (L.z * L.z + L.w * L.w), 0.0)));
                amax = fmax(amax, fmax(fabs(L.x), fabs(L.y)));
            }
            q[0] = slo; q[1] = shi; q[2] = tlo; q[3] = thi;
            q[4] = (1.0 + 1e-12) / fmax(cmin, 1e-3);
            q[5] = 1.0 + amax;
            const double dwm = fmax(fabs(f.dwlo), fabs(f.dwhi));
            sp = fmax((shi - slo) * dwm * f.inv_cs, (thi - tlo) * dwm * f.inv_ct);
            ne = 1.0;
        }
        for (int k = 0; k < kTileSl; ++k) tsl[t * kTileSl + k] = q[k];
    }
    red[0][threadIdx.x] = sp;
    red[1][threadIdx.x] = ne;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            red[0][threadIdx.x] += red[0][threadIdx.x + w];
            red[1][threadIdx.x] += red[1][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        atomicAdd(&spread[0], red[0][0]);
        atomicAdd(&spread[1], red[1][0]);
    }
}

__global__ void __launch_bounds__(256) k_gather_slopes(const int32_t* __restrict__ perm, const double4* __restrict__ sl,
                                                       int64_t nl, double4* __restrict__ sl_sorted) {
    for (int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x; j < nl; j += (int64_t)gridDim.x * 256)
        sl_sorted[j] = sl[perm[j]];
}

// sort key of a line: tile-major, then cell inside the tile
__global__ void __launch_bounds__(256) k_line_keys(const double* __restrict__ st, int64_t nl, Frame f,
                                                   uint64_t* __restrict__ keys, int32_t* __restrict__ vals) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < nl; i += (int64_t)gridDim.x * 256) {
        const int cx = cell_of(st[i], f.so, f.inv_cs, f.gx);
        const int cy = cell_of(st[nl + i], f.to, f.inv_ct, f.gy);
        const uint64_t tile = (uint64_t)(cy / kTS) * (uint64_t)f.tx + (uint64_t)(cx / kTS);
        keys[i] = tile * kCells + (uint64_t)((cy % kTS) * kTS + (cx % kTS));
        vals[i] = (int32_t)i;
    }
}

// CSR start of every cell over the sorted lines: start[c] = first sorted position with key >= c
// (empty cells get the start of the next non-empty one; start[ncell] = nl), and the sorted
// (s, t) of each line.
__global__ void __launch_bounds__(256) k_line_cells(const uint64_t* __restrict__ keys, const int32_t* __restrict__ perm,
                                                    const double* __restrict__ st, int64_t nl, int64_t ncell,
                                                    int32_t* __restrict__ start, double2* __restrict__ st_sorted) {
    for (int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x; j < nl; j += (int64_t)gridDim.x * 256) {
        const int64_t k = (int64_t)keys[j];
        const int64_t kp = j == 0 ? -1 : (int64_t)keys[j - 1];
        for (int64_t c = kp + 1; c <= k && c <= ncell; ++c) start[c] = (int32_t)j;
        if (j == nl - 1)
            for (int64_t c = k + 1; c <= ncell; ++c) start[c] = (int32_t)nl;
        const int32_t i = perm[j];
        st_sorted[j] = make_double2(st[i], st[nl + i]);
    }
}

// Clears counters and flags (a kernel rather than a memset node, so callers may capture it).
__global__ void k_zero_counter(unsigned* __restrict__ p, int n) {
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0u;
}

// Workgroups of kTileWaves wavefronts share one LDS copy of the Inorm table: its reads are
// per-lane gathers over 16 KB, which the vector L1 serves one cache line at a time while LDS
// serves them in a few cycles.  Each wavefront takes tiles from a device counter, so the
// wavefronts of a workgroup stay busy until the map is done.
constexpr int kTileWaves = 8;
// Line sets of at least this many lines take the union walk.
constexpr int64_t kMaxTiledLines = (int64_t)1 << 28;

// H2: weight Inorm / h^2 instead of Inorm / h (a compile-time choice, so the default kernel
// carries no test of it).  TILT: a nearly parallel set (tv).  Parallel sets run at 4 wavefronts
// per SIMD (at most 128 VGPRs: two workgroups per CU); the nearly parallel kernels need more
// registers and run at 2.
template <int K, int MODE, bool H2 = false, bool TILT = false>
__global__ void __launch_bounds__(64 * kTileWaves, TILT ? 2 : 4) k_lines_tiled(T32 T, Frame f, TileCat cat,
                                                                 const int32_t* __restrict__ cstart,
                                                                 const int32_t* __restrict__ perm,
                                                                 const double2* __restrict__ st_sorted,
                                                                 const double* __restrict__ o, int64_t nl,
                                                                 const float* __restrict__ lut, double inv_dq,
                                                                 float* __restrict__ out, int64_t ldo,
                                                                 uint64_t* __restrict__ digest,
                                                                 unsigned* __restrict__ next_tile, TiltView tv) {
    extern __shared__ double2 dyn2[];             // per wavefront: accumulators, particle fields
    __shared__ LineTileLDS Ls[kTileWaves];
    __shared__ float lut_s[kLutN + 2];             // the table and a zero after it
    for (int i = threadIdx.x; i < kLutN + 2; i += 64 * kTileWaves) lut_s[i] = lut[i];
    __syncthreads();
    const int w = threadIdx.x >> 6;
    double* dyn = reinterpret_cast<double*>(dyn2) + (size_t)w * (lines_tile_lds(cat.n, MODE) / sizeof(double));
    const int64_t ntiles = (int64_t)f.tx * f.ty;
    for (;;) {
        unsigned t = 0;
        if ((threadIdx.x & 63) == 0) t = atomicAdd(next_tile, 1u);
        const int64_t tile = (int64_t)(unsigned)__builtin_amdgcn_readfirstlane((int)__shfl((int)t, 0));
        if (tile >= ntiles) break;
        lines_tile<K, MODE, false, H2, TILT>(T, f, cat, cstart, perm, st_sorted, o, nl, lut_s, inv_dq, out, ldo, digest,
                                             tile, Ls[w], dyn, tv);
    }
}

// Line binning shared by interpolation and digest: returns false (and no error) when the
// lines are neither a parallel nor a nearly parallel set of unit-length lines.
struct Binned {
    Frame f;
    const int32_t* cells;
    const int32_t* perm;
    const double2* sts;
    unsigned* next_tile;       // tile counter of the line-map kernel
    bool tilt = false;         // nearly parallel: tv holds the slopes
    TiltView tv;
    double amax = 0.0;         // largest magnitude of the lines' points
};

template <int K, int MODE, bool H2, bool TILT>
int launch_tiled(pmx_ctx* c, const Binned& B, const TileCat& cat, const double* o, int64_t nl, float* out, int64_t ldo,
                 uint64_t* digest) {
    const Frame& f = B.f;
    const int64_t ntiles = (int64_t)f.tx * f.ty;
    const size_t lds = lines_tile_lds(cat.n, MODE) * kTileWaves;
    int per_cu = 0;
    PMX_HIP(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_lines_tiled<K, MODE, H2, TILT>, 64 * kTileWaves,
                                                            lds));
    const int64_t resident = (int64_t)std::max(1, per_cu) * c->ncu;
    const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(resident, (ntiles + kTileWaves - 1) / kTileWaves));
    T32 T = tree32(c, 64, kMaxRunLines);
    T.bucket = c->tune.tiled_bucket;
    hipLaunchKernelGGL(k_zero_counter, dim3(1), dim3(64), 0, c->stream, B.next_tile, 1);
    hipLaunchKernelGGL((k_lines_tiled<K, MODE, H2, TILT>), dim3(grid), dim3(64 * kTileWaves), lds, c->stream, T, f, cat,
                       B.cells, B.perm, B.sts, o, nl, c->lutf.as<float>() + (size_t)K * kLutF, 1.0 / c->lut_dq[K], out,
                       ldo, digest, B.next_tile, B.tv);
    PMX_HIP(c, hipGetLastError());
    return PMX_OK;
}

template <int MODE, bool H2, bool TILT>
int launch_tiled_k(pmx_ctx* c, const Binned& B, const TileCat& cat, const double* o, int64_t nl, float* out,
                   int64_t ldo, uint64_t* digest) {
    switch (c->tree.kernel) {
    case kM4: return launch_tiled<kM4, MODE, H2, TILT>(c, B, cat, o, nl, out, ldo, digest);
    case kM5: return launch_tiled<kM5, MODE, H2, TILT>(c, B, cat, o, nl, out, ldo, digest);
    case kM6: return launch_tiled<kM6, MODE, H2, TILT>(c, B, cat, o, nl, out, ldo, digest);
    case kC2: return launch_tiled<kC2, MODE, H2, TILT>(c, B, cat, o, nl, out, ldo, digest);
    case kC4: return launch_tiled<kC4, MODE, H2, TILT>(c, B, cat, o, nl, out, ldo, digest);
    default: return launch_tiled<kC6, MODE, H2, TILT>(c, B, cat, o, nl, out, ldo, digest);
    }
}

template <int MODE>
int launch_binned(pmx_ctx* c, const Binned& B, const TileCat& cat, const double* o, int64_t nl, float* out,
                  int64_t ldo, uint64_t* digest, bool h2) {
    if (B.tilt) {
        if (h2) return launch_tiled_k<MODE, true, true>(c, B, cat, o, nl, out, ldo, digest);
        return launch_tiled_k<MODE, false, true>(c, B, cat, o, nl, out, ldo, digest);
    }
    if (h2) return launch_tiled_k<MODE, true, false>(c, B, cat, o, nl, out, ldo, digest);
    return launch_tiled_k<MODE, false, false>(c, B, cat, o, nl, out, ldo, digest);
}

// Orthonormal basis (u, v) of the plane normal to the unit vector d; exact axes for an
// axis-aligned d.
void plane_basis(const double d0[3], Frame& f) {
    for (int k = 0; k < 3; ++k) f.d[k] = d0[k];
    const double ax = fabs(d0[0]), ay = fabs(d0[1]), az = fabs(d0[2]);
    if (d0[0] == 0.0 && d0[1] == 0.0) {
        f.u[0] = 1; f.u[1] = 0; f.u[2] = 0; f.v[0] = 0; f.v[1] = 1; f.v[2] = 0;
    } else if (d0[0] == 0.0 && d0[2] == 0.0) {
        f.u[0] = 1; f.u[1] = 0; f.u[2] = 0; f.v[0] = 0; f.v[1] = 0; f.v[2] = 1;
    } else if (d0[1] == 0.0 && d0[2] == 0.0) {
        f.u[0] = 0; f.u[1] = 1; f.u[2] = 0; f.v[0] = 0; f.v[1] = 0; f.v[2] = 1;
    } else {
        double e[3] = {0, 0, 0};
        e[(ax <= ay && ax <= az) ? 0 : (ay <= az ? 1 : 2)] = 1.0;
        double u[3] = {d0[1] * e[2] - d0[2] * e[1], d0[2] * e[0] - d0[0] * e[2], d0[0] * e[1] - d0[1] * e[0]};
        const double un = std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
        for (int k = 0; k < 3; ++k) f.u[k] = u[k] / un;
        f.v[0] = d0[1] * f.u[2] - d0[2] * f.u[1];
        f.v[1] = d0[2] * f.u[0] - d0[0] * f.u[2];
        f.v[2] = d0[0] * f.u[1] - d0[1] * f.u[0];
    }
    f.au = unit_axis(f.u);
    f.av = unit_axis(f.v);
    f.ad = unit_axis(f.d);
}

// Nearly parallel sets: a line at most acos(0.7) = 45.6 degrees from the mean direction (its
// footprint on the reference plane widens by at most 1 / 0.7), and the tiles' slope spread adds
// at most kTiltMaxSpread cells of image width on average over the particles' depth range.
constexpr double kTiltMinCos = 0.7;
constexpr double kTiltMaxSpread = 2.0;

int bin_lines(pmx_ctx* c, const LineWork* ws, const double* o, const double* d, int64_t nl, bool* ok,
              Binned* out, bool allow_tilt = true) {
    *ok = false;
    // (the tile kernel addresses the sorted lines with 32-bit byte offsets: fewer than 2^28 lines)
    if (nl < 1 || nl >= kMaxTiledLines || c->tree.n < 1) return PMX_OK;
    // direction of line 0 and the particle box (root node, or the single leaf)
    double d0[3];
    for (int k = 0; k < 3; ++k)
        PMX_HIP(c, hipMemcpyAsync(&d0[k], d + (int64_t)k * nl, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    Node64 root{};
    if (c->tree.n >= 2)
        PMX_HIP(c, hipMemcpyAsync(&root, c->nodes64.as<Node64>() + 1, sizeof(Node64), hipMemcpyDeviceToHost, c->stream));
    else {
        Leaf64 L{};
        if (c->tree.f32in) {
            float4 H;
            PMX_HIP(c, hipMemcpyAsync(&H, c->leaf32.p, sizeof(float4), hipMemcpyDeviceToHost, c->stream));
            PMX_HIP(c, hipStreamSynchronize(c->stream));
            L.x = H.x; L.y = H.y; L.z = H.z;
        } else {
            PMX_HIP(c, hipMemcpyAsync(&L.x, c->cols64.as<double>(), 8, hipMemcpyDeviceToHost, c->stream));
            PMX_HIP(c, hipMemcpyAsync(&L.y, c->cols64.as<double>() + 1, 8, hipMemcpyDeviceToHost, c->stream));
            PMX_HIP(c, hipMemcpyAsync(&L.z, c->cols64.as<double>() + 2, 8, hipMemcpyDeviceToHost, c->stream));
        }
        root.mn[0] = root.mx[0] = L.x; root.mn[1] = root.mx[1] = L.y; root.mn[2] = root.mx[2] = L.z;
    }
    PMX_HIP(c, hipStreamSynchronize(c->stream));
    const double len2 = d0[0] * d0[0] + d0[1] * d0[1] + d0[2] * d0[2];
    if (!(fabs(len2 - 1.0) <= 1e-12)) return PMX_OK;
    Frame f{};
    plane_basis(d0, f);

    const int blocks = (int)std::min<int64_t>(kBinBlocks, (nl + 255) / 256);
    double* st;
    double* partd;
    if (ws) {
        st = ws->st;
        partd = ws->part;
    } else {
        PMX_TRY(ensure(c, c->tl_st, sizeof(double) * 2 * (size_t)nl + 64));
        PMX_TRY(ensure(c, c->tl_part, kBinPartBytes));
        st = c->tl_st.as<double>();
        partd = c->tl_part.as<double>();
    }
    // after the partial bounds: the flags, then the line-map kernel's tile counter
    int* flag = reinterpret_cast<int*>(partd + 5 * kBinBlocks);
    hipLaunchKernelGGL(k_zero_counter, dim3(1), dim3(64), 0, c->stream, reinterpret_cast<unsigned*>(flag), 2);
    hipLaunchKernelGGL(k_line_project, dim3(blocks), dim3(256), 0, c->stream, o, d, nl, f, st, partd, flag);
    PMX_HIP(c, hipGetLastError());
    double part[5 * kBinBlocks];
    int hflag = 0;
    PMX_HIP(c, hipMemcpyAsync(part, partd, sizeof(double) * 5 * (size_t)blocks, hipMemcpyDeviceToHost, c->stream));
    PMX_HIP(c, hipMemcpyAsync(&hflag, flag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    PMX_HIP(c, hipStreamSynchronize(c->stream));
    double4* sl = nullptr;
    bool tilt = false;
    if (hflag) {
        // not one direction: a nearly parallel set of unit lines?  (not with caller-supplied work
        // buffers, which hold no slopes)
        if (ws || !allow_tilt) return PMX_OK;
        hipLaunchKernelGGL(k_zero_counter, dim3(1), dim3(64), 0, c->stream, reinterpret_cast<unsigned*>(flag), 1);
        hipLaunchKernelGGL(k_line_dirsum, dim3(blocks), dim3(256), 0, c->stream, d, nl, partd, flag);
        PMX_HIP(c, hipGetLastError());
        PMX_HIP(c, hipMemcpyAsync(part, partd, sizeof(double) * 3 * (size_t)blocks, hipMemcpyDeviceToHost, c->stream));
        PMX_HIP(c, hipMemcpyAsync(&hflag, flag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        PMX_HIP(c, hipStreamSynchronize(c->stream));
        if (hflag) return PMX_OK;
        double m[3] = {0.0, 0.0, 0.0};
        for (int b = 0; b < blocks; ++b)
            for (int k = 0; k < 3; ++k) m[k] += part[3 * b + k];
        const double mn = std::sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
        if (!(mn > 0.0) || !(mn < 1e300)) return PMX_OK;
        for (int k = 0; k < 3; ++k) m[k] /= mn;
        f = Frame{};
        plane_basis(m, f);
        // reference plane through the particles' mid-depth; their depth range around it
        double wlo = 0.0, whi = 0.0;
        for (int k = 0; k < 3; ++k) {
            const double p = f.d[k] * root.mn[k], q = f.d[k] * root.mx[k];
            wlo += std::min(p, q);
            whi += std::max(p, q);
        }
        f.w0 = 0.5 * (wlo + whi);
        f.dwlo = wlo - f.w0;
        f.dwhi = whi - f.w0;
        PMX_TRY(ensure(c, c->tl_sl, 2 * sizeof(double4) * (size_t)nl + 64));
        sl = c->tl_sl.as<double4>();
        hipLaunchKernelGGL(k_zero_counter, dim3(1), dim3(64), 0, c->stream, reinterpret_cast<unsigned*>(flag), 1);
        hipLaunchKernelGGL(k_line_project_tilt, dim3(blocks), dim3(256), 0, c->stream, o, d, nl, f, kTiltMinCos, st, sl,
                           partd, flag);
        PMX_HIP(c, hipGetLastError());
        PMX_HIP(c, hipMemcpyAsync(part, partd, sizeof(double) * 5 * (size_t)blocks, hipMemcpyDeviceToHost, c->stream));
        PMX_HIP(c, hipMemcpyAsync(&hflag, flag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        PMX_HIP(c, hipStreamSynchronize(c->stream));
        if (hflag) return PMX_OK;
        tilt = true;
    }
    double smin = INFINITY, smax = -INFINITY, tmin = INFINITY, tmax = -INFINITY, amax = 0.0;
    for (int b = 0; b < blocks; ++b) {
        smin = std::min(smin, part[5 * b]);
        smax = std::max(smax, -part[5 * b + 1]);
        tmin = std::min(tmin, part[5 * b + 2]);
        tmax = std::max(tmax, -part[5 * b + 3]);
        amax = std::max(amax, -part[5 * b + 4]);
    }
    // scale of all coordinates: particle box (the single particle when n = 1) and line origins
    // (and the lines' reference-plane points)
    double E = amax;
    for (int k = 0; k < 3; ++k) E = std::max(E, std::max(std::fabs(root.mn[k]), std::fabs(root.mx[k])));
    if (!(E < 1e150)) return PMX_OK;
    frame_scale(f, 2.0 * E);
    // cells: about one line per cell, lines at cell centres for a regular lattice
    const double es = smax - smin, et = tmax - tmin;
    int64_t gx, gy;
    if (es > 0.0 && et > 0.0) {
        gx = std::max<int64_t>(1, std::llround(std::sqrt((double)nl * es / et)));
        gy = std::max<int64_t>(1, std::llround((double)nl / (double)gx));
    } else if (es > 0.0) {
        gx = nl; gy = 1;
    } else if (et > 0.0) {
        gx = 1; gy = nl;
    } else {
        gx = 1; gy = 1;
    }
    gx = std::min<int64_t>(gx, 1 << 20);
    gy = std::min<int64_t>(gy, 1 << 20);
    f.cs = gx > 1 ? es / (double)(gx - 1) : std::max(es, 1.0);
    f.ct = gy > 1 ? et / (double)(gy - 1) : std::max(et, 1.0);
    if (!(f.cs > 0.0) || !(f.ct > 0.0)) return PMX_OK;
    f.near = (double)c->tune.tiled_near;
    f.inv_cs = 1.0 / f.cs;
    f.inv_ct = 1.0 / f.ct;
    f.so = smin - 0.5 * f.cs;
    f.to = tmin - 0.5 * f.ct;
    f.gx = (int)gx;
    f.gy = (int)gy;
    f.tx = (int)((gx + kTS - 1) / kTS);
    f.ty = (int)((gy + kTS - 1) / kTS);
    const int64_t ncell = (int64_t)f.tx * f.ty * kCells;
    int bits = 1;
    while (bits < 62 && ((int64_t)1 << bits) < ncell) ++bits;

    const size_t cells_bytes = ((sizeof(int32_t) * (size_t)(ncell + 1) + 15) & ~(size_t)15) + sizeof(double2) * (size_t)nl;
    uint64_t *k0, *k1;
    int32_t *v0, *v1;
    uint32_t* hist = nullptr;
    char* cbuf;
    if (ws) {
        if (cells_bytes > ws->cells_bytes) {
            char msg[192];
            snprintf(msg, sizeof(msg),
                     "line map: the line set needs %lld binning cells, beyond the scratch capacity (lines far from a "
                     "square lattice)", (long long)ncell);
            return set_error(c, PMX_ERR_CAPACITY, msg);
        }
        k0 = ws->keys0; v0 = ws->vals0; k1 = ws->keys1; v1 = ws->vals1;
        hist = ws->hist;
        cbuf = ws->cells;
    } else {
        PMX_TRY(ensure(c, c->scratch_a, 12 * (size_t)nl + 64));
        PMX_TRY(ensure(c, c->scratch_b, 12 * (size_t)nl + 64));
        PMX_TRY(ensure(c, c->tl_cells, cells_bytes + 64));
        k0 = c->scratch_a.as<uint64_t>();
        v0 = reinterpret_cast<int32_t*>(k0 + nl);
        k1 = c->scratch_b.as<uint64_t>();
        v1 = reinterpret_cast<int32_t*>(k1 + nl);
        cbuf = c->tl_cells.as<char>();
    }
    const dim3 g(grid_for(nl, 256, 8192)), b(256);
    hipLaunchKernelGGL(k_line_keys, g, b, 0, c->stream, st, nl, f, k0, v0);
    bool alt = false;
    PMX_TRY(radix_sort_pairs(c, k0, v0, nl, k1, v1, 0, bits, &alt, hist));
    const uint64_t* ks = alt ? k1 : k0;
    const int32_t* perm = alt ? v1 : v0;
    int32_t* cells = reinterpret_cast<int32_t*>(cbuf);
    double2* sts = reinterpret_cast<double2*>(cbuf + ((sizeof(int32_t) * (size_t)(ncell + 1) + 15) & ~(size_t)15));
    hipLaunchKernelGGL(k_line_cells, g, b, 0, c->stream, ks, perm, st, nl, ncell, cells, sts);
    PMX_HIP(c, hipGetLastError());
    if (tilt) {
        // slopes in sorted order, per-tile bounds, and the decision on the slope spread
        const int64_t ntiles = (int64_t)f.tx * f.ty;
        PMX_TRY(ensure(c, c->tl_tsl, sizeof(double) * kTileSl * (size_t)ntiles + 64));
        double4* sls = sl + nl;
        double* spread = partd;   // (the bounds are read already)
        hipLaunchKernelGGL(k_gather_slopes, g, b, 0, c->stream, perm, sl, nl, sls);
        hipLaunchKernelGGL(k_zero_counter, dim3(1), dim3(64), 0, c->stream, reinterpret_cast<unsigned*>(spread), 4);
        hipLaunchKernelGGL(k_tile_slopes, dim3((unsigned)((ntiles + 255) / 256)), b, 0, c->stream, cells, sls, ntiles, f,
                           c->tl_tsl.as<double>(), spread);
        PMX_HIP(c, hipGetLastError());
        double hs[2] = {0.0, 0.0};
        PMX_HIP(c, hipMemcpyAsync(hs, spread, sizeof(hs), hipMemcpyDeviceToHost, c->stream));
        PMX_HIP(c, hipStreamSynchronize(c->stream));
        if (!(hs[1] > 0.0) || !(hs[0] <= kTiltMaxSpread * hs[1])) return PMX_OK;
        out->tilt = true;
        out->tv.sl = sls;
        out->tv.tb = c->tl_tsl.as<double>();
        out->tv.d = d;
    }
    out->amax = amax;
    out->f = f;
    out->cells = cells;
    out->perm = perm;
    out->sts = sts;
    out->next_tile = reinterpret_cast<unsigned*>(flag + 1);
    *ok = true;
    return PMX_OK;
}

}  // namespace


int interp_lines_tiled(pmx_ctx* c, const int* slots, const int* shep, int ncols, const double* o, const double* d,
                       int64_t nl, float* out, int64_t ldo, bool* used, const LineWork* ws) {
    *used = false;
    if (!c->tune.lines_tiled) return PMX_OK;
    bool ok = false;
    Binned B;
    PMX_TRY(bin_lines(c, ws, o, d, nl, &ok, &B));
    if (!ok) return PMX_OK;
    PMX_HIP(c, hipEventRecord(c->ev[8], c->stream));
    for (int base = 0; base < ncols || (ncols == 0 && base == 0); base += kMaxCols) {
        TileCat cat{};
        cat.n = std::min(kMaxCols, ncols - base);
        for (int j = 0; j < cat.n; ++j) { cat.fcol[j] = 1 + slots[base + j]; cat.shep[j] = shep[base + j]; }
        float* ob = out + (int64_t)base * ldo;
        PMX_TRY(launch_binned<0>(c, B, cat, o, nl, ob, ldo, nullptr, c->line_inv_h2));
        if (ncols == 0) break;
    }
    PMX_HIP(c, hipGetLastError());
    *used = true;
    return PMX_OK;
}

int lines_plan(pmx_ctx* c, const double* o, const double* d, int64_t nl, lmap::LinePlan* plan, bool* ok) {
    Binned B;
    PMX_TRY(bin_lines(c, nullptr, o, d, nl, ok, &B));
    if (*ok) {
        plan->f = B.f;
        plan->cstart = B.cells;
        plan->perm = B.perm;
        plan->sts = B.sts;
        plan->amax = B.amax;
        plan->tilt = B.tilt;
        plan->tv = B.tv;
    }
    return PMX_OK;
}

int line_digest_tiled(pmx_ctx* c, const double* o, const double* d, int64_t nl, uint64_t* digest, bool* used) {
    *used = false;
    bool ok = false;
    Binned B;
    PMX_TRY(bin_lines(c, nullptr, o, d, nl, &ok, &B));
    if (!ok) return PMX_OK;
    TileCat cat{};
    PMX_TRY(launch_binned<1>(c, B, cat, o, nl, nullptr, 0, digest, false));
    PMX_HIP(c, hipGetLastError());
    *used = true;
    return PMX_OK;
}

}  // namespace pmx
