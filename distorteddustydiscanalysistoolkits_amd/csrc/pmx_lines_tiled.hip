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
//   * one wavefront per tile walks the LBVH once (wave-uniform node test: the node box
//     projected on the plane against the tile rectangle, radius K hmax), and each bucket of
//     up to 64 particles (contiguous in Morton order) is taken one particle per lane;
//   * each lane visits only the cells its disc covers, tests those lines in FP64 on the
//     projected coordinates with a guard band, re-checks band cases with the reference's
//     FP64 formula on the FP64 columns, and adds m/rho * W_line and the field terms to FP64
//     accumulators of the tile's lines in LDS (ds_add_f64);
//   * the wave then applies the Shepard division and writes the tile's outputs.
// The accepted pair set is the reference's FP64 predicate set; only the order of the FP64
// additions into one line differs from the per-line traversal (not bit-reproducible across
// runs, agreement ~1e-15 relative).  Lines that are not all parallel and unit-length use the
// per-tile union walk of pmx_interp32.hip instead.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "pmx_internal.h"
#include "pmx_tree32.h"

namespace pmx {

namespace {

constexpr int kTS = 8;                 // cells per tile side
constexpr int kCells = kTS * kTS;      // one wavefront per tile: lane = cell when loading
constexpr int kCap = 64;               // lines per LDS chunk of a tile (a lattice tile holds 64)
constexpr int kMaxCols = 8;            // field columns per launch

struct Frame {
    double u[3], v[3], d[3];
    double so, to, cs, ct;             // cell origin and size along s and t
    int gx, gy, tx, ty;                // cells and tiles along s and t
    double band;                       // absolute band on d^2 around (K h)^2 -> exact re-check
    double pad;                        // absolute padding of footprint radii
    double sqrt_band;                  // sqrt(band): R + sqrt(band) >= sqrt(R^2 + band)
    double inv_cs, inv_ct;             // 1 / cell size (binning and footprints use the same products)
    double near;                       // cells: nodes whose box is this close are read whole
};

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
    if (bad) atomicOr(flag, 1);
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

__device__ inline int clamp_cell(double c, int g) { return c < 0.0 ? 0 : (c > (double)(g - 1) ? g - 1 : (int)c); }
__device__ inline int cell_of(double v, double org, double inv_size, int g) { return clamp_cell(floor((v - org) * inv_size), g); }

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

struct TileCat {
    int n;
    int fcol[kMaxCols];   // payload column of each output (1.. = quant)
    int shep[kMaxCols];
};

// Extent of an outward-rounded FP32 box projected on a unit axis a.
__device__ inline void project_box(const Node32& b, const double a[3], double& lo, double& hi) {
    const double mn[3] = {b.mnx, b.mny, b.mnz}, mx[3] = {b.mxx, b.mxy, b.mxz};
    lo = 0.0;
    hi = 0.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double p = a[k] * mn[k], q = a[k] * mx[k];
        lo += fmin(p, q);
        hi += fmax(p, q);
    }
}

// MODE 0: interpolation (FP64 weight sums).  MODE 1: neighbour digest per line (count, sum
// of sorted particle indices, sum of their squares; unsigned 64-bit, wrapping).
template <int K, int MODE>
__global__ void __launch_bounds__(64) k_lines_tiled(T32 T, Frame f, TileCat cat, const int32_t* __restrict__ cstart,
                                                    const int32_t* __restrict__ perm,
                                                    const double2* __restrict__ st_sorted,
                                                    const double* __restrict__ o, int64_t nl,
                                                    const float* __restrict__ lut, double inv_dq,
                                                    float* __restrict__ out, int64_t ldo,
                                                    uint64_t* __restrict__ digest) {
    // dynamic LDS: line s, t [kCap]; accumulators [nacc][kCap]; particle fields [ncol][64]
    extern __shared__ double dyn[];
    __shared__ int32_t cs[kCells + 1];
    // sweep state: one particle per lane, and the owner lane of each (particle, row) item
    __shared__ double w_sb[64], w_tb[64], w_r2[64], w_wb[64], w_ih[64];
    __shared__ int32_t w_p[64], w_cx0[64], w_cx1[64], w_cy0[64], w_start[64];
    __shared__ unsigned char w_own[64 * kTS];
    const int lane = threadIdx.x;
    const int nacc = MODE == 0 ? 1 + cat.n : 3;
    double* S = dyn;
    double* Tt = dyn + kCap;
    double* acc = dyn + 2 * kCap;
    unsigned long long* dg = reinterpret_cast<unsigned long long*>(dyn + 2 * kCap);
    float* w_A = reinterpret_cast<float*>(dyn + (2 + nacc) * kCap);     // [ncol][64]
    const double Kv = support(K);
    const int64_t n = T.n;
    const int64_t ntiles = (int64_t)f.tx * f.ty;

    for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        __syncthreads();
        cs[lane] = cstart[tile * kCells + lane];
        if (lane == 0) cs[kCells] = cstart[(tile + 1) * kCells];
        __syncthreads();
        const int L0 = cs[0], L1 = cs[kCells];
        if (L0 >= L1) continue;
        const int txi = (int)(tile % f.tx), tyi = (int)(tile / f.tx);
        const int cx_base = txi * kTS, cy_base = tyi * kTS;
        // tile rectangle in the plane; border tiles also hold lines clamped into the grid
        const double rs0 = cx_base == 0 ? -__builtin_inf() : f.so + (double)cx_base * f.cs;
        const double rs1 = cx_base + kTS >= f.gx ? __builtin_inf() : f.so + (double)(cx_base + kTS) * f.cs;
        const double rt0 = cy_base == 0 ? -__builtin_inf() : f.to + (double)cy_base * f.ct;
        const double rt1 = cy_base + kTS >= f.gy ? __builtin_inf() : f.to + (double)(cy_base + kTS) * f.ct;
        const double near_s = f.near * f.cs, near_t = f.near * f.ct;

        for (int C0 = L0; C0 < L1; C0 += kCap) {
            const int m = min(kCap, L1 - C0);
            for (int k = lane; k < kCap; k += 64) {
                if (k < m) {
                    const double2 v = st_sorted[C0 + k];
                    S[k] = v.x;
                    Tt[k] = v.y;
                }
                for (int a = 0; a < nacc; ++a) {
                    if (MODE == 0) acc[a * kCap + k] = 0.0;
                    else dg[a * kCap + k] = 0ull;
                }
            }
            __syncthreads();

            auto need = [&](const Node32& nd) {
                double s0, s1, t0, t1;
                project_box(nd, f.u, s0, s1);
                project_box(nd, f.v, t0, t1);
                const double gs = fmax(0.0, fmax(s0 - rs1, rs0 - s1));
                const double gt = fmax(0.0, fmax(t0 - rt1, rt0 - t1));
                const double R = Kv * (double)nd.hmax;
                const double Rp = R + f.sqrt_band + f.pad;
                return gs * gs + gt * gt <= Rp * Rp;
            };
            // One sweep = up to 64 particles of a leaf run, one per lane.  Each lane finds the
            // rows and columns of this tile its footprint covers; the (particle, row) items of the
            // sweep are dealt round-robin to the lanes (particles with large h do not hold the
            // other lanes idle), and a row's lines are one contiguous run of sorted lines.
            auto sweep = [&](int64_t p, bool valid) {
                int nrow = 0;
                if (valid) {
                    const float4 H = T.leaf[p];
                    const float4 Lo = T.leaflo[p];
                    const double x = (double)H.x + (double)Lo.x, y = (double)H.y + (double)Lo.y,
                                 z = (double)H.z + (double)Lo.z;
                    const double sb = (x * f.u[0] + y * f.u[1]) + z * f.u[2];
                    const double tb = (x * f.v[0] + y * f.v[1]) + z * f.v[2];
                    // reject on the FP32 h (round-to-nearest, widened by 1e-6) before more loads
                    const double Ru = Kv * (double)H.w * (1.0 + 1e-6);
                    const double Rpu = Ru + f.sqrt_band + f.pad;
                    if (!(sb + Rpu < rs0 || sb - Rpu > rs1 || tb + Rpu < rt0 || tb - Rpu > rt1)) {
                        const double hb = T.cols64[4 * n + p];
                        const double rr = Kv * hb;
                        const double r2 = rr * rr;
                        const double Rp = rr + f.sqrt_band + f.pad;
                        // global cell range of the footprint, clamped like the binning, then tile-local
                        const int gx0 = clamp_cell(floor((sb - Rp - f.so) * f.inv_cs), f.gx);
                        const int gx1 = clamp_cell(floor((sb + Rp - f.so) * f.inv_cs), f.gx);
                        const int gy0 = clamp_cell(floor((tb - Rp - f.to) * f.inv_ct), f.gy);
                        const int gy1 = clamp_cell(floor((tb + Rp - f.to) * f.inv_ct), f.gy);
                        const int cx0 = max(gx0 - cx_base, 0), cx1 = min(gx1 - cx_base, kTS - 1);
                        const int cy0 = max(gy0 - cy_base, 0), cy1 = min(gy1 - cy_base, kTS - 1);
                        if (cx0 <= cx1 && cy0 <= cy1) {
                            nrow = cy1 - cy0 + 1;
                            w_sb[lane] = sb;
                            w_tb[lane] = tb;
                            w_r2[lane] = r2;
                            w_p[lane] = (int32_t)p;
                            w_cx0[lane] = cx0;
                            w_cx1[lane] = cx1;
                            w_cy0[lane] = cy0;
                            if (MODE == 0) {
                                w_wb[lane] = T.w64[p];
                                w_ih[lane] = T.ih64[p];
                                for (int c = 0; c < cat.n; ++c) w_A[c * 64 + lane] = T.pay[(int64_t)cat.fcol[c] * n + p];
                            }
                        }
                    }
                }
                // exclusive prefix of the item counts across the wave
                int incl = nrow;
                for (int o2 = 1; o2 < 64; o2 <<= 1) {
                    const int u = __shfl_up(incl, o2);
                    if (lane >= o2) incl += u;
                }
                const int total = __shfl(incl, 63);
                if (total == 0) return;
                const int excl = incl - nrow;
                w_start[lane] = excl;
                for (int i = 0; i < nrow; ++i) w_own[excl + i] = (unsigned char)lane;
                __syncthreads();
                for (int t = lane; t < total; t += 64) {
                    const int j = w_own[t];
                    const int row = w_cy0[j] + (t - w_start[j]);
                    const int a = max(cs[row * kTS + w_cx0[j]], C0);
                    const int b = min(cs[row * kTS + w_cx1[j] + 1], C0 + m);
                    if (a >= b) continue;
                    const double sb = w_sb[j], tb = w_tb[j], r2 = w_r2[j];
                    const int64_t pj = w_p[j];
                    // lanes on the same row start at different columns (fewer LDS bank conflicts)
                    const int len = b - a;
                    int k = a + (lane % len);
                    for (int it = 0; it < len; ++it, k = (k + 1 == b) ? a : k + 1) {
                        const int kk = k - C0;
                        const double ds = sb - S[kk], dt = tb - Tt[kk];
                        const double d2 = ds * ds + dt * dt;
                        if (d2 > r2 + f.band) continue;
                        bool ok = d2 < r2 - f.band;
                        if (!ok) {
                            // reference formula on the FP64 columns (_squared_distance_line_coords)
                            const int64_t li = perm[k];
                            double D2 = 0.0, Dm = 0.0, e;
                            e = T.cols64[pj] - o[li]; D2 += e * e; Dm += e * f.d[0];
                            e = T.cols64[n + pj] - o[nl + li]; D2 += e * e; Dm += e * f.d[1];
                            e = T.cols64[2 * n + pj] - o[2 * nl + li]; D2 += e * e; Dm += e * f.d[2];
                            ok = D2 - Dm * Dm <= r2;
                        }
                        if (!ok) continue;
                        if (MODE == 0) {
                            // line_integrated_kernel_function: invh * Inorm(r invh).  r from an FP32
                            // reciprocal square root and one FP64 correction step (relative error
                            // ~1e-14), q and the table interpolation in FP64 on FP32 table values:
                            // FP32 q would carry ~1e-7 relative error, amplified by dlnI/dlnq
                            const double invh = w_ih[j];
                            const float d2f = (float)d2;
                            const float rs = d2f > 0.0f ? __builtin_amdgcn_rsqf(d2f) : 0.0f;
                            const double r0 = (double)(d2f * rs);
                            const double r = r0 + (d2 - r0 * r0) * (0.5 * (double)rs);
                            const double q = r * invh;
                            double I = 0.0;
                            if (q < Kv) {
                                const double x = q * inv_dq;
                                int i = (int)x;
                                i = i < 0 ? 0 : (i > kLutN - 2 ? kLutN - 2 : i);
                                const double t = x - (double)i;
                                const double L0 = (double)lut[i];
                                I = L0 + t * ((double)lut[i + 1] - L0);
                            }
                            const double w = w_wb[j] * (invh * I);
                            atomicAdd(&acc[kk], w);
                            for (int cc = 0; cc < cat.n; ++cc)
                                atomicAdd(&acc[(1 + cc) * kCap + kk], (double)w_A[cc * 64 + j] * w);
                        } else {
                            const unsigned long long ip = (unsigned long long)pj;
                            atomicAdd(&dg[kk], 1ull);
                            atomicAdd(&dg[kCap + kk], ip);
                            atomicAdd(&dg[2 * kCap + kk], ip * ip);
                        }
                    }
                }
                __syncthreads();
            };
            // runs of leaves (contiguous in Morton order), 64 per sweep
            auto bucket = [&](int lo, int cnt) {
                for (int base = 0; base < cnt; base += 64)
                    sweep((int64_t)lo + base + lane, base + lane < cnt);
            };
            // A node is taken whole (up to T.bucket leaves) when its box itself lies near the
            // tile; a node that reaches the tile only through a large hmax is descended down to
            // one sweep of 64, so that only the subtrees holding the large-h particles are read.
            auto take = [&](const Node32& nd, int cnt) {
                if (cnt <= 64) return true;
                if (cnt > T.bucket) return false;
                double s0, s1, t0, t1;
                project_box(nd, f.u, s0, s1);
                project_box(nd, f.v, t0, t1);
                const double gs = fmax(0.0, fmax(s0 - rs1, rs0 - s1));
                const double gt = fmax(0.0, fmax(t0 - rt1, rt0 - t1));
                return gs <= near_s && gt <= near_t;
            };
            walk(T, true, need, bucket, take);
            __syncthreads();
            for (int k = lane; k < m; k += 64) {
                const int64_t li = perm[C0 + k];
                if (MODE == 0) {
                    const double S1 = acc[k];
                    for (int c = 0; c < cat.n; ++c) {
                        float v;
                        if (S1 == 0.0) v = __builtin_nanf("");
                        else v = (float)(cat.shep[c] ? acc[(1 + c) * kCap + k] / S1 : acc[(1 + c) * kCap + k]);
                        out[(int64_t)c * ldo + li] = v;
                    }
                } else {
                    digest[li] = dg[k];
                    digest[nl + li] = dg[kCap + k];
                    digest[2 * nl + li] = dg[2 * kCap + k];
                }
            }
        }
    }
}

template <int K, int MODE>
void launch_tiled(pmx_ctx* c, const Frame& f, const TileCat& cat, const int32_t* cells, const int32_t* perm,
                  const double2* sts, const double* o, int64_t nl, float* out, int64_t ldo, uint64_t* digest) {
    const int64_t ntiles = (int64_t)f.tx * f.ty;
    const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(ntiles, 1 << 20));
    const size_t lds = sizeof(double) * (size_t)kCap * (2 + (MODE == 0 ? 1 + cat.n : 3)) +
                       sizeof(float) * 64 * (size_t)(MODE == 0 ? cat.n : 0);
    T32 T = tree32(c, 64);
    T.bucket = g_tiled_bucket;
    hipLaunchKernelGGL((k_lines_tiled<K, MODE>), dim3(grid), dim3(64), lds, c->stream, T, f, cat, cells,
                       perm, sts, o, nl, c->lutf.as<float>() + (size_t)K * kLutN, 1.0 / c->lut_dq[K], out, ldo,
                       digest);
}

template <int MODE>
void launch_tiled_k(pmx_ctx* c, const Frame& f, const TileCat& cat, const int32_t* cells, const int32_t* perm,
                    const double2* sts, const double* o, int64_t nl, float* out, int64_t ldo, uint64_t* digest) {
    switch (c->tree.kernel) {
    case kM4: launch_tiled<kM4, MODE>(c, f, cat, cells, perm, sts, o, nl, out, ldo, digest); break;
    case kM5: launch_tiled<kM5, MODE>(c, f, cat, cells, perm, sts, o, nl, out, ldo, digest); break;
    case kM6: launch_tiled<kM6, MODE>(c, f, cat, cells, perm, sts, o, nl, out, ldo, digest); break;
    case kC2: launch_tiled<kC2, MODE>(c, f, cat, cells, perm, sts, o, nl, out, ldo, digest); break;
    case kC4: launch_tiled<kC4, MODE>(c, f, cat, cells, perm, sts, o, nl, out, ldo, digest); break;
    default: launch_tiled<kC6, MODE>(c, f, cat, cells, perm, sts, o, nl, out, ldo, digest); break;
    }
}

// Line binning shared by interpolation and digest: returns false (and no error) when the
// lines are not a parallel set of unit-length lines the tiled path handles.
struct Binned {
    Frame f;
    const int32_t* cells;
    const int32_t* perm;
    const double2* sts;
};

int bin_lines(pmx_ctx* c, const double* o, const double* d, int64_t nl, bool* ok, Binned* out) {
    *ok = false;
    if (nl < 1 || nl > (int64_t)INT32_MAX - kCap || c->tree.n < 1) return PMX_OK;
    // direction of line 0 and the particle box (root node, or the single leaf)
    double d0[3];
    for (int k = 0; k < 3; ++k)
        PMX_HIP(c, hipMemcpyAsync(&d0[k], d + (int64_t)k * nl, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    Node64 root{};
    if (c->tree.n >= 2)
        PMX_HIP(c, hipMemcpyAsync(&root, c->nodes64.as<Node64>() + 1, sizeof(Node64), hipMemcpyDeviceToHost, c->stream));
    PMX_HIP(c, hipStreamSynchronize(c->stream));
    const double len2 = d0[0] * d0[0] + d0[1] * d0[1] + d0[2] * d0[2];
    if (!(fabs(len2 - 1.0) <= 1e-12)) return PMX_OK;
    Frame f{};
    for (int k = 0; k < 3; ++k) f.d[k] = d0[k];
    // orthonormal basis (u, v) of the plane normal to d; exact axes for axis-aligned d
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

    const int blocks = (int)std::min<int64_t>(1024, (nl + 255) / 256);
    PMX_TRY(ensure(c, c->tl_st, sizeof(double) * 2 * (size_t)nl + 64));
    PMX_TRY(ensure(c, c->tl_part, sizeof(double) * 5 * (size_t)blocks + sizeof(int) * 4));
    int* flag = reinterpret_cast<int*>(c->tl_part.as<double>() + 5 * blocks);
    PMX_HIP(c, hipMemsetAsync(flag, 0, sizeof(int), c->stream));
    hipLaunchKernelGGL(k_line_project, dim3(blocks), dim3(256), 0, c->stream, o, d, nl, f, c->tl_st.as<double>(),
                       c->tl_part.as<double>(), flag);
    PMX_HIP(c, hipGetLastError());
    std::vector<double> part(5 * (size_t)blocks);
    int hflag = 0;
    PMX_HIP(c, hipMemcpyAsync(part.data(), c->tl_part.p, sizeof(double) * part.size(), hipMemcpyDeviceToHost, c->stream));
    PMX_HIP(c, hipMemcpyAsync(&hflag, flag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    PMX_HIP(c, hipStreamSynchronize(c->stream));
    if (hflag) return PMX_OK;
    double smin = INFINITY, smax = -INFINITY, tmin = INFINITY, tmax = -INFINITY, amax = 0.0;
    for (int b = 0; b < blocks; ++b) {
        smin = std::min(smin, part[5 * b]);
        smax = std::max(smax, -part[5 * b + 1]);
        tmin = std::min(tmin, part[5 * b + 2]);
        tmax = std::max(tmax, -part[5 * b + 3]);
        amax = std::max(amax, -part[5 * b + 4]);
    }
    // scale of all coordinates: particle box and line origins
    double E = amax;
    if (c->tree.n >= 2)
        for (int k = 0; k < 3; ++k) E = std::max(E, std::max(std::fabs(root.mn[k]), std::fabs(root.mx[k])));
    else
        E = std::max(E, 1.0);
    if (!(E < 1e150)) return PMX_OK;
    E *= 2.0;
    // |FP64 reference d^2 - projected d^2| <= ~1e-14 E^2 (rounding of both forms, direction
    // normalisation within 1e-12); band 1e-11 E^2 leaves a wide margin
    f.band = 1e-11 * E * E;
    f.pad = 1e-12 * E;
    f.sqrt_band = std::sqrt(f.band);
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
    f.near = (double)g_tiled_near;
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

    PMX_TRY(ensure(c, c->scratch_a, 12 * (size_t)nl + 64));
    PMX_TRY(ensure(c, c->scratch_b, 12 * (size_t)nl + 64));
    PMX_TRY(ensure(c, c->tl_cells, sizeof(int32_t) * (size_t)(ncell + 1) + sizeof(double2) * (size_t)nl + 64));
    uint64_t* k0 = c->scratch_a.as<uint64_t>();
    int32_t* v0 = reinterpret_cast<int32_t*>(k0 + nl);
    uint64_t* k1 = c->scratch_b.as<uint64_t>();
    int32_t* v1 = reinterpret_cast<int32_t*>(k1 + nl);
    const dim3 g(grid_for(nl, 256, 8192)), b(256);
    hipLaunchKernelGGL(k_line_keys, g, b, 0, c->stream, c->tl_st.as<double>(), nl, f, k0, v0);
    bool alt = false;
    PMX_TRY(radix_sort_pairs(c, k0, v0, nl, k1, v1, 0, bits, &alt));
    const uint64_t* ks = alt ? k1 : k0;
    const int32_t* perm = alt ? v1 : v0;
    int32_t* cells = c->tl_cells.as<int32_t>();
    double2* sts = reinterpret_cast<double2*>(c->tl_cells.as<char>() + ((sizeof(int32_t) * (size_t)(ncell + 1) + 15) & ~(size_t)15));
    hipLaunchKernelGGL(k_line_cells, g, b, 0, c->stream, ks, perm, c->tl_st.as<double>(), nl, ncell, cells, sts);
    PMX_HIP(c, hipGetLastError());
    out->f = f;
    out->cells = cells;
    out->perm = perm;
    out->sts = sts;
    *ok = true;
    return PMX_OK;
}

}  // namespace

int g_lines_tiled = 1;
int g_tiled_bucket = 512;
int g_tiled_near = 4;

int interp_lines_tiled(pmx_ctx* c, const int* slots, const int* shep, int ncols, const double* o, const double* d,
                       int64_t nl, float* out, int64_t ldo, bool* used) {
    *used = false;
    if (!g_lines_tiled) return PMX_OK;
    bool ok = false;
    Binned B;
    PMX_TRY(bin_lines(c, o, d, nl, &ok, &B));
    if (!ok) return PMX_OK;
    PMX_HIP(c, hipEventRecord(c->ev[8], c->stream));
    for (int base = 0; base < ncols || (ncols == 0 && base == 0); base += kMaxCols) {
        TileCat cat{};
        cat.n = std::min(kMaxCols, ncols - base);
        for (int j = 0; j < cat.n; ++j) { cat.fcol[j] = 1 + slots[base + j]; cat.shep[j] = shep[base + j]; }
        launch_tiled_k<0>(c, B.f, cat, B.cells, B.perm, B.sts, o, nl, out + (int64_t)base * ldo, ldo, nullptr);
        if (ncols == 0) break;
    }
    PMX_HIP(c, hipGetLastError());
    *used = true;
    return PMX_OK;
}

int line_digest_tiled(pmx_ctx* c, const double* o, const double* d, int64_t nl, uint64_t* digest, bool* used) {
    *used = false;
    bool ok = false;
    Binned B;
    PMX_TRY(bin_lines(c, o, d, nl, &ok, &B));
    if (!ok) return PMX_OK;
    TileCat cat{};
    launch_tiled_k<1>(c, B.f, cat, B.cells, B.perm, B.sts, o, nl, nullptr, 0, digest);
    PMX_HIP(c, hipGetLastError());
    *used = true;
    return PMX_OK;
}

}  // namespace pmx
