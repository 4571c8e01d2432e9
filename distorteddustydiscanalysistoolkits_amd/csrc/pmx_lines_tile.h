// pmx_lines_tile.h — the particle-centric line-map tile procedure (one wavefront per 8 x 8
// cells of a parallel line set) shared by the per-call kernel and the batched persistent
// kernel.  Reference semantics: line_integrated_scalar_interpolation.jl:123-166 (itpScatter),
// line_integrated_kernel.jl:13-19, 132-137 (table lookup, 1/h prefactor).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pmx_internal.h"
#include "pmx_tree32.h"

namespace pmx {
namespace lmap {

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
    double pad32;                      // pad + bound on the projection of a position's FP32 residual
    double sqrt_band;                  // sqrt(band): R + sqrt(band) >= sqrt(R^2 + band)
    double inv_cs, inv_ct;             // 1 / cell size (binning and footprints use the same products)
    double near;                       // cells: nodes whose box is this close are read whole
    // nearly parallel sets: d is the lines' mean direction, lines are binned where they cross the
    // reference plane at depth w0 along d, and the particles' depths lie in w0 + [dwlo, dwhi]
    double w0, dwlo, dwhi;
    // u, v, d: index of the coordinate axis each is (exactly +e_k), -1 otherwise; a projection on
    // such an axis is the coordinate itself (wave-uniform branches instead of FP64 products)
    int au, av, ad;
};

// Nearly parallel line sets (the TILT kernels): per sorted line its slopes against the mean
// direction d, (d_i.u / d_i.d, d_i.v / d_i.d), and its direction's components on u and v; per
// tile the bounds of its lines' slopes, 1 / the smallest d_i.d, and 1 + the largest |slope|.
constexpr int kTileSl = 6;
struct TiltView {
    const double4* sl = nullptr;       // (slope s, slope t, d_i.u, d_i.v) per sorted line
    const double* tb = nullptr;        // kTileSl doubles per tile
    const double* d = nullptr;         // line directions [3][nl] (exact re-checks)
};

// Tolerances of a frame from the scale E of every coordinate (twice the largest magnitude of
// the particle box and the line points): |FP64 reference d^2 - projected d^2| <= ~1e-14 E^2
// (rounding of both forms, direction normalisation within 1e-12), so a band of 1e-11 E^2 around
// (K h)^2 sends every doubtful pair to the reference formula; |x - hi| <= 2^-24 |x| <= 2^-25 E per
// coordinate, so a projection from the FP32 position alone is within sqrt(3) 2^-25 E = 5.2e-8 E
// of the exact one (pad32).
__host__ __device__ inline void frame_scale(Frame& f, double E) {
    f.band = 1e-11 * E * E;
    f.pad = 1e-12 * E;
    f.pad32 = f.pad + 2e-7 * E;
    f.sqrt_band = sqrt(f.band);
}

__device__ inline int clamp_cell(double c, int g) { return c < 0.0 ? 0 : (c > (double)(g - 1) ? g - 1 : (int)c); }
__device__ inline int cell_of(double v, double org, double inv_size, int g) { return clamp_cell(floor((v - org) * inv_size), g); }

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

// Index of the coordinate axis a is exactly (+e_k), or -1.
__host__ __device__ inline int unit_axis(const double a[3]) {
    for (int k = 0; k < 3; ++k)
        if (a[k] == 1.0 && a[(k + 1) % 3] == 0.0 && a[(k + 2) % 3] == 0.0) return k;
    return -1;
}

// project_box on an axis given by its index (ax >= 0: the box's own extent, the same values).
__device__ inline void project_axis(const Node32& b, int ax, const double a[3], double& lo, double& hi) {
    if (ax == 0) { lo = b.mnx; hi = b.mxx; }
    else if (ax == 1) { lo = b.mny; hi = b.mxy; }
    else if (ax == 2) { lo = b.mnz; hi = b.mxz; }
    else project_box(b, a, lo, hi);
}

// A position's coordinate along a (ax >= 0: the coordinate itself, the same value).
__device__ inline double dot_axis(double x, double y, double z, int ax, const double a[3]) {
    if (ax == 0) return x;
    if (ax == 1) return y;
    if (ax == 2) return z;
    return (x * a[0] + y * a[1]) + z * a[2];
}

// LDS of one line-map tile: CSR starts of the tile's cells and the sweep state (one particle
// per lane, and the owner lane of each (particle, row) item).  The accumulators and particle
// fields live in the caller's dynamic LDS (`dyn`, lines_tile_lds bytes).  The lines' (s, t)
// are read from the sorted global array (1 KB per chunk, cache-resident): a global read waits
// on the vector-memory counter, so prefetching the next line does not wait for the LDS
// accumulator atomics issued after it.
struct alignas(16) LineTileLDS {
    double w_sb[64], w_tb[64], w_r2[64], w_wih[64], w_xs[64], w_dw[64];
    int32_t cs[kCells + 1];
    int32_t w_p[64];
    int32_t pend[128];                  // particles that reach the tile, not yet swept
    uint16_t w_item[64 * kTS];          // items: lane | row << 6 | first col << 9 | last col << 12 (tile-local)
};

// Dynamic LDS of one tile: accumulators (and, for one-wavefront workgroups, the chunk's line
// coordinates) and the particle fields of a sweep.
__host__ __device__ inline size_t lines_tile_lds(int ncol, int mode, bool wg1 = false) {
    return sizeof(double) * (size_t)kCap * ((wg1 ? 2 : 0) + (mode == 0 ? 1 + ncol : 3)) +
           sizeof(float) * 64 * (size_t)(mode == 0 ? ncol : 0);
}

// One 8 x 8-cell tile of a line map (one wavefront).  `lut` is the kernel's Inorm table (in LDS
// or global memory) followed by at least one zero.  MODE 0: interpolation (FP64 weight sums).  MODE 1: neighbour digest per
// line (count, sum of sorted particle indices, sum of their squares; unsigned 64-bit, wrapping).
// WG1: the workgroup is this one wavefront (the batch kernel, whose table stays in global
// memory): the chunk's line coordinates are staged in LDS, so the table has the vector L1 to
// itself, and barriers are workgroup barriers.  Otherwise other wavefronts of the workgroup work
// on other tiles: wavefront barriers only, and each line is read from global memory one pair
// ahead (a global read waits on the vector-memory counter, not behind the LDS atomics).
// Always inlined: it is called from several kernels of one file, and an out-of-line call
// makes each caller's register allocation conservative.
// TILT: a nearly parallel set (tv): a particle at depth w0 + dw meets line i where the line's
// reference-plane point shifted by dw times its slopes lies, so the particle's image on the
// reference plane spans its position minus dw times the tile's slope range, and its footprint
// there widens by 1 / d_i.d; each pair's squared distance is |e|^2 - (e.d_i)^2 with e the offset
// from the line's point at the particle's depth.
template <int K, int MODE, bool WG1 = false, bool H2 = false, bool TILT = false>
__device__ __attribute__((always_inline)) inline void lines_tile(const T32& T, const Frame& f, const TileCat& cat, const int32_t* __restrict__ cstart,
                                  const int32_t* __restrict__ perm, const double2* __restrict__ st_sorted,
                                  const double* __restrict__ o, int64_t nl, const float* __restrict__ lut,
                                  double inv_dq, float* __restrict__ out, int64_t ldo, uint64_t* __restrict__ digest,
                                  int64_t tile, LineTileLDS& L, double* dyn, const TiltView& tv = TiltView{}) {
    int32_t* cs = L.cs;
    double *w_sb = L.w_sb, *w_tb = L.w_tb, *w_r2 = L.w_r2, *w_wih = L.w_wih, *w_xs = L.w_xs, *w_dw = L.w_dw;
    int32_t* w_p = L.w_p;
    uint16_t* w_item = L.w_item;
    const int lane = WG1 ? (int)threadIdx.x : (int)(threadIdx.x & 63);
    auto sync = [] {
        if constexpr (WG1) __syncthreads();
        else wave_sync();
    };
    const int nacc = MODE == 0 ? 1 + cat.n : 3;
    double2* ST = reinterpret_cast<double2*>(dyn);          // WG1: (s, t) of the chunk's lines
    double* acc = dyn + (WG1 ? 2 * kCap : 0);
    unsigned long long* dg = reinterpret_cast<unsigned long long*>(acc);
    float* w_A = reinterpret_cast<float*>(acc + nacc * kCap);     // [ncol][64]
    const double Kv = support(K);
    const int64_t n = T.n;

    sync();
    cs[lane] = cstart[tile * kCells + lane];
    if (lane == 0) cs[kCells] = cstart[(tile + 1) * kCells];
    sync();
    const int L0 = cs[0], L1 = cs[kCells];
    if (L0 >= L1) return;
    const int txi = (int)(tile % f.tx), tyi = (int)(tile / f.tx);
    const int cx_base = txi * kTS, cy_base = tyi * kTS;
    // tile rectangle in the plane; border tiles also hold lines clamped into the grid
    const double rs0 = cx_base == 0 ? -__builtin_inf() : f.so + (double)cx_base * f.cs;
    const double rs1 = cx_base + kTS >= f.gx ? __builtin_inf() : f.so + (double)(cx_base + kTS) * f.cs;
    const double rt0 = cy_base == 0 ? -__builtin_inf() : f.to + (double)cy_base * f.ct;
    const double rt1 = cy_base + kTS >= f.gy ? __builtin_inf() : f.to + (double)(cy_base + kTS) * f.ct;
    const double near_s = f.near * f.cs, near_t = f.near * f.ct;
    // TILT: the tile's slope bounds, footprint widening and residual factor (uniform)
    double ss_lo = 0.0, ss_hi = 0.0, st_lo = 0.0, st_hi = 0.0, icm = 1.0, padm = 1.0;
    if constexpr (TILT) {
        const double* q = tv.tb + tile * kTileSl;
        ss_lo = q[0]; ss_hi = q[1]; st_lo = q[2]; st_hi = q[3]; icm = q[4]; padm = q[5];
    }
    // image interval of a position at depth offset dw: [v - hi, v - lo] with [lo, hi] the range of
    // slope x dw over the tile's slopes (dw a point or an interval [d0, d1])
    auto shift_range = [&](double lo_sl, double hi_sl, double d0, double d1, double& mn, double& mx) {
        const double a = lo_sl * d0, b = lo_sl * d1, c = hi_sl * d0, e = hi_sl * d1;
        mn = fmin(fmin(a, b), fmin(c, e));
        mx = fmax(fmax(a, b), fmax(c, e));
    };

    for (int C0 = L0; C0 < L1; C0 += kCap) {
        const int m = min(kCap, L1 - C0);
        // the chunk's LDS arrays indexed by sorted line (one address operation per access)
        double* accC = acc - C0;
        unsigned long long* dgC = dg - C0;
        const double2* STc = ST - C0;
        for (int k = lane; k < kCap; k += 64) {
            if (WG1 && k < m) ST[k] = st_sorted[C0 + k];
            for (int a = 0; a < nacc; ++a) {
                if (MODE == 0) acc[a * kCap + k] = 0.0;
                else dg[a * kCap + k] = 0ull;
            }
        }
        sync();

        // node box projected on the plane (TILT: its image over the box's depths)
        auto image_box = [&](const Node32& nd, double& s0, double& s1, double& t0, double& t1) {
            // (the axis shortcut only for parallel sets: in the TILT kernels it measured slower,
            // 15.3 -> 17.3 ms for a camera's rays)
            if constexpr (TILT) {
                project_box(nd, f.u, s0, s1);
                project_box(nd, f.v, t0, t1);
            } else {
                project_axis(nd, f.au, f.u, s0, s1);
                project_axis(nd, f.av, f.v, t0, t1);
            }
            if constexpr (TILT) {
                double w0n, w1n, mn, mx;
                project_box(nd, f.d, w0n, w1n);
                shift_range(ss_lo, ss_hi, w0n - f.w0, w1n - f.w0, mn, mx);
                s0 -= mx;
                s1 -= mn;
                shift_range(st_lo, st_hi, w0n - f.w0, w1n - f.w0, mn, mx);
                t0 -= mx;
                t1 -= mn;
            }
        };
        // need(nd) also notes whether the node's box lies near the tile, for take(nd) on the same
        // node right after it (the walk always tests need first)
        bool near_last = false;
        auto need = [&](const Node32& nd) {
            double s0, s1, t0, t1;
            image_box(nd, s0, s1, t0, t1);
            const double gs = fmax(0.0, fmax(s0 - rs1, rs0 - s1));
            const double gt = fmax(0.0, fmax(t0 - rt1, rt0 - t1));
            near_last = gs <= near_s && gt <= near_t;
            const double R = Kv * (double)nd.hmax * icm;
            const double Rp = R + f.sqrt_band + f.pad;
            return gs * gs + gt * gt <= Rp * Rp;
        };
        // One sweep = up to 64 particles of a leaf run, one per lane.  Each lane finds the
        // rows and columns of this tile its footprint covers; the (particle, row) items of the
        // sweep are dealt round-robin to the lanes (particles with large h do not hold the
        // other lanes idle), and a row's lines are one contiguous run of sorted lines.
        auto sweep = [&](int64_t p, bool valid) {
            int nrow = 0, ncolc = 0, cell0 = 0;   // cell0: the item code of the first row
            if (valid) {
                // every load of the particle issued up front (one latency instead of three); an
                // FP32 input's position has no low part
                const float4 H = T.leaf[p];
                float4 Lo = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                if (!T.f32in) Lo = T.leaflo[p];
                double invh_l = 0.0, w64_l = 0.0;
                float A_l[2] = {0.0f, 0.0f};
                if (MODE == 0) {
                    invh_l = T.ih64[p];
                    w64_l = T.w64[p];
                    if (cat.n > 0) A_l[0] = T.pay[(int64_t)cat.fcol[0] * n + p];
                    if (cat.n > 1) A_l[1] = T.pay[(int64_t)cat.fcol[1] * n + p];
                }
                const double x = (double)H.x + (double)Lo.x, y = (double)H.y + (double)Lo.y,
                             z = (double)H.z + (double)Lo.z;
                const double sb = TILT ? (x * f.u[0] + y * f.u[1]) + z * f.u[2] : dot_axis(x, y, z, f.au, f.u);
                const double tb = TILT ? (x * f.v[0] + y * f.v[1]) + z * f.v[2] : dot_axis(x, y, z, f.av, f.v);
                // image of the position on the reference plane (TILT), else the position itself
                double dw = 0.0, s_lo = sb, s_hi = sb, t_lo = tb, t_hi = tb;
                if constexpr (TILT) {
                    dw = ((x * f.d[0] + y * f.d[1]) + z * f.d[2]) - f.w0;
                    double mn, mx;
                    shift_range(ss_lo, ss_hi, dw, dw, mn, mx);
                    s_lo = sb - mx;
                    s_hi = sb - mn;
                    shift_range(st_lo, st_hi, dw, dw, mn, mx);
                    t_lo = tb - mx;
                    t_hi = tb - mn;
                }
                // reject on the FP32 h (round-to-nearest, widened by 1e-6) before more loads
                const double Ru = Kv * (double)H.w * (1.0 + 1e-6) * icm;
                const double Rpu = Ru + f.sqrt_band + f.pad;
                if (!(s_hi + Rpu < rs0 || s_lo - Rpu > rs1 || t_hi + Rpu < rt0 || t_lo - Rpu > rt1)) {
                    const double hb = T.f32in ? (double)H.w : T.cols64[4 * n + p];
                    const double rr = Kv * hb;
                    const double r2 = rr * rr;
                    const double Rp = rr * icm + f.sqrt_band + f.pad;
                    // global cell range of the footprint, clamped like the binning, then tile-local
                    const int gx0 = clamp_cell(floor((s_lo - Rp - f.so) * f.inv_cs), f.gx);
                    const int gx1 = clamp_cell(floor((s_hi + Rp - f.so) * f.inv_cs), f.gx);
                    const int gy0 = clamp_cell(floor((t_lo - Rp - f.to) * f.inv_ct), f.gy);
                    const int gy1 = clamp_cell(floor((t_hi + Rp - f.to) * f.inv_ct), f.gy);
                    const int cx0 = max(gx0 - cx_base, 0), cx1 = min(gx1 - cx_base, kTS - 1);
                    const int cy0 = max(gy0 - cy_base, 0), cy1 = min(gy1 - cy_base, kTS - 1);
                    if (cx0 <= cx1 && cy0 <= cy1) {
                        nrow = cy1 - cy0 + 1;
                        ncolc = cx1 - cx0 + 1;
                        w_sb[lane] = sb;
                        w_tb[lane] = tb;
                        w_r2[lane] = r2;
                        if constexpr (TILT) w_dw[lane] = dw;
                        w_p[lane] = (int32_t)p;
                        cell0 = lane | cy0 << 6 | cx0 << 9 | cx1 << 12;
                        if (MODE == 0) {
                            // (m/rho) / h and 1 / (h dq) of line_integrated_kernel_function
                            w_wih[lane] = H2 ? (w64_l * invh_l) * invh_l : w64_l * invh_l;
                            w_xs[lane] = invh_l * inv_dq;
                            for (int c = 0; c < cat.n && c < 2; ++c) w_A[c * 64 + lane] = A_l[c];
                            for (int c = 2; c < cat.n; ++c) w_A[c * 64 + lane] = T.pay[(int64_t)cat.fcol[c] * n + p];
                        }
                    }
                }
            }
            // Item order: grouped by the number of covered columns (the pair loop's length),
            // longest first, so the 64 items dealt to the lanes at once have similar lengths.
            // Exclusive prefix of the item counts per group across the wave: three words of
            // 10-bit fields (groups 8 7 6 | 5 4 3 | 2 1; at most 512 items per group).
            const int gw = ncolc >= 6 ? 0 : (ncolc >= 3 ? 1 : 2);
            const int gs = gw == 0 ? 8 - ncolc : (gw == 1 ? 5 - ncolc : 2 - ncolc);
            uint32_t iw[3] = {0u, 0u, 0u};
            if (nrow > 0) iw[gw] = (uint32_t)nrow << (10 * gs);
#pragma unroll
            for (int q = 0; q < 3; ++q) iw[q] = wave_inclusive_add(iw[q]);
            uint32_t tw[3];
#pragma unroll
            for (int q = 0; q < 3; ++q) tw[q] = (uint32_t)__builtin_amdgcn_readlane((int)iw[q], 63);
            int total = 0, excl = 0;
#pragma unroll
            for (int c = 8; c >= 1; --c) {           // groups in decreasing column count
                const int cw = c >= 6 ? 0 : (c >= 3 ? 1 : 2);
                const int cs_ = cw == 0 ? 8 - c : (cw == 1 ? 5 - c : 2 - c);
                const int tot_c = (int)((tw[cw] >> (10 * cs_)) & 1023u);
                if (c == ncolc) excl = total + (int)((iw[cw] >> (10 * cs_)) & 1023u) - nrow;
                total += tot_c;
            }
            if (total == 0) return;
            for (int i = 0; i < nrow; ++i) w_item[excl + i] = (uint16_t)(cell0 + (i << 6));
            sync();
            for (int t = lane; t < total; t += 64) {
                const int itm = w_item[t];
                const int j = itm & 63, row = (itm >> 6) & 7;
                const int a = max(cs[row * kTS + ((itm >> 9) & 7)], C0);
                const int b = min(cs[row * kTS + (itm >> 12) + 1], C0 + m);
                if (a >= b) continue;
                const double sb = w_sb[j], tb = w_tb[j], r2 = w_r2[j];
                const double dwj = TILT ? w_dw[j] : 0.0;
                const double r2hi = r2 + f.band, r2lo = r2 - f.band;
                const int64_t pj = w_p[j];
                // per-item factors in registers: (m/rho) / h, 1 / (h dq), the item's fields
                double wih = 0.0, xscale = 0.0;
                float A0 = 0.0f, A1 = 0.0f;
                if (MODE == 0) {
                    wih = w_wih[j];
                    xscale = w_xs[j];
                    if (cat.n > 0) A0 = w_A[j];
                    if (cat.n > 1) A1 = w_A[64 + j];
                    // take the LDS values here: waited for at their first use inside the pair
                    // loop, the wait would also cover the accumulator atomics of every pair
                    if constexpr (!WG1) asm volatile("" ::"v"(wih), "v"(xscale), "v"(A0), "v"(A1));
                }
                // lanes on the same row start at different columns (fewer LDS bank conflicts)
                const int len = b - a;
                int k = a + (lane % len);
                // the next line's (s, t) is read one pair ahead (WG1: from LDS, in the pair), at a
                // 32-bit byte offset from the array (line sets on this path hold < 2^28 lines)
                auto st_at = [&](int kq) {
                    return *reinterpret_cast<const double2*>(reinterpret_cast<const char*>(st_sorted) + ((uint32_t)kq << 4));
                };
                double2 Lnext = WG1 ? make_double2(0.0, 0.0) : st_at(k);
                double4 Snext = make_double4(0.0, 0.0, 0.0, 0.0);
                if constexpr (TILT) Snext = tv.sl[k];
#pragma unroll 2   // two pairs per trip: no register copies of the prefetched line
                for (int it = 0; it < len; ++it) {
                    const int kc = k;                      // this pair's sorted line
                    const double2 Ln = WG1 ? STc[kc] : Lnext;
                    double4 Sl = Snext;
                    k = (k + 1 == b) ? a : k + 1;
                    if constexpr (!WG1) Lnext = st_at(k);
                    if constexpr (TILT) Snext = tv.sl[k];
                    double d2;
                    if constexpr (TILT) {
                        // offset from the line's point at the particle's depth, minus its part along
                        // the line (d_i.d > 0: e has no component along d)
                        const double es = fma(-Sl.x, dwj, sb - Ln.x), et = fma(-Sl.y, dwj, tb - Ln.y);
                        const double ed = Sl.z * es + Sl.w * et;
                        d2 = fmax((es * es + et * et) - ed * ed, 0.0);
                    } else {
                        const double ds = sb - Ln.x, dt = tb - Ln.y;
                        d2 = fma(ds, ds, dt * dt);   // within ~1e-16 d2 of the unfused form, far inside the band
                    }
                    if (d2 > r2hi) continue;
                    bool ok = d2 < r2lo;
                    if (!ok) {
                        // reference formula on the FP64 columns (_squared_distance_line_coords)
                        const int64_t li = perm[kc];
                        double px, py, pz;
                        pos64(T, pj, px, py, pz);
                        double dl0 = f.d[0], dl1 = f.d[1], dl2 = f.d[2];
                        if constexpr (TILT) {
                            dl0 = tv.d[li];
                            dl1 = tv.d[nl + li];
                            dl2 = tv.d[2 * nl + li];
                        }
                        double D2 = 0.0, Dm = 0.0, e;
                        e = px - o[li]; D2 += e * e; Dm += e * dl0;
                        e = py - o[nl + li]; D2 += e * e; Dm += e * dl1;
                        e = pz - o[2 * nl + li]; D2 += e * e; Dm += e * dl2;
                        ok = D2 - Dm * Dm <= r2;
                    }
                    if (!ok) continue;
                    if (MODE == 0) {
                        // line_integrated_kernel_function: invh * Inorm(r invh).  r from the
                        // hardware FP64 reciprocal square root (relative error ~2^-23) and one
                        // Newton step (~1e-14; d2 + 1e-300 equals d2 for any d2 above ~1e-284,
                        // and a particle on the line gets r ~ 1e-150, i.e. Inorm(0)), q and the
                        // table interpolation in FP64 on FP32 table values: FP32 q would carry
                        // ~1e-7 relative error, amplified by dlnI/dlnq.  The table ends with
                        // Inorm(K) = 0 followed by zeros (kLutF), so a clamped index needs no
                        // test of q >= K: there both table values are 0.
                        const double dd = d2 + 1e-300;
                        const double y = __builtin_amdgcn_rsq(dd);
                        const double r0 = dd * y;
                        const double r = fma(fma(-r0, r0, dd), 0.5 * y, r0);
                        const double x = fmin(r * xscale, (double)(kLutN - 1));   // q / dq
                        const int i = (int)x;
                        const double ft = __builtin_amdgcn_fract(x);
                        float2 Lv;                                 // lut[i], lut[i + 1]
                        __builtin_memcpy(&Lv, lut + i, sizeof(Lv));
                        const double I0 = (double)Lv.x;
                        const double I = fma(ft, (double)Lv.y - I0, I0);
                        const double w = wih * I;
                        atomicAdd(&accC[kc], w);
                        if (cat.n > 0) atomicAdd(&accC[kCap + kc], (double)A0 * w);
                        if (cat.n > 1) atomicAdd(&accC[2 * kCap + kc], (double)A1 * w);
                        for (int c = 2; c < cat.n; ++c)
                            atomicAdd(&accC[(1 + c) * kCap + kc], (double)w_A[c * 64 + j] * w);
                    } else {
                        const unsigned long long ip = (unsigned long long)pj;
                        atomicAdd(&dgC[kc], 1ull);
                        atomicAdd(&dgC[kCap + kc], ip);
                        atomicAdd(&dgC[2 * kCap + kc], ip * ip);
                    }
                }
            }
            sync();
        };
        // Runs of leaves (contiguous in Morton order): the particles whose footprint box can
        // reach the tile are queued, and every 64 queued particles make one sweep, so that the
        // sweeps stay full where most of a run misses the tile.
        int npend = 0;                 // wave-uniform
        // (on the FP32 position alone: its residual moves the projection by less than pad32)
        // (TILT: the image's residual is at most (1 + max |slope|) times a projection's)
        auto reaches = [&](const float4& H) {
            const double x = (double)H.x, y = (double)H.y, z = (double)H.z;
            const double sb = TILT ? (x * f.u[0] + y * f.u[1]) + z * f.u[2] : dot_axis(x, y, z, f.au, f.u);
            const double tb = TILT ? (x * f.v[0] + y * f.v[1]) + z * f.v[2] : dot_axis(x, y, z, f.av, f.v);
            double s_lo = sb, s_hi = sb, t_lo = tb, t_hi = tb;
            if constexpr (TILT) {
                const double dw = ((x * f.d[0] + y * f.d[1]) + z * f.d[2]) - f.w0;
                double mn, mx;
                shift_range(ss_lo, ss_hi, dw, dw, mn, mx);
                s_lo = sb - mx;
                s_hi = sb - mn;
                shift_range(st_lo, st_hi, dw, dw, mn, mx);
                t_lo = tb - mx;
                t_hi = tb - mn;
            }
            const double Rpu = Kv * (double)H.w * (1.0 + 1e-6) * icm + f.sqrt_band + f.pad32 * padm;
            return !(s_hi + Rpu < rs0 || s_lo - Rpu > rs1 || t_hi + Rpu < rt0 || t_lo - Rpu > rt1);
        };
        auto flush = [&](int cnt) {
            sync();
            const int q = L.pend[lane];
            const int rest = npend - cnt;
            const int moved = lane < rest ? L.pend[cnt + lane] : 0;
            sync();
            if (lane < rest) L.pend[lane] = moved;
            npend = rest;
            sweep((int64_t)q, lane < cnt);
        };
        // (the next 64 particles' positions are loaded while the current ones are tested)
        auto bucket = [&](int lo, int cnt) {
            float4 Hn = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (lane < cnt) Hn = T.leaf[(int64_t)lo + lane];
            for (int base = 0; base < cnt; base += 64) {
                const float4 H = Hn;
                if (base + 64 + lane < cnt) Hn = T.leaf[(int64_t)lo + base + 64 + lane];
                const bool hit = base + lane < cnt && reaches(H);
                const uint64_t mk = ballot(hit);
                if (hit)
                    L.pend[npend + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(mk >> 32),
                                                                  __builtin_amdgcn_mbcnt_lo((uint32_t)mk, 0u))] =
                        lo + base + lane;
                npend += __builtin_popcountll(mk);
                if (npend >= 64) flush(64);
            }
        };
        // A node is taken whole (up to T.bucket leaves) when its box itself lies near the
        // tile; a node that reaches the tile only through a large hmax is descended down to
        // one sweep of 64, so that only the subtrees holding the large-h particles are read.
        auto take = [&](const Node32&, int cnt) {
            if (cnt <= 64) return true;
            if (cnt > T.bucket) return false;
            return near_last;
        };
        walk(T, true, need, bucket, take);
        if (npend > 0) flush(npend);
        sync();
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

// A binned parallel line set ready for lines_tile (kept by its context).
struct LinePlan {
    Frame f;
    const int32_t* cstart;
    const int32_t* perm;
    const double2* sts;
    double amax;          // largest magnitude of the lines' points (the frame's scale with the particles)
    bool tilt;            // nearly parallel set (tv)
    TiltView tv;
};

}  // namespace lmap

// Bin a parallel unit-direction line set (o, d packed [3][nl] on the device); *ok = false when
// the lines do not qualify.  Synchronises the context stream.
int lines_plan(pmx_ctx* c, const double* o, const double* d, int64_t nl, lmap::LinePlan* plan, bool* ok);

}  // namespace pmx
