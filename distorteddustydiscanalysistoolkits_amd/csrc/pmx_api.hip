// pmx_api.hip — the extern "C" entry points of libpartia_mi355x (declared in
// include/partia_mi355x.h): argument validation with the reference's error classes,
// host/device staging, catalog chunking and output delivery.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <memory>
#include <new>
#include <vector>

#include "pmx_internal.h"

namespace pmx {

int set_error(pmx_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    return code;
}

int set_error(pmx_ctx* c, int code, const char* msg) {
    if (c) c->err.assign(msg);
    return code;
}

void note_exception(pmx_ctx* c, const char* msg) noexcept {
    if (!c) return;
    try {
        c->err = msg;
    } catch (...) {
    }
}
void note_exception(const pmx_ctx* c, const char* msg) noexcept { note_exception(const_cast<pmx_ctx*>(c), msg); }
void note_exception(std::string* s, const char* msg) noexcept {
    if (!s) return;
    try {
        *s = msg;
    } catch (...) {
    }
}

void throw_test_fault() {
    const char* v = std::getenv("PMX_TEST_THROW");
    if (!v || !*v) return;
    if (std::strcmp(v, "bad_alloc") == 0) throw std::bad_alloc();
    if (std::strcmp(v, "exception") == 0) throw std::runtime_error("PMX_TEST_THROW: injected exception");
}

// Knobs of a new context (see Tuning); out-of-range values are clamped.
void read_tuning(Tuning& t) {
    if (const char* v = std::getenv("PMX_LINES_TILED")) t.lines_tiled = std::atoi(v) != 0;
    if (const char* v = std::getenv("PMX_BUCKET_POINTS")) t.bucket_points = std::max(kMinRun, std::min(64, std::atoi(v)));
    if (const char* v = std::getenv("PMX_BUCKET_LINES")) t.bucket_lines = std::max(kMinRun, std::min(64, std::atoi(v)));
    if (const char* v = std::getenv("PMX_TILED_BUCKET")) t.tiled_bucket = std::max(1, std::min(1 << 16, std::atoi(v)));
    if (const char* v = std::getenv("PMX_TILED_NEAR")) t.tiled_near = std::max(0, std::min(1 << 20, std::atoi(v)));
    if (const char* v = std::getenv("PMX_REFIT_SMALL")) t.refit_small = std::max(1, std::min(1024, std::atoi(v)));
}

int hip_fail(pmx_ctx* c, hipError_t e, const char* what) {
    char buf[512];
    snprintf(buf, sizeof(buf), "HIP error %d (%s) in %s", (int)e, hipGetErrorString(e), what);
    return set_error(c, e == hipErrorOutOfMemory ? PMX_ERR_NOMEM : PMX_ERR_HIP, buf);
}

int ensure(pmx_ctx* c, DevBuf& b, size_t bytes) {
    if (bytes == 0) bytes = 8;
    if (b.cap >= bytes) return PMX_OK;
    if (b.p) {
        hipError_t e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) return hip_fail(c, e, "hipStreamSynchronize");
        (void)hipFree(b.p);
        b.p = nullptr;
        b.cap = 0;
    }
    size_t want = bytes + bytes / 8;
    hipError_t e = hipMalloc(&b.p, want);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        b.p = nullptr;
        return set_error(c, PMX_ERR_NOMEM, "device allocation of " + std::to_string(want) + " bytes failed");
    }
    b.cap = want;
    return PMX_OK;
}

namespace {

constexpr int kBlock = 256;

__global__ void k_f64_to_f32(const double* __restrict__ in, float* __restrict__ out, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
        out[i] = (float)in[i];
}
// coordinate_grid(Cartesian, grid) (StructuredGrid.jl:126-141): point i = ix + nx (iy + ny iz)
// takes (ax[ix], ay[iy], az[iz]) — copies, no arithmetic on the values.
__global__ void k_cartesian_points(const double* __restrict__ ax, const double* __restrict__ ay,
                                   const double* __restrict__ az, int64_t nx, int64_t ny, int64_t npts,
                                   double* __restrict__ x, double* __restrict__ y, double* __restrict__ z) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < npts; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t ix = i % nx, r = i / nx;
        const int64_t iy = r % ny, iz = r / ny;
        x[i] = ax[ix];
        y[i] = ay[iy];
        z[i] = az[iz];
    }
}

__global__ void k_f32_to_f64(const float* __restrict__ in, double* __restrict__ out, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
        out[i] = (double)in[i];
}

int check_catalog(pmx_ctx* c, const pmx_catalog* cat, bool lines, FullCat& fc) {
    if (!cat) return set_error(c, PMX_ERR_ARGUMENT, "catalog: null pointer");
    const int ncol = c->tree.ncol;
    if (cat->n_scalar < 0 || cat->n_grad < 0 || cat->n_div < 0 || cat->n_curl < 0)
        return set_error(c, PMX_ERR_DIMENSION, "catalog: negative count");
    if (lines && (cat->n_grad || cat->n_div || cat->n_curl))
        return set_error(c, PMX_ERR_ARGUMENT,
                         "LineSamples_interpolation: catalog must hold scalars only (InterpolationCatalog{3,N,0,0,0,N})");
    auto slot_ok = [&](int s) { return s >= 0 && s < ncol; };
    for (int j = 0; j < cat->n_scalar; ++j) {
        int s = cat->scalar_slots[j];
        if (!slot_ok(s)) return set_error(c, PMX_ERR_DIMENSION, "catalog: scalar slot out of range");
        fc.ss.push_back(s);
        fc.sh.push_back(cat->scalar_shepard ? (cat->scalar_shepard[j] ? 1 : 0) : 1);
    }
    for (int j = 0; j < cat->n_grad; ++j) {
        int s = cat->grad_slots[j];
        if (s != -1 && !slot_ok(s)) return set_error(c, PMX_ERR_DIMENSION, "catalog: gradient slot out of range");
        fc.gs.push_back(s);
    }
    for (int j = 0; j < 3 * cat->n_div; ++j) {
        int s = cat->div_slots[j];
        if (!slot_ok(s)) return set_error(c, PMX_ERR_DIMENSION, "catalog: divergence slot out of range");
        fc.ds.push_back(s);
    }
    for (int j = 0; j < 3 * cat->n_curl; ++j) {
        int s = cat->curl_slots[j];
        if (!slot_ok(s)) return set_error(c, PMX_ERR_DIMENSION, "catalog: curl slot out of range");
        fc.cs.push_back(s);
    }
    return PMX_OK;
}

}  // namespace

// Split a catalog into chunks the fixed-size kernels hold (<= 4 scalars, 2 gradients,
// 2 divergences, 2 curls each); every chunk writes its outputs at their catalog positions.
std::vector<CatChunk> make_chunks(const FullCat& fc) {
    const int N = (int)fc.ss.size(), G = (int)fc.gs.size(), D = (int)fc.ds.size() / 3, Cc = (int)fc.cs.size() / 3;
    std::vector<CatChunk> out;
    for (int k = 0;; ++k) {
        CatChunk ch{};
        ch.ns = std::max(0, std::min(4, N - 4 * k));
        ch.ng = std::max(0, std::min(2, G - 2 * k));
        ch.nd = std::max(0, std::min(2, D - 2 * k));
        ch.nc = std::max(0, std::min(2, Cc - 2 * k));
        if (k > 0 && ch.ns + ch.ng + ch.nd + ch.nc == 0) break;
        for (int j = 0; j < ch.ns; ++j) { ch.ss[j] = fc.ss[4 * k + j]; ch.sh[j] = fc.sh[4 * k + j]; }
        for (int j = 0; j < ch.ng; ++j) ch.gs[j] = fc.gs[2 * k + j];
        for (int j = 0; j < 3 * ch.nd; ++j) ch.ds[j] = fc.ds[6 * k + j];
        for (int j = 0; j < 3 * ch.nc; ++j) ch.cs[j] = fc.cs[6 * k + j];
        ch.os = 4 * k;
        ch.og = N + 3 * 2 * k;
        ch.od = N + 3 * G + 2 * k;
        ch.oc = N + 3 * G + D + 3 * 2 * k;
        out.push_back(ch);
        if (k == 0 && N + G + D + Cc == 0) break;
    }
    return out;
}

namespace {

int check_points(pmx_ctx* c, const pmx_points* pts) {
    if (!pts) return set_error(c, PMX_ERR_ARGUMENT, "points: null pointer");
    if (pts->n < 0) return set_error(c, PMX_ERR_DIMENSION, "points: negative count");
    if (pts->n > 0 && (!pts->x || !pts->y || !pts->z)) return set_error(c, PMX_ERR_ARGUMENT, "points: null coordinate");
    return PMX_OK;
}

int check_lines(pmx_ctx* c, const pmx_lines* l) {
    if (!l) return set_error(c, PMX_ERR_ARGUMENT, "lines: null pointer");
    if (l->n < 0) return set_error(c, PMX_ERR_DIMENSION, "lines: negative count");
    if (l->n > 0 && (!l->ox || !l->oy || !l->oz || !l->dx || !l->dy || !l->dz))
        return set_error(c, PMX_ERR_ARGUMENT, "lines: null pointer");
    return PMX_OK;
}

// Query coordinates on the device: host arrays are copied to qbuf.
int stage_points(pmx_ctx* c, const pmx_points* pts, const double** qx, const double** qy, const double** qz) {
    const int64_t n = pts->n;
    if (pts->mem == PMX_MEM_DEVICE) {
        *qx = pts->x; *qy = pts->y; *qz = pts->z;
        return PMX_OK;
    }
    PMX_TRY(ensure(c, c->qbuf, 24 * (size_t)std::max<int64_t>(n, 1)));
    double* q = c->qbuf.as<double>();
    PMX_HIP(c, hipMemcpyAsync(q, pts->x, 8 * (size_t)n, hipMemcpyHostToDevice, c->stream));
    PMX_HIP(c, hipMemcpyAsync(q + n, pts->y, 8 * (size_t)n, hipMemcpyHostToDevice, c->stream));
    PMX_HIP(c, hipMemcpyAsync(q + 2 * n, pts->z, 8 * (size_t)n, hipMemcpyHostToDevice, c->stream));
    *qx = q; *qy = q + n; *qz = q + 2 * n;
    return PMX_OK;
}

// Lines as two packed device arrays o[3n], d[3n].
int stage_lines(pmx_ctx* c, const pmx_lines* l, const double** o, const double** d) {
    const int64_t n = l->n;
    PMX_TRY(ensure(c, c->qbuf, 48 * (size_t)std::max<int64_t>(n, 1)));
    double* q = c->qbuf.as<double>();
    const double* src[6] = {l->ox, l->oy, l->oz, l->dx, l->dy, l->dz};
    hipMemcpyKind kind = l->mem == PMX_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    for (int j = 0; j < 6; ++j) PMX_HIP(c, hipMemcpyAsync(q + j * n, src[j], 8 * (size_t)n, kind, c->stream));
    *o = q;
    *d = q + 3 * n;
    return PMX_OK;
}

// Copy L columns of `n` results (FP64 or FP32 in `buf`, column stride n) to the caller.
int deliver(pmx_ctx* c, const void* buf, bool buf_is_f32, int L, int64_t n, const pmx_output* out) {
    if (!out || (L > 0 && !out->cols)) return set_error(c, PMX_ERR_ARGUMENT, "output: null pointer");
    if (out->dtype != PMX_F32 && out->dtype != PMX_F64) return set_error(c, PMX_ERR_ARGUMENT, "output: dtype");
    for (int j = 0; j < L; ++j)
        if (!out->cols[j] && n > 0) return set_error(c, PMX_ERR_ARGUMENT, "output: null column");
    const bool want_f32 = out->dtype == PMX_F32;
    const size_t esz = want_f32 ? 4 : 8;
    const void* src = buf;
    if (want_f32 != buf_is_f32) {
        PMX_TRY(ensure(c, c->scratch_c, esz * (size_t)n * std::max(L, 1)));
        const dim3 g(grid_for((int64_t)L * n, kBlock, 16384)), b(kBlock);
        if (want_f32)
            hipLaunchKernelGGL(k_f64_to_f32, g, b, 0, c->stream, (const double*)buf, c->scratch_c.as<float>(), (int64_t)L * n);
        else
            hipLaunchKernelGGL(k_f32_to_f64, g, b, 0, c->stream, (const float*)buf, c->scratch_c.as<double>(), (int64_t)L * n);
        src = c->scratch_c.p;
    }
    hipMemcpyKind kind = out->mem == PMX_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    for (int j = 0; j < L; ++j)
        PMX_HIP(c, hipMemcpyAsync(out->cols[j], (const char*)src + esz * (size_t)n * j, esz * (size_t)n, kind, c->stream));
    return PMX_OK;
}

void upload_tables(pmx_ctx* c, std::vector<double>& h64) {
    h64.resize(6 * kLutN);
    std::vector<double> Q(kLutN);
    for (int k = 0; k < 6; ++k) {
        line_table(k, Q.data(), h64.data() + k * kLutN);
        c->lut_dq[k] = Q[1] - Q[0];
    }
}

int require_tree(pmx_ctx* c) {
    if (!c) return PMX_ERR_ARGUMENT;
    if (!c->tree.built) return set_error(c, PMX_ERR_STATE, "no LBVH built in this context (call pmx_build_lbvh)");
    return PMX_OK;
}

}  // namespace
}  // namespace pmx

using namespace pmx;

extern "C" {

const char* pmx_version(void) { return "partia_mi355x 0.1 (gfx950)"; }

int pmx_create(pmx_ctx** out, int32_t device_ordinal) {
    return guard(nullptr, [&]() -> int {
        if (!out) return PMX_ERR_ARGUMENT;
        *out = nullptr;
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
            (void)hipGetLastError();
            return PMX_ERR_HIP;
        }
        if (device_ordinal < 0 || device_ordinal >= ndev) return PMX_ERR_ARGUMENT;
        // owned by `hold` until it is handed out, so an exception on the way frees it
        std::unique_ptr<pmx_ctx, void (*)(pmx_ctx*)> hold(new pmx_ctx(), pmx_destroy);
        pmx_ctx* c = hold.get();
        c->device = device_ordinal;
        // error texts fit in the reserved capacity: setting one does not allocate
        c->err.reserve(1024);
        read_tuning(c->tune);
        {
            hipDeviceProp_t prop;
            if (hipGetDeviceProperties(&prop, device_ordinal) == hipSuccess) c->ncu = std::max(1, prop.multiProcessorCount);
        }
        if (const char* v = std::getenv("PMX_SPLIT_BUDGET")) c->split_budget = std::max(0, std::atoi(v));
        if (const char* v = std::getenv("PMX_SPLIT_ITEM")) c->split_item = std::max(64, std::atoi(v));
        if (hipSetDevice(device_ordinal) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
            return PMX_ERR_HIP;
        for (auto& e : c->ev)
            if (hipEventCreate(&e) != hipSuccess) return PMX_ERR_HIP;
        if (noalloc_init_device(c) != PMX_OK) return PMX_ERR_HIP;
        std::vector<double> h64;
        upload_tables(c, h64);
        std::vector<float> hf((size_t)6 * kLutF, 0.0f);
        for (int k = 0; k < 6; ++k)
            for (int i = 0; i < kLutN; ++i) hf[(size_t)k * kLutF + i] = (float)h64[(size_t)k * kLutN + i];
        if (ensure(c, c->lut64, 8 * h64.size()) != PMX_OK || ensure(c, c->lutf, 4 * hf.size()) != PMX_OK ||
            hipMemcpy(c->lut64.p, h64.data(), 8 * h64.size(), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(c->lutf.p, hf.data(), 4 * hf.size(), hipMemcpyHostToDevice) != hipSuccess)
            return PMX_ERR_HIP;
        *out = hold.release();
        return PMX_OK;
    });
}

void pmx_destroy(pmx_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    DevBuf* bufs[] = {&c->lut64, &c->lutf, &c->codes, &c->order, &c->keys_tmp, &c->vals_tmp, &c->cols64,
                      &c->leaf64, &c->leaf_esc, &c->nodes64, &c->nodes32, &c->leaf32, &c->pay32,
                      &c->lr, &c->parent, &c->escape, &c->visited, &c->unread, &c->range_far, &c->scratch_a,
                      &c->scratch_b, &c->scratch_c, &c->hist, &c->staging, &c->outbuf, &c->qbuf, &c->misc,
                      &c->tl_st, &c->tl_part, &c->tl_cells, &c->tl_sl, &c->tl_tsl, &c->records, &c->split, &c->gridpts};
    for (DevBuf* b : bufs)
        if (b->p) (void)hipFree(b->p);
    for (auto& e : c->ev)
        if (e) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

const char* pmx_last_error(const pmx_ctx* c) { return c ? c->err.c_str() : "null context"; }

int pmx_build_lbvh(pmx_ctx* c, const pmx_particles* p, int32_t flags, uint64_t* codes_out, int64_t* order_out) {
    return guard(c, [&]() -> int {
        if (!c) return PMX_ERR_ARGUMENT;
        PMX_HIP(c, hipSetDevice(c->device));
        return build_lbvh(c, p, flags, codes_out, order_out);
    });
}

int pmx_interpolate_cartesian_grid(pmx_ctx* c, int32_t strategy, int32_t precision, const pmx_catalog* cat,
                                   const pmx_axes* axes, const pmx_output* out) {
    return guard(c, [&]() -> int {
        if (!c) return PMX_ERR_ARGUMENT;
        if (!axes) return set_error(c, PMX_ERR_ARGUMENT, "axes: null pointer");
        if (axes->mem != PMX_MEM_HOST && axes->mem != PMX_MEM_DEVICE) return set_error(c, PMX_ERR_ARGUMENT, "axes: mem");
        int64_t npts = 1;
        for (int k = 0; k < 3; ++k) {
            if (axes->n[k] < 1) return set_error(c, PMX_ERR_DIMENSION, "axes: every axis needs at least one value");
            if (!axes->a[k]) return set_error(c, PMX_ERR_ARGUMENT, "axes: null axis pointer");
            if (npts > ((int64_t)1 << 40) / axes->n[k]) return set_error(c, PMX_ERR_ARGUMENT, "axes: grid too large");
            npts *= axes->n[k];
        }
        PMX_TRY(require_tree(c));
        PMX_HIP(c, hipSetDevice(c->device));
        const int64_t naxes = axes->n[0] + axes->n[1] + axes->n[2];
        PMX_TRY(ensure(c, c->gridpts, 8 * (size_t)(3 * npts + naxes)));
        double* xyz = c->gridpts.as<double>();
        const double* dax[3];
        double* tail = xyz + 3 * npts;
        for (int k = 0; k < 3; ++k) {
            if (axes->mem == PMX_MEM_DEVICE) {
                dax[k] = axes->a[k];
            } else {
                PMX_HIP(c, hipMemcpyAsync(tail, axes->a[k], 8 * (size_t)axes->n[k], hipMemcpyHostToDevice, c->stream));
                dax[k] = tail;
                tail += axes->n[k];
            }
        }
        hipLaunchKernelGGL(k_cartesian_points, dim3(grid_for(npts, 256, 16384)), dim3(256), 0, c->stream, dax[0], dax[1],
                           dax[2], axes->n[0], axes->n[1], npts, xyz, xyz + npts, xyz + 2 * npts);
        PMX_HIP(c, hipGetLastError());
        pmx_points pts;
        pts.mem = PMX_MEM_DEVICE;
        pts.n = npts;
        pts.x = xyz;
        pts.y = xyz + npts;
        pts.z = xyz + 2 * npts;
        return pmx_interpolate_points(c, strategy, precision, cat, &pts, out);
    });
}

int pmx_interpolate_points(pmx_ctx* c, int32_t strategy, int32_t precision, const pmx_catalog* cat,
                           const pmx_points* pts, const pmx_output* out) {
    return guard(c, [&]() -> int {
        PMX_TRY(require_tree(c));
        PMX_HIP(c, hipSetDevice(c->device));
        if (strategy < 0 || strategy > 2) return set_error(c, PMX_ERR_ARGUMENT, "unknown interpolation strategy");
        if (precision != PMX_PREC_F64 && precision != PMX_PREC_F32C) return set_error(c, PMX_ERR_ARGUMENT, "unknown precision");
        FullCat fc;
        PMX_TRY(check_catalog(c, cat, false, fc));
        PMX_TRY(check_points(c, pts));
        const int L = fc.L();
        const int64_t n = pts->n;
        if (n == 0) return PMX_OK;
        const double *qx, *qy, *qz;
        PMX_HIP(c, hipEventRecord(c->ev[4], c->stream));
        PMX_TRY(stage_points(c, pts, &qx, &qy, &qz));
        std::vector<CatChunk> chunks = make_chunks(fc);
        const bool f32 = precision == PMX_PREC_F32C;
        PMX_TRY(ensure(c, c->outbuf, (f32 ? 4 : 8) * (size_t)n * std::max(L, 1)));
        PMX_HIP(c, hipEventRecord(c->ev[5], c->stream));
        if (f32) {
            PMX_TRY(prepare_f32(c));
            int32_t* perm = nullptr;
            PMX_TRY(sort_queries(c, qx, qy, qz, n, &perm));
            PMX_HIP(c, hipEventRecord(c->ev[8], c->stream));
            for (const CatChunk& ch : chunks)
                PMX_TRY(interp_points_f32c(c, strategy, ch, perm, qx, qy, qz, n, c->outbuf.as<float>(), n));
            c->timings[8] = c->timings[9] = 0.0;
            if (strategy != kScatter && c->split_budget > 0 && c->split.p) {
                // split counters of the last pass (queued items, slots taken), capped as the kernels cap them
                unsigned int cnt[4] = {0, 0, 0, 0};
                PMX_HIP(c, hipMemcpyAsync(cnt, c->split.p, sizeof(cnt), hipMemcpyDeviceToHost, c->stream));
                PMX_HIP(c, hipStreamSynchronize(c->stream));
                c->timings[8] = std::min<double>(cnt[1], std::min<int64_t>(kSplitSlots, (n + 63) / 64));
                c->timings[9] = std::min<double>(cnt[0], kSplitItems);
            }
        } else {
            PMX_HIP(c, hipEventRecord(c->ev[8], c->stream));
            // ha from the first chunk is reused by later chunks (same nearest-h search result)
            double* ha = nullptr;
            if (chunks.size() > 1 && strategy != kScatter) {
                PMX_TRY(ensure(c, c->scratch_a, 8 * (size_t)n));
                ha = c->scratch_a.as<double>();
            }
            for (size_t k = 0; k < chunks.size(); ++k)
                PMX_TRY(interp_points_f64(c, strategy, chunks[k], qx, qy, qz, n, c->outbuf.as<double>(), n,
                                          k > 0 ? ha : nullptr, k == 0 ? ha : nullptr));
        }
        PMX_HIP(c, hipEventRecord(c->ev[6], c->stream));
        PMX_TRY(deliver(c, c->outbuf.p, f32, L, n, out));
        PMX_HIP(c, hipEventRecord(c->ev[7], c->stream));
        PMX_HIP(c, hipStreamSynchronize(c->stream));
        float a = 0, b = 0, k = 0;
        PMX_HIP(c, hipEventElapsedTime(&a, c->ev[5], c->ev[6]));
        PMX_HIP(c, hipEventElapsedTime(&b, c->ev[4], c->ev[7]));
        PMX_HIP(c, hipEventElapsedTime(&k, c->ev[8], c->ev[6]));
        c->timings[7] = k;
        c->timings[4] = 0.0;
        c->timings[5] = a;
        c->timings[6] = b;
        return PMX_OK;
    });
}

int pmx_interpolate_lines(pmx_ctx* c, int32_t strategy, int32_t precision, const pmx_catalog* cat,
                          const pmx_lines* lines, const pmx_output* out) {
    return guard(c, [&]() -> int {
        PMX_TRY(require_tree(c));
        PMX_HIP(c, hipSetDevice(c->device));
        if (strategy != PMX_SCATTER)
            return set_error(c, PMX_ERR_ARGUMENT,
                             "LineSamples_interpolation only supports itpScatter. Line-integrated samples do not have a "
                             "well-defined query smoothing length ha, so itpGather and itpSymmetric are not supported.");
        c->line_inv_h2 = (precision & PMX_LINE_INV_H2) != 0;
        precision &= ~PMX_LINE_INV_H2;
        if (precision != PMX_PREC_F64 && precision != PMX_PREC_F32C) return set_error(c, PMX_ERR_ARGUMENT, "unknown precision");
        FullCat fc;
        PMX_TRY(check_catalog(c, cat, true, fc));
        PMX_TRY(check_lines(c, lines));
        const int L = fc.L();
        const int64_t n = lines->n;
        if (n == 0) return PMX_OK;
        PMX_HIP(c, hipEventRecord(c->ev[4], c->stream));
        const double *o, *d;
        PMX_TRY(stage_lines(c, lines, &o, &d));
        const bool f32 = precision == PMX_PREC_F32C;
        PMX_TRY(ensure(c, c->outbuf, (f32 ? 4 : 8) * (size_t)n * std::max(L, 1)));
        PMX_HIP(c, hipEventRecord(c->ev[5], c->stream));
        if (f32) {
            PMX_TRY(prepare_f32(c));
            bool tiled = false;
            PMX_TRY(interp_lines_tiled(c, fc.ss.data(), fc.sh.data(), L, o, d, n, c->outbuf.as<float>(), n, &tiled));
            if (!tiled) {
                int32_t* perm = nullptr;
                PMX_TRY(sort_queries(c, o, o + n, o + 2 * n, n, &perm));
                PMX_HIP(c, hipEventRecord(c->ev[8], c->stream));
                PMX_TRY(interp_lines_f32c(c, fc.ss.data(), fc.sh.data(), L, perm, o, d, n, c->outbuf.as<float>(), n));
            }
        } else {
            PMX_HIP(c, hipEventRecord(c->ev[8], c->stream));
            PMX_TRY(interp_lines_f64(c, strategy, fc.ss.data(), fc.sh.data(), L, o, d, n, c->outbuf.as<double>(), n));
        }
        PMX_HIP(c, hipEventRecord(c->ev[6], c->stream));
        PMX_TRY(deliver(c, c->outbuf.p, f32, L, n, out));
        PMX_HIP(c, hipEventRecord(c->ev[7], c->stream));
        PMX_HIP(c, hipStreamSynchronize(c->stream));
        float a = 0, b = 0, k = 0;
        PMX_HIP(c, hipEventElapsedTime(&a, c->ev[5], c->ev[6]));
        PMX_HIP(c, hipEventElapsedTime(&b, c->ev[4], c->ev[7]));
        PMX_HIP(c, hipEventElapsedTime(&k, c->ev[8], c->ev[6]));
        c->timings[7] = k;
        c->timings[5] = a;
        c->timings[6] = b;
        return PMX_OK;
    });
}

int pmx_point_samples_interpolation(pmx_ctx* c, const pmx_particles* p, int32_t strategy, int32_t precision,
                                    const pmx_catalog* cat, const pmx_points* pts, const pmx_output* out) {
    return guard(c, [&]() -> int {
        if (!c) return PMX_ERR_ARGUMENT;
        if (strategy < 0 || strategy > 2) return set_error(c, PMX_ERR_ARGUMENT, "unknown interpolation strategy");
        PMX_TRY(pmx_build_lbvh(c, p, PMX_BUILD_WRITEBACK, nullptr, nullptr));
        return pmx_interpolate_points(c, strategy, precision, cat, pts, out);
    });
}

int pmx_line_samples_interpolation(pmx_ctx* c, const pmx_particles* p, int32_t strategy, int32_t precision,
                                   const pmx_catalog* cat, const pmx_lines* lines, const pmx_output* out) {
    return guard(c, [&]() -> int {
        if (!c) return PMX_ERR_ARGUMENT;
        if (strategy != PMX_SCATTER)
            return set_error(c, PMX_ERR_ARGUMENT,
                             "LineSamples_interpolation only supports itpScatter. Line-integrated samples do not have a "
                             "well-defined query smoothing length ha, so itpGather and itpSymmetric are not supported.");
        PMX_TRY(pmx_build_lbvh(c, p, PMX_BUILD_WRITEBACK, nullptr, nullptr));
        return pmx_interpolate_lines(c, strategy, precision, cat, lines, out);
    });
}

int pmx_export_tree(pmx_ctx* c, int32_t* left, int32_t* right, int32_t* parent, int32_t* escape) {
    return guard(c, [&]() -> int {
        PMX_TRY(require_tree(c));
        PMX_HIP(c, hipSetDevice(c->device));
        const int64_t n = c->tree.n, tot = 2 * n - 1;
        std::vector<int32_t> lr((size_t)(4 * n + 4)), pa((size_t)(2 * n + 2)), es((size_t)(2 * n + 2)), le((size_t)n);
        PMX_HIP(c, hipStreamSynchronize(c->stream));
        PMX_HIP(c, hipMemcpy(lr.data(), c->lr.p, 4 * lr.size(), hipMemcpyDeviceToHost));
        PMX_HIP(c, hipMemcpy(pa.data(), c->parent.p, 4 * pa.size(), hipMemcpyDeviceToHost));
        PMX_HIP(c, hipMemcpy(es.data(), c->escape.p, 4 * es.size(), hipMemcpyDeviceToHost));
        PMX_HIP(c, hipMemcpy(le.data(), c->leaf_esc.p, 4 * le.size(), hipMemcpyDeviceToHost));
        // ids 1 .. n-1 are internal nodes, n .. 2n-1 the leaves (no children; escapes in leaf_esc)
        for (int64_t id = 1; id <= tot; ++id) {
            const bool leaf = id >= n;
            if (left) left[id - 1] = leaf ? 0 : lr[(size_t)(2 * id)];
            if (right) right[id - 1] = leaf ? 0 : lr[(size_t)(2 * id + 1)];
            if (parent) parent[id - 1] = pa[(size_t)id];
            if (escape) escape[id - 1] = leaf ? le[(size_t)(id - n)] : es[(size_t)id];
        }
        return PMX_OK;
    });
}

int pmx_export_nodes(pmx_ctx* c, double* node_min, double* node_max, double* node_hmax) {
    return guard(c, [&]() -> int {
        PMX_TRY(require_tree(c));
        PMX_HIP(c, hipSetDevice(c->device));
        const int64_t n = c->tree.n;
        if (n < 2) return PMX_OK;
        PMX_TRY(ensure_leaf64(c));                  // the FP64 nodes are refitted on first use
        std::vector<Node64> nodes((size_t)n);
        PMX_HIP(c, hipStreamSynchronize(c->stream));
        PMX_HIP(c, hipMemcpy(nodes.data(), c->nodes64.p, sizeof(Node64) * (size_t)n, hipMemcpyDeviceToHost));
        for (int64_t i = 1; i < n; ++i) {
            for (int k = 0; k < 3; ++k) {
                if (node_min) node_min[k * (n - 1) + (i - 1)] = nodes[(size_t)i].mn[k];
                if (node_max) node_max[k * (n - 1) + (i - 1)] = nodes[(size_t)i].mx[k];
            }
            if (node_hmax) node_hmax[i - 1] = nodes[(size_t)i].hmax;
        }
        return PMX_OK;
    });
}

int pmx_export_particles(pmx_ctx* c, double* const* cols, int32_t ncols) {
    return guard(c, [&]() -> int {
        PMX_TRY(require_tree(c));
        PMX_HIP(c, hipSetDevice(c->device));
        if (!cols || ncols < 0 || ncols > 6 + c->tree.ncol) return set_error(c, PMX_ERR_DIMENSION, "export_particles: ncols");
        const int64_t n = c->tree.n;
        PMX_TRY(ensure_cols64(c));
        PMX_HIP(c, hipStreamSynchronize(c->stream));
        for (int j = 0; j < ncols; ++j)
            if (cols[j]) PMX_HIP(c, hipMemcpy(cols[j], c->cols64.as<double>() + (size_t)j * n, 8 * (size_t)n, hipMemcpyDeviceToHost));
        return PMX_OK;
    });
}

int pmx_nearest(pmx_ctx* c, int32_t precision, const pmx_points* pts, int64_t* idx_out) {
    return guard(c, [&]() -> int {
        PMX_TRY(require_tree(c));
        PMX_HIP(c, hipSetDevice(c->device));
        if (precision != PMX_PREC_F64 && precision != PMX_PREC_F32C) return set_error(c, PMX_ERR_ARGUMENT, "unknown precision");
        PMX_TRY(check_points(c, pts));
        if (!idx_out) return set_error(c, PMX_ERR_ARGUMENT, "nearest: null output");
        const int64_t n = pts->n;
        if (n == 0) return PMX_OK;
        const double *qx, *qy, *qz;
        PMX_TRY(stage_points(c, pts, &qx, &qy, &qz));
        PMX_TRY(ensure(c, c->scratch_c, 8 * (size_t)n));
        if (precision == PMX_PREC_F32C) PMX_TRY(nearest_f32c(c, qx, qy, qz, n, c->scratch_c.as<int64_t>()));
        else PMX_TRY(nearest_f64(c, qx, qy, qz, n, c->scratch_c.as<int64_t>()));
        PMX_HIP(c, hipMemcpyAsync(idx_out, c->scratch_c.p, 8 * (size_t)n, hipMemcpyDeviceToHost, c->stream));
        PMX_HIP(c, hipStreamSynchronize(c->stream));
        return PMX_OK;
    });
}

static int neighbors_common(pmx_ctx* c, int64_t n, const double* ha, int64_t* counts, int64_t* offsets,
                            int32_t* idx, int64_t capacity,
                            const std::function<int(const double*, int64_t*, const int64_t*, int32_t*)>& run) {
    if (!counts) return set_error(c, PMX_ERR_ARGUMENT, "neighbors: counts is required");
    if (n == 0) {
        if (offsets) offsets[0] = 0;
        return PMX_OK;
    }
    const double* dha = nullptr;
    PMX_TRY(ensure(c, c->scratch_a, 8 * (size_t)n));
    if (ha) {
        PMX_HIP(c, hipMemcpyAsync(c->scratch_a.p, ha, 8 * (size_t)n, hipMemcpyHostToDevice, c->stream));
        dha = c->scratch_a.as<double>();
    }
    PMX_TRY(ensure(c, c->scratch_b, 8 * (size_t)(n + 1)));
    int64_t* dcounts = c->scratch_b.as<int64_t>();
    PMX_TRY(run(dha, dcounts, nullptr, nullptr));
    PMX_HIP(c, hipMemcpyAsync(counts, dcounts, 8 * (size_t)n, hipMemcpyDeviceToHost, c->stream));
    PMX_HIP(c, hipStreamSynchronize(c->stream));
    if (!offsets || !idx) return PMX_OK;
    offsets[0] = 0;
    for (int64_t i = 0; i < n; ++i) offsets[i + 1] = offsets[i] + counts[i];
    if (offsets[n] > capacity) return set_error(c, PMX_ERR_CAPACITY, "neighbors: idx capacity too small");
    PMX_TRY(ensure(c, c->scratch_c, 8 * (size_t)(n + 1) + 4 * (size_t)std::max<int64_t>(offsets[n], 1)));
    int64_t* doff = c->scratch_c.as<int64_t>();
    int32_t* didx = reinterpret_cast<int32_t*>(doff + n + 1);
    PMX_HIP(c, hipMemcpyAsync(doff, offsets, 8 * (size_t)(n + 1), hipMemcpyHostToDevice, c->stream));
    PMX_TRY(run(dha, dcounts, doff, didx));
    PMX_HIP(c, hipMemcpyAsync(idx, didx, 4 * (size_t)offsets[n], hipMemcpyDeviceToHost, c->stream));
    PMX_HIP(c, hipStreamSynchronize(c->stream));
    return PMX_OK;
}

int pmx_point_neighbors(pmx_ctx* c, int32_t strategy, int32_t precision, const pmx_points* pts, const double* ha,
                        int64_t* counts, int64_t* offsets, int32_t* idx, int64_t capacity) {
    return guard(c, [&]() -> int {
        PMX_TRY(require_tree(c));
        PMX_HIP(c, hipSetDevice(c->device));
        if (precision != PMX_PREC_F64 && precision != PMX_PREC_F32C) return set_error(c, PMX_ERR_ARGUMENT, "unknown precision");
        if (strategy < 0 || strategy > 2) return set_error(c, PMX_ERR_ARGUMENT, "unknown interpolation strategy");
        PMX_TRY(check_points(c, pts));
        const int64_t n = pts->n;
        const double *qx = nullptr, *qy = nullptr, *qz = nullptr;
        if (n > 0) PMX_TRY(stage_points(c, pts, &qx, &qy, &qz));
        return neighbors_common(c, n, ha, counts, offsets, idx, capacity,
                                [&](const double* dha, int64_t* cnt, const int64_t* off, int32_t* ix) {
                                    if (precision == PMX_PREC_F32C)
                                        return neighbors_points_f32c(c, strategy, qx, qy, qz, dha, n, cnt, off, ix);
                                    return neighbors_points_f64(c, strategy, qx, qy, qz, dha, n, cnt, off, ix);
                                });
    });
}

int pmx_line_neighbors(pmx_ctx* c, int32_t strategy, int32_t precision, const pmx_lines* lines, const double* ha,
                       int64_t* counts, int64_t* offsets, int32_t* idx, int64_t capacity) {
    return guard(c, [&]() -> int {
        PMX_TRY(require_tree(c));
        PMX_HIP(c, hipSetDevice(c->device));
        if (precision != PMX_PREC_F64 && precision != PMX_PREC_F32C) return set_error(c, PMX_ERR_ARGUMENT, "unknown precision");
        if (strategy < 0 || strategy > 2) return set_error(c, PMX_ERR_ARGUMENT, "unknown interpolation strategy");
        PMX_TRY(check_lines(c, lines));
        const int64_t n = lines->n;
        const double *o = nullptr, *d = nullptr;
        if (n > 0) PMX_TRY(stage_lines(c, lines, &o, &d));
        return neighbors_common(c, n, ha, counts, offsets, idx, capacity,
                                [&](const double* dha, int64_t* cnt, const int64_t* off, int32_t* ix) {
                                    if (precision == PMX_PREC_F32C) {
                                        if (strategy != kScatter)
                                            return set_error(c, PMX_ERR_ARGUMENT,
                                                             "compensated line neighbours support itpScatter only");
                                        return neighbors_lines_f32c(c, o, d, n, cnt, off, ix);
                                    }
                                    return neighbors_lines_f64(c, strategy, o, d, dha, n, cnt, off, ix);
                                });
    });
}

int pmx_point_traversal_stats(pmx_ctx* c, int32_t strategy, const pmx_points* pts, int64_t* stats) {
    return guard(c, [&]() -> int {
        PMX_TRY(require_tree(c));
        PMX_HIP(c, hipSetDevice(c->device));
        if (strategy < 0 || strategy > 2) return set_error(c, PMX_ERR_ARGUMENT, "unknown interpolation strategy");
        PMX_TRY(check_points(c, pts));
        if (!stats) return set_error(c, PMX_ERR_ARGUMENT, "stats: null");
        for (int k = 0; k < 5; ++k) stats[k] = 0;
        if (pts->n == 0) return PMX_OK;
        const double *qx, *qy, *qz;
        PMX_TRY(stage_points(c, pts, &qx, &qy, &qz));
        PMX_TRY(ensure(c, c->scratch_b, 64));
        unsigned long long* tot = c->scratch_b.as<unsigned long long>();
        PMX_HIP(c, hipMemsetAsync(tot, 0, 40, c->stream));
        PMX_TRY(traversal_stats_f64(c, 0, strategy, qx, qy, qz, pts->n, tot));
        PMX_HIP(c, hipMemcpyAsync(stats, tot, 40, hipMemcpyDeviceToHost, c->stream));
        PMX_HIP(c, hipStreamSynchronize(c->stream));
        return PMX_OK;
    });
}

int pmx_line_traversal_stats(pmx_ctx* c, const pmx_lines* lines, int64_t* stats) {
    return guard(c, [&]() -> int {
        PMX_TRY(require_tree(c));
        PMX_HIP(c, hipSetDevice(c->device));
        PMX_TRY(check_lines(c, lines));
        if (!stats) return set_error(c, PMX_ERR_ARGUMENT, "stats: null");
        for (int k = 0; k < 5; ++k) stats[k] = 0;
        if (lines->n == 0) return PMX_OK;
        const double *o, *d;
        PMX_TRY(stage_lines(c, lines, &o, &d));
        PMX_TRY(ensure(c, c->scratch_b, 64));
        unsigned long long* tot = c->scratch_b.as<unsigned long long>();
        PMX_HIP(c, hipMemsetAsync(tot, 0, 40, c->stream));
        PMX_TRY(traversal_stats_f64(c, 1, kScatter, o, d, nullptr, lines->n, tot));
        PMX_HIP(c, hipMemcpyAsync(stats, tot, 40, hipMemcpyDeviceToHost, c->stream));
        PMX_HIP(c, hipStreamSynchronize(c->stream));
        return PMX_OK;
    });
}

int pmx_line_neighbor_digest(pmx_ctx* c, const pmx_lines* lines, uint64_t* digest, int32_t* used) {
    return guard(c, [&]() -> int {
        PMX_TRY(require_tree(c));
        PMX_HIP(c, hipSetDevice(c->device));
        PMX_TRY(check_lines(c, lines));
        if (!digest || !used) return set_error(c, PMX_ERR_ARGUMENT, "digest/used: null");
        *used = 0;
        const int64_t n = lines->n;
        if (n == 0) return PMX_OK;
        const double *o, *d;
        PMX_TRY(stage_lines(c, lines, &o, &d));
        PMX_TRY(prepare_f32(c));
        PMX_TRY(ensure(c, c->outbuf, 24 * (size_t)n));
        bool tiled = false;
        PMX_TRY(line_digest_tiled(c, o, d, n, c->outbuf.as<uint64_t>(), &tiled));
        if (tiled) {
            PMX_HIP(c, hipMemcpyAsync(digest, c->outbuf.p, 24 * (size_t)n, hipMemcpyDeviceToHost, c->stream));
            *used = 1;
        }
        PMX_HIP(c, hipStreamSynchronize(c->stream));
        return PMX_OK;
    });
}

int pmx_kernel_eval(pmx_ctx* c, int32_t kernel, int64_t n, const double* q, double* f, double* df) {
    return guard(c, [&]() -> int {
        if (!c) return PMX_ERR_ARGUMENT;
        PMX_HIP(c, hipSetDevice(c->device));
        if (kernel < 0 || kernel > 5) return set_error(c, PMX_ERR_ARGUMENT, "unknown kernel");
        if (n < 0 || (n > 0 && (!q || !f || !df))) return set_error(c, PMX_ERR_ARGUMENT, "kernel_eval: arguments");
        if (n == 0) return PMX_OK;
        PMX_TRY(ensure(c, c->scratch_a, 24 * (size_t)n));
        double* dq = c->scratch_a.as<double>();
        PMX_HIP(c, hipMemcpyAsync(dq, q, 8 * (size_t)n, hipMemcpyHostToDevice, c->stream));
        PMX_TRY(kernel_eval_f64(c, kernel, dq, n, dq + n, dq + 2 * n));
        PMX_HIP(c, hipMemcpyAsync(f, dq + n, 8 * (size_t)n, hipMemcpyDeviceToHost, c->stream));
        PMX_HIP(c, hipMemcpyAsync(df, dq + 2 * n, 8 * (size_t)n, hipMemcpyDeviceToHost, c->stream));
        PMX_HIP(c, hipStreamSynchronize(c->stream));
        return PMX_OK;
    });
}

int pmx_line_kernel_table(int32_t kernel, double* Q, double* Inorm) {
    return guard(nullptr, [&]() -> int {
        if (kernel < 0 || kernel > 5 || !Q || !Inorm) return PMX_ERR_ARGUMENT;
        line_table(kernel, Q, Inorm);
        return PMX_OK;
    });
}

int pmx_last_timings(const pmx_ctx* c, double* ms, int32_t n) {
    return guard(c, [&]() -> int {
        if (!c || !ms) return PMX_ERR_ARGUMENT;
        pmx_ctx* m = const_cast<pmx_ctx*>(c);
        PMX_TRY(settle_build_timings(m));
        for (int i = 0; i < n && i < 10; ++i) ms[i] = c->timings[i];
        return PMX_OK;
    });
}

}  // extern "C"
