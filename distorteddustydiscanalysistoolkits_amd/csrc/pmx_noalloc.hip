// pmx_noalloc.hip — the line-map entry point that allocates nothing.
//
// pmx_interpolate_lines_noalloc runs the kernels of pmx_interpolate_lines' FP32-compensated path
// (line binning, radix sort, the particle-centric tiled kernel, or the union walk for
// non-parallel sets) with every work buffer carved from one file-scope device array instead of
// the context's grow-only buffers.  Each device holds its own copy of the array from the moment
// the library's code object is loaded there until the process ends; its address on a device is
// looked up once, when the first context on that device is created (noalloc_init_device), and
// only that address is handed to kernels and copies.  Calls on one device share the array, so
// calls that overlap on a device (two host threads, or two contexts on one device) race.
// Reference semantics: those of pmx_interpolate_lines (LineSamples_interpolation,
// line_integrated_scalar_interpolation.jl:123-166, itpScatter only).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <mutex>

#include "pmx_internal.h"

namespace pmx {
namespace {

// Limits of one call: 2^20 lines (a 1024 x 1024 column map) and 4 scalar columns.
constexpr int64_t kMaxLines = int64_t(1) << 20;
constexpr int kMaxCols = 4;
// binning cells the scratch holds: about one per line plus the rounding of a lattice to 8 x 8 tiles
constexpr int64_t kMaxCells = 2 * kMaxLines + 4096;

constexpr size_t up256(size_t v) { return (v + 255) & ~size_t(255); }

// Regions of the scratch array (byte offsets).
constexpr size_t kOffLines = 0;                                                  // o[3n], d[3n]
constexpr size_t kOffOut = kOffLines + up256(48 * (size_t)kMaxLines);            // FP32 results [L][n]
constexpr size_t kOffSt = kOffOut + up256(4 * (size_t)kMaxCols * kMaxLines);     // (s, t) per line
constexpr size_t kOffPart = kOffSt + up256(16 * (size_t)kMaxLines);              // bounds, flag, tile counter
constexpr size_t kOffKeys0 = kOffPart + up256(kBinPartBytes);
constexpr size_t kOffVals0 = kOffKeys0 + up256(8 * (size_t)kMaxLines);
constexpr size_t kOffKeys1 = kOffVals0 + up256(4 * (size_t)kMaxLines);
constexpr size_t kOffVals1 = kOffKeys1 + up256(8 * (size_t)kMaxLines);
constexpr size_t kOffHist = kOffVals1 + up256(4 * (size_t)kMaxLines);
constexpr size_t kHistBytes = radix_hist_bytes(kMaxLines);
constexpr size_t kOffCells = kOffHist + up256(kHistBytes);
constexpr size_t kCellsBytes = ((4 * (size_t)(kMaxCells + 1) + 15) & ~size_t(15)) + 16 * (size_t)kMaxLines;
constexpr size_t kScratchBytes = kOffCells + up256(kCellsBytes);                // about 128 MiB

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_dev_ptr[kMaxDevices] = {};

__global__ void __launch_bounds__(256) k_widen(const float* __restrict__ in, double* __restrict__ out, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
        out[i] = (double)in[i];
}

}  // namespace

int noalloc_init_device(pmx_ctx* c) {
    if (c->device < 0 || c->device >= kMaxDevices) return set_error(c, PMX_ERR_ARGUMENT, "device ordinal beyond 63");
    std::lock_guard<std::mutex> lock(g_mu);
    void*& p = g_dev_ptr[c->device];
    if (!p) {
        // the device is current (pmx_create); each device has its own copy of the array
        void* q = nullptr;
        const hipError_t e = hipGetSymbolAddress(&q, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_fail(c, e, "hipGetSymbolAddress(scratch)");
        if (!q) return set_error(c, PMX_ERR_HIP, "hipGetSymbolAddress(scratch) returned a null address");
        p = q;
    }
    c->noalloc_scratch = p;
    return PMX_OK;
}

}  // namespace pmx

using namespace pmx;

extern "C" {

int pmx_interpolate_lines_noalloc(pmx_ctx* c, int32_t strategy, int32_t precision, const pmx_catalog* cat,
                                  const pmx_lines* lines, const pmx_output* out) {
    return guard(c, [&]() -> int {
        if (!c) return PMX_ERR_ARGUMENT;
        if (!c->tree.built) return set_error(c, PMX_ERR_STATE, "no LBVH built in this context (call pmx_build_lbvh)");
        PMX_HIP(c, hipSetDevice(c->device));
        if (strategy != PMX_SCATTER)
            return set_error(c, PMX_ERR_ARGUMENT,
                             "LineSamples_interpolation only supports itpScatter. Line-integrated samples do not have a "
                             "well-defined query smoothing length ha, so itpGather and itpSymmetric are not supported.");
        const bool inv_h2 = (precision & PMX_LINE_INV_H2) != 0;
        if ((precision & ~PMX_LINE_INV_H2) != PMX_PREC_F32C)
            return set_error(c, PMX_ERR_ARGUMENT, "pmx_interpolate_lines_noalloc: the FP32-compensated path only (PMX_PREC_F32C)");
        // catalog: scalars only, at most kMaxCols of them
        if (!cat) return set_error(c, PMX_ERR_ARGUMENT, "catalog: null pointer");
        if (cat->n_grad || cat->n_div || cat->n_curl)
            return set_error(c, PMX_ERR_ARGUMENT,
                             "LineSamples_interpolation: catalog must hold scalars only (InterpolationCatalog{3,N,0,0,0,N})");
        const int L = cat->n_scalar;
        if (L < 0) return set_error(c, PMX_ERR_DIMENSION, "catalog: negative count");
        if (L > kMaxCols) return set_error(c, PMX_ERR_CAPACITY, "pmx_interpolate_lines_noalloc: at most 4 scalar columns");
        if (L > 0 && !cat->scalar_slots) return set_error(c, PMX_ERR_ARGUMENT, "catalog: null scalar slots");
        int slots[kMaxCols], shep[kMaxCols];
        for (int j = 0; j < L; ++j) {
            slots[j] = cat->scalar_slots[j];
            if (slots[j] < 0 || slots[j] >= c->tree.ncol) return set_error(c, PMX_ERR_DIMENSION, "catalog: scalar slot out of range");
            shep[j] = cat->scalar_shepard ? (cat->scalar_shepard[j] ? 1 : 0) : 1;
        }
        // lines and outputs: device memory, within the scratch limits
        if (!lines) return set_error(c, PMX_ERR_ARGUMENT, "lines: null pointer");
        const int64_t n = lines->n;
        if (n < 0) return set_error(c, PMX_ERR_DIMENSION, "lines: negative count");
        if (n > kMaxLines) return set_error(c, PMX_ERR_CAPACITY, "pmx_interpolate_lines_noalloc: at most 1048576 lines");
        if (lines->mem != PMX_MEM_DEVICE) return set_error(c, PMX_ERR_ARGUMENT, "pmx_interpolate_lines_noalloc: lines must be in device memory");
        if (n > 0 && (!lines->ox || !lines->oy || !lines->oz || !lines->dx || !lines->dy || !lines->dz))
            return set_error(c, PMX_ERR_ARGUMENT, "lines: null pointer");
        if (!out || (L > 0 && !out->cols)) return set_error(c, PMX_ERR_ARGUMENT, "output: null pointer");
        if (out->mem != PMX_MEM_DEVICE) return set_error(c, PMX_ERR_ARGUMENT, "pmx_interpolate_lines_noalloc: outputs must be in device memory");
        if (out->dtype != PMX_F32 && out->dtype != PMX_F64) return set_error(c, PMX_ERR_ARGUMENT, "output: dtype");
        for (int j = 0; j < L; ++j)
            if (!out->cols[j] && n > 0) return set_error(c, PMX_ERR_ARGUMENT, "output: null column");
        if (!c->noalloc_scratch) return set_error(c, PMX_ERR_STATE, "no-allocation scratch array missing on this device");
        PMX_TRY(prepare_f32(c));
        if (n == 0) return PMX_OK;

        char* base = static_cast<char*>(c->noalloc_scratch);
        double* lbuf = reinterpret_cast<double*>(base + kOffLines);
        float* obuf = reinterpret_cast<float*>(base + kOffOut);
        LineWork ws;
        ws.st = reinterpret_cast<double*>(base + kOffSt);
        ws.part = reinterpret_cast<double*>(base + kOffPart);
        ws.keys0 = reinterpret_cast<uint64_t*>(base + kOffKeys0);
        ws.vals0 = reinterpret_cast<int32_t*>(base + kOffVals0);
        ws.keys1 = reinterpret_cast<uint64_t*>(base + kOffKeys1);
        ws.vals1 = reinterpret_cast<int32_t*>(base + kOffVals1);
        ws.hist = reinterpret_cast<uint32_t*>(base + kOffHist);
        ws.cells = base + kOffCells;
        ws.cells_bytes = kCellsBytes;

        PMX_HIP(c, hipEventRecord(c->ev[4], c->stream));
        // lines as two packed arrays o[3n], d[3n], as the line kernels read them
        const double* src[6] = {lines->ox, lines->oy, lines->oz, lines->dx, lines->dy, lines->dz};
        for (int j = 0; j < 6; ++j)
            PMX_HIP(c, hipMemcpyAsync(lbuf + j * n, src[j], 8 * (size_t)n, hipMemcpyDeviceToDevice, c->stream));
        const double* o = lbuf;
        const double* d = lbuf + 3 * n;
        c->line_inv_h2 = inv_h2;
        PMX_HIP(c, hipEventRecord(c->ev[5], c->stream));
        bool tiled = false;
        PMX_TRY(interp_lines_tiled(c, slots, shep, L, o, d, n, obuf, n, &tiled, &ws));
        if (!tiled) {
            int32_t* perm = nullptr;
            PMX_TRY(sort_queries(c, o, o + n, o + 2 * n, n, &perm, &ws));
            PMX_HIP(c, hipEventRecord(c->ev[8], c->stream));
            PMX_TRY(interp_lines_f32c(c, slots, shep, L, perm, o, d, n, obuf, n));
        }
        PMX_HIP(c, hipEventRecord(c->ev[6], c->stream));
        for (int j = 0; j < L; ++j) {
            if (out->dtype == PMX_F32) {
                PMX_HIP(c, hipMemcpyAsync(out->cols[j], obuf + (size_t)j * n, 4 * (size_t)n, hipMemcpyDeviceToDevice,
                                          c->stream));
            } else {
                hipLaunchKernelGGL(k_widen, dim3(grid_for(n, 256, 4096)), dim3(256), 0, c->stream, obuf + (size_t)j * n,
                                   static_cast<double*>(out->cols[j]), n);
            }
        }
        PMX_HIP(c, hipGetLastError());
        PMX_HIP(c, hipEventRecord(c->ev[7], c->stream));
        PMX_HIP(c, hipStreamSynchronize(c->stream));
        float a = 0, b = 0, k = 0;
        PMX_HIP(c, hipEventElapsedTime(&a, c->ev[5], c->ev[6]));
        PMX_HIP(c, hipEventElapsedTime(&b, c->ev[4], c->ev[7]));
        PMX_HIP(c, hipEventElapsedTime(&k, c->ev[8], c->ev[6]));
        c->timings[5] = a;
        c->timings[6] = b;
        c->timings[7] = k;
        return PMX_OK;
    });
}

}  // extern "C"
