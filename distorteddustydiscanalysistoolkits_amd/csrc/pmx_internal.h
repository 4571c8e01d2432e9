// pmx_internal.h — host-side context and internal entry points of libpartia_mi355x.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/partia_mi355x.h"
#include "pmx_device.h"

struct pmx_phantom;   // Phantom dump reader (pmx_phantom.hip, include/partia_phantom.h)

namespace pmx {

// Grow-only device buffer owned by a context.
struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
    template <typename T>
    T* as() const { return static_cast<T*>(p); }
};

struct Error {
    int code;
    std::string msg;
};

// FP32 tree for the compensated path: node boxes rounded outward from the FP64 boxes,
// hmax rounded up, so every FP32 node test is conservative w.r.t. the FP64 geometry.
struct __align__(16) Node32 {
    float mnx, mny, mnz, hmax;
    float mxx, mxy, mxz;
    int32_t left;
    int32_t escape;
    int32_t lo;       // first leaf (0-based sorted index) of the node's range
    int32_t cnt;      // number of leaves in the range
    int32_t pad;
};

// Launch-shape knobs of a context, read from the environment when it is created:
// PMX_LINES_TILED=0 sends every line set through the union walk; PMX_BUCKET_POINTS /
// PMX_BUCKET_LINES set the leaf-run size the sample-tile walks hand over at once (kMinRun..64;
// derivative point tiles clamp it to 32, their staging capacity);
// PMX_TILED_BUCKET the leaf-run size of the line-map tile walk (1..65536); PMX_TILED_NEAR how
// close (in cells) a node's box must be to the tile for the walk to read it whole;
// PMX_REFIT_SMALL the leaf-run length up to which the refit computes a node directly from its
// leaves.
// Every FP32 walk hands over a node of at most kMinRun leaves as one leaf run (leaf-run knobs
// are clamped to at least kMinRun), so the build writes no FP32 record for a node whose parent
// covers at most kMinRun leaves: no walk reads it.
constexpr int kMinRun = 16;

struct Tuning {
    int lines_tiled = 1;
    int bucket_points = 64;
    int bucket_lines = 64;
    int tiled_bucket = 1024;
    int tiled_near = 16;
    int refit_small = 128;
};

struct Tree {
    int64_t n = 0;
    int32_t root = 0;
    int32_t ncol = 0;
    int32_t kernel = 0;
    bool built = false;
    bool f32_ready = false;
    bool leaf64_ready = false;    // plain FP64 path: Leaf64 records and FP64 node boxes, built on first use
    // FP32 input: the FP32 leaf and payload arrays hold every input value exactly, so the build
    // writes no FP64 column copy; cols64 is gathered from the packed records on first use
    bool f32in = false;
    bool cols64_ready = false;
};

}  // namespace pmx

struct pmx_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    pmx::Tree tree;
    // kernel tables (device): Inorm per kernel, FP64 and FP32, and grid step dq
    pmx::DevBuf lut64;
    pmx::DevBuf lutf;          // Inorm per kernel rounded to FP32, stride kLutF (zeros after each table; the line-map kernel copies it to LDS)
    double lut_dq[6] = {0, 0, 0, 0, 0, 0};
    // Morton-ordered particle data (device)
    pmx::DevBuf codes, order, keys_tmp, vals_tmp;
    pmx::DevBuf cols64;        // (6 + ncol) x n doubles: x y z m h rho quant...
    pmx::DevBuf leaf64;        // Leaf64[n]
    pmx::DevBuf leaf_esc;      // int32[n]  escape of leaf li (unified id n + li)
    pmx::DevBuf nodes64;       // Node64[n] by internal id
    pmx::DevBuf nodes32;       // Node32[n] by internal id
    pmx::DevBuf leaf32;        // float4[n] x y z h
    pmx::DevBuf pay32;         // float payload for the compensated path
    pmx::DevBuf lr, parent, escape;  // int32 [2n] by unified id (left/right interleaved in lr)
    pmx::DevBuf visited, range_far;   // range_far: other end of each internal node's leaf range
    pmx::DevBuf unread;        // u8 by internal id: 1 when no FP32 walk reads the node (its parent covers <= kMinRun leaves)
    pmx::DevBuf scratch_a, scratch_b, scratch_c, hist, staging, outbuf, qbuf, misc;
    int ncu = 256;             // compute units of the device
    pmx::Tuning tune;          // launch-shape knobs (environment at creation)
    void* noalloc_scratch = nullptr;   // this device's address of the no-allocation scratch array
    bool line_inv_h2 = false;  // the current line call weights by Inorm / h^2 (PMX_LINE_INV_H2)
    int split_budget = 4096;   // heavy-tile splitting of FP32c gather / symmetric points (0 = off)
    int split_item = 2048;     // largest subtree of a split tile queued as one item
    pmx::DevBuf split;         // split-tile counters, item queue and slots (FP32c points)
    pmx::DevBuf gridpts;       // samples of a Cartesian structured grid (x | y | z), then its axes
    pmx::DevBuf records;       // particle records (AoS) for the Morton-order permutation
    pmx::DevBuf tl_st, tl_part, tl_cells;  // line-map binning (pmx_lines_tiled.hip)
    pmx::DevBuf tl_sl, tl_tsl;             // nearly parallel sets: per-line slopes (unsorted, sorted), per-tile bounds
    hipEvent_t ev[12] = {};
    double timings[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    bool build_timings_pending = false;   // timings[0..3] not yet read from the last build's events
};

namespace pmx {

int set_error(pmx_ctx* c, int code, const std::string& msg);
// The same for a plain text: copied into the context's reserved error buffer (no allocation for
// texts that fit it).
int set_error(pmx_ctx* c, int code, const char* msg);

// No C++ exception crosses the C ABI: every int-returning entry point runs its body through
// guard(owner, body).  std::bad_alloc becomes PMX_ERR_NOMEM, any other exception
// PMX_ERR_INTERNAL, and the message goes to the owner's last-error text (a context, a batch, a
// Phantom reader, the open-error string, or nothing).  The environment variable PMX_TEST_THROW
// ("bad_alloc" or "exception") makes the guard throw before the body runs: the tests of this
// conversion use it.
void note_exception(pmx_ctx* c, const char* msg) noexcept;
void note_exception(const pmx_ctx* c, const char* msg) noexcept;
void note_exception(pmx_batch* b, const char* msg) noexcept;
void note_exception(const pmx_batch* b, const char* msg) noexcept;
void note_exception(pmx_phantom* ph, const char* msg) noexcept;
void note_exception(const pmx_phantom* ph, const char* msg) noexcept;
void note_exception(std::string* s, const char* msg) noexcept;
inline void note_exception(std::nullptr_t, const char*) noexcept {}
void throw_test_fault();

template <class Owner, class Body>
int guard(Owner owner, Body&& body) noexcept {
    try {
        throw_test_fault();
        return body();
    } catch (const std::bad_alloc&) {
        note_exception(owner, "host memory allocation failed (std::bad_alloc)");
        return PMX_ERR_NOMEM;
    } catch (const std::exception& e) {
        note_exception(owner, e.what());
        return PMX_ERR_INTERNAL;
    } catch (...) {
        note_exception(owner, "unknown C++ exception");
        return PMX_ERR_INTERNAL;
    }
}
int hip_fail(pmx_ctx* c, hipError_t e, const char* what);
int ensure(pmx_ctx* c, DevBuf& b, size_t bytes);

#define PMX_HIP(ctx, expr)                                     \
    do {                                                       \
        hipError_t _e = (expr);                                \
        if (_e != hipSuccess) return ::pmx::hip_fail(ctx, _e, #expr); \
    } while (0)

#define PMX_TRY(expr)              \
    do {                           \
        int _rc = (expr);          \
        if (_rc != PMX_OK) return _rc; \
    } while (0)

inline unsigned grid_for(int64_t n, int block, int64_t max_blocks = 1 << 20) {
    int64_t g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > max_blocks) g = max_blocks;
    return (unsigned)g;
}

// build (pmx_build.hip)
// build_lbvh flag (internal): capture-safe build, see pmx_build.hip
// capacities of the split-tile storage (slots of FP64 sums, queued subtree items: 8 B each; an
// item that finds the queue full is walked by its tile's wavefront alone, which is far slower)
constexpr int kSplitSlots = 65536, kSplitItems = 1 << 23;
constexpr int kBuildAsync = 1 << 30;
int* nan_count_device(pmx_ctx* c);
int build_lbvh(pmx_ctx* c, const pmx_particles* p, int flags, uint64_t* codes_out,
               int64_t* order_out);
int settle_build_timings(pmx_ctx* c);
// Stable LSD radix sort of (key, value) pairs on bits [begin_bit, end_bit), 8 bits per pass; the
// result is in keys_alt / vals_alt when *result_in_alt.  Both key arrays 16-byte aligned.  hist: the histogram buffer of
// radix_hist_bytes(n) bytes (null: the context's); first_hist_ready: it already holds the first
// pass's per-tile digit counts (k_encode writes them); vals_are_indices: the input values are
// 0 .. n-1 and are not read (vals is then only written, by the second pass).
int radix_sort_pairs(pmx_ctx* c, uint64_t* keys, int32_t* vals, int64_t n, uint64_t* keys_alt,
                     int32_t* vals_alt, int begin_bit, int end_bit, bool* result_in_alt, uint32_t* hist = nullptr,
                     bool first_hist_ready = false, bool vals_are_indices = false);
constexpr size_t radix_hist_bytes(int64_t n) { return sizeof(uint32_t) * 256 * (size_t)((n + 4095) / 4096 + 1); }

// Device work memory of one line-map call.  The per-call entry points take it from the
// context's grow-only buffers (ws == nullptr); the no-allocation entry point passes regions of
// its static scratch array instead.
constexpr int kBinBlocks = 1024;                                        // blocks of the line projection
constexpr size_t kBinPartBytes = sizeof(double) * 5 * kBinBlocks + 64;  // bounds, flag, tile counter
struct LineWork {
    double* st = nullptr;                  // (s, t) of each line origin: 2 nl doubles
    double* part = nullptr;                // kBinPartBytes
    uint64_t* keys0 = nullptr;             // sort pairs, nl each
    int32_t* vals0 = nullptr;
    uint64_t* keys1 = nullptr;
    int32_t* vals1 = nullptr;
    uint32_t* hist = nullptr;              // radix_hist_bytes(nl)
    char* cells = nullptr;                 // cell starts (ncell + 1 int32, 16-aligned) then sorted (s, t)
    size_t cells_bytes = 0;
};

// The no-allocation scratch array of the context's device (pmx_noalloc.hip): its address on
// this device, looked up once per device at context creation.
int noalloc_init_device(pmx_ctx* c);

// interpolation (pmx_interp64.hip / pmx_interp32.hip)
struct CatChunk {
    int ns, ng, nd, nc;
    int ss[4];
    int sh[4];
    int gs[2];
    int ds[6];
    int cs[6];
    int os, og, od, oc;   // catalog positions of this chunk's first scalar/gradient/div/curl output
};
// A whole catalog as 0-based slot lists (div / curl slots three per field).
struct FullCat {
    std::vector<int> ss, sh, gs, ds, cs;
    int L() const { return (int)(ss.size() + 3 * gs.size() + ds.size() / 3 + cs.size()); }
};
std::vector<CatChunk> make_chunks(const FullCat& fc);
int interp_points_f64(pmx_ctx* c, int strategy, const CatChunk& cat, const double* qx,
                      const double* qy, const double* qz, int64_t npts, double* out, int64_t ldo,
                      const double* ha_in, double* ha_out);
int interp_lines_f64(pmx_ctx* c, int strategy, const int* slots, const int* shep, int ncols,
                     const double* o, const double* d, int64_t nl, double* out, int64_t ldo);
int nearest_f64(pmx_ctx* c, const double* qx, const double* qy, const double* qz, int64_t npts,
                int64_t* idx);
int neighbors_points_f64(pmx_ctx* c, int strategy, const double* qx, const double* qy,
                         const double* qz, const double* ha, int64_t npts, int64_t* counts,
                         const int64_t* offsets, int32_t* idx);
int neighbors_lines_f64(pmx_ctx* c, int strategy, const double* o, const double* d,
                        const double* ha, int64_t nl, int64_t* counts, const int64_t* offsets,
                        int32_t* idx);
int kernel_eval_f64(pmx_ctx* c, int kernel, const double* q, int64_t n, double* f, double* df);
int traversal_stats_f64(pmx_ctx* c, int kind, int strategy, const double* a, const double* b, const double* cc,
                        int64_t n, unsigned long long* tot);

// Morton order of the samples (FP32-compensated path tiles); the permutation lives in the
// context's scratch buffers until the next call that uses them.
int sort_queries(pmx_ctx* c, const double* x, const double* y, const double* z, int64_t npts, int32_t** perm,
                 const LineWork* ws = nullptr);
// Leaf64 records of the current tree from cols64 (the plain FP64 path packs them on first use).
int ensure_leaf64(pmx_ctx* c);
// cols64 of the current tree (a gather of the packed input records for an FP32 input).
int ensure_cols64(pmx_ctx* c);
int interp_points_f32c(pmx_ctx* c, int strategy, const CatChunk& cat, const int32_t* perm,
                       const double* qx, const double* qy, const double* qz, int64_t npts, float* out, int64_t ldo);
int interp_lines_f32c(pmx_ctx* c, const int* slots, const int* shep, int ncols, const int32_t* perm,
                      const double* o, const double* d, int64_t nl, float* out, int64_t ldo);
int prepare_f32(pmx_ctx* c);
// Parallel unit-direction line sets: particle-centric tiled path.  *used is false (and
// nothing is written) when the lines do not qualify; the caller then uses the union walk.
int interp_lines_tiled(pmx_ctx* c, const int* slots, const int* shep, int ncols, const double* o, const double* d,
                       int64_t nl, float* out, int64_t ldo, bool* used, const LineWork* ws = nullptr);
int line_digest_tiled(pmx_ctx* c, const double* o, const double* d, int64_t nl, uint64_t* digest, bool* used);
int nearest_f32c(pmx_ctx* c, const double* qx, const double* qy, const double* qz, int64_t npts, int64_t* idx);
int neighbors_points_f32c(pmx_ctx* c, int strategy, const double* qx, const double* qy, const double* qz,
                          const double* ha, int64_t npts, int64_t* counts, const int64_t* offsets, int32_t* idx);
int neighbors_lines_f32c(pmx_ctx* c, const double* o, const double* d, int64_t nl, int64_t* counts,
                         const int64_t* offsets, int32_t* idx);

// line-integrated kernel tables (pmx_lut.cpp)
void line_table(int kernel, double* Q, double* Inorm);

}  // namespace pmx
