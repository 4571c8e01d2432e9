// pmx_batch.hip — many independent snapshots in one persistent interpolation kernel.
//
// Each job (a particle set with its own samples) keeps its LBVH in a context of its own.  A
// run rebuilds every tree (capture-safe builds: no host synchronisation), then one persistent
// kernel pops tiles from a device work queue: sample tiles of 64 Morton-ordered points
// (pts::points_tile) and 8 x 8-cell line-map tiles (lmap::lines_tile) of every job, heaviest
// jobs first.  The whole sequence can be captured into one hipGraph and replayed.  The
// autotuner times candidate launch shapes (waves per CU, leaf-run sizes, job order) with HIP
// events and keeps the fastest.  No counterpart in the reference, which interpolates one
// snapshot per call (initialize_interpolation.jl:49-65).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "pmx_internal.h"
#include "pmx_lines_tile.h"
#include "pmx_points_tile.h"
#include "pmx_tree32.h"

namespace pmx {
namespace {

// One work unit: a job (or one sub-catalog of it: <= 2 scalars, 1 gradient, 1 divergence and
// 1 curl for samples, <= 8 scalars for lines) and its tile range in the queue.
struct BatchUnit {
    int kind;                  // PMX_JOB_POINTS / PMX_JOB_LINES
    int strategy;
    int deriv;                 // points: the sub-catalog holds derivative terms (dynamic tile layout)
    T32 T;
    pts::PointCat pcat;
    const double* qx;
    const double* qy;
    const double* qz;
    const int32_t* qperm;
    int64_t npts;
    lmap::Frame f;
    lmap::TileCat lcat;
    const int32_t* cstart;
    const int32_t* lperm;
    const double2* sts;
    const double* o;
    int64_t nl;
    const float* lut;
    double inv_dq;
    float* out;
    int64_t ldo;
    int64_t tile0;
    int64_t ntiles;
    // line units: the tolerances of f are re-derived from the rebuilt particle box before every
    // interpolation (k_frame_scale): root box (n >= 2) or the single leaf, and the lines' scale
    const Node64* root;
    double amax;
    int tilt;                  // nearly parallel line set (tv)
    lmap::TiltView tv;
};

// Line units: the frame's tolerances from the current particle box (the particles may move
// between runs; the lines stay as planned).
__global__ void k_frame_scale(BatchUnit* __restrict__ units, int nunits) {
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= nunits) return;
    BatchUnit& U = units[u];
    if (U.kind != PMX_JOB_LINES) return;
    double E = U.amax;
    if (U.T.n >= 2) {
        for (int k = 0; k < 3; ++k) E = fmax(E, fmax(fabs(U.root->mn[k]), fabs(U.root->mx[k])));
    } else {
        // the single particle: its FP32 position is within 2^-24 of the FP64 one
        const float4 H = U.T.leaf[0];
        E = fmax(E, (1.0 + 1e-6) * fmax(fabs((double)H.x), fmax(fabs((double)H.y), fabs((double)H.z))));
    }
    lmap::frame_scale(U.f, 2.0 * E);
}

// Sample tiles of derivative sub-catalogs: the tile layout is sized by the unit's catalog.
// (P: the scalar tiles' staging area, at least as large as the derivative tiles' runs need)
template <int K, int S>
__device__ void batch_deriv_tile(const BatchUnit& U, int64_t t, pts::BPos& P, void* dyn, pts::PointSplit sp) {
    pts::BTile<true> Y(dyn, U.pcat.nder, U.pcat.nacc);
    auto& Pd = *reinterpret_cast<pts::BPosT<kRunDeriv>*>(&P);
    pts::points_tile<K, S, true>(U.T, U.pcat, U.qx, U.qy, U.qz, U.qperm, U.npts, U.out, U.ldo, t, Pd, Y, sp);
}

// DERIV: the batch holds derivative units (an instantiation of its own, launched only then, so
// that scalar batches keep the smaller register budget).
// TILT: the batch holds nearly parallel line sets (likewise an instantiation of its own).
template <int K, bool DERIV, bool TILT>
__global__ void __launch_bounds__(64) k_batch(const BatchUnit* __restrict__ units, int nunits, int64_t total,
                                              unsigned long long* __restrict__ counter, pts::PointSplit split) {
    // dynamic LDS shared by the tile kinds: points {BPos, BTile}, lines {LineTileLDS, dyn}
    extern __shared__ double2 lds2[];             // 16-byte aligned dynamic LDS
    double* lds = reinterpret_cast<double*>(lds2);
    __shared__ long long item_s;
    pts::BPos& P = *reinterpret_cast<pts::BPos*>(lds);
    pts::BTile<false>& Y = *reinterpret_cast<pts::BTile<false>*>(reinterpret_cast<char*>(lds) + sizeof(pts::BPos));
    lmap::LineTileLDS& L = *reinterpret_cast<lmap::LineTileLDS*>(lds);
    double* ldyn = reinterpret_cast<double*>(reinterpret_cast<char*>(lds) + sizeof(lmap::LineTileLDS));
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) item_s = (long long)atomicAdd(counter, 1ull);
        __syncthreads();
        const int64_t item = item_s;
        if (item >= total) break;
        int lo = 0, hi = nunits - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (units[mid].tile0 <= item) lo = mid;
            else hi = mid - 1;
        }
        const BatchUnit& U = units[lo];
        const int64_t t = item - U.tile0;
        if (U.kind == PMX_JOB_POINTS) {
            pts::PointSplit sp = split;
            sp.unit = lo;
            if constexpr (DERIV) {
                if (U.deriv) {
                    void* dyn = reinterpret_cast<char*>(lds) + sizeof(pts::BPos);
                    if (U.strategy == kGather) batch_deriv_tile<K, kGather>(U, t, P, dyn, sp);
                    else if (U.strategy == kScatter) batch_deriv_tile<K, kScatter>(U, t, P, dyn, sp);
                    else batch_deriv_tile<K, kSymmetric>(U, t, P, dyn, sp);
                    continue;
                }
            }
            if (U.strategy == kGather)
                pts::points_tile<K, kGather, false>(U.T, U.pcat, U.qx, U.qy, U.qz, U.qperm, U.npts, U.out, U.ldo, t, P, Y, sp);
            else if (U.strategy == kScatter)
                pts::points_tile<K, kScatter, false>(U.T, U.pcat, U.qx, U.qy, U.qz, U.qperm, U.npts, U.out, U.ldo, t, P, Y, sp);
            else
                pts::points_tile<K, kSymmetric, false>(U.T, U.pcat, U.qx, U.qy, U.qz, U.qperm, U.npts, U.out, U.ldo, t,
                                                       P, Y, sp);
        } else {
            if constexpr (TILT) {
                if (U.tilt) {
                    lmap::lines_tile<K, 0, true, false, true>(U.T, U.f, U.lcat, U.cstart, U.lperm, U.sts, U.o, U.nl, U.lut,
                                                              U.inv_dq, U.out, U.ldo, nullptr, t, L, ldyn, U.tv);
                    continue;
                }
            }
            lmap::lines_tile<K, 0, true>(U.T, U.f, U.lcat, U.cstart, U.lperm, U.sts, U.o, U.nl, U.lut, U.inv_dq, U.out, U.ldo,
                                   nullptr, t, L, ldyn);
        }
    }
}

// Items of split point tiles of every job (after k_batch), then the split tiles' outputs.
// The work unit of a queued split item.
__device__ inline const BatchUnit& U_of(const BatchUnit* __restrict__ units, const pts::PointSplit& sp,
                                        unsigned int item) {
    return units[uniform(sp.slot_unit[uniform(sp.items[item].x)])];
}

// One queued item of a split scalar tile.
template <int K>
__device__ void batch_item(const BatchUnit& U, const pts::PointSplit& sp, unsigned int item, pts::BPos& P,
                           pts::BTile<false>& Y) {
    if (U.strategy == kGather)
        pts::points_item<K, kGather, false>(U.T, U.pcat, U.qx, U.qy, U.qz, U.qperm, U.npts, sp, item, P, Y);
    else
        pts::points_item<K, kSymmetric, false>(U.T, U.pcat, U.qx, U.qy, U.qz, U.qperm, U.npts, sp, item, P, Y);
}

// DERIV: dynamic LDS of max(sizeof(BTile<false>), the largest derivative tile) bytes, shared by
// the two tile layouts; scalar batches keep the static layout.
template <int K, bool DERIV>
__global__ void __launch_bounds__(64) k_batch_items(const BatchUnit* __restrict__ units, pts::PointSplit sp) {
    __shared__ pts::BPos P;
    const unsigned int nitems = min(sp.counters[0], (unsigned int)sp.cap_items);
    if constexpr (DERIV) {
        extern __shared__ double2 dyn2[];
        for (;;) {
            const unsigned int item = pts::lane0_atomic_add(&sp.counters[2], 1u);
            if (item >= nitems) break;
            const BatchUnit& U = U_of(units, sp, item);
            if (U.deriv) {
                pts::BTile<true> Yd(dyn2, U.pcat.nder, U.pcat.nacc);
                auto& Pd = *reinterpret_cast<pts::BPosT<kRunDeriv>*>(&P);
                if (U.strategy == kGather)
                    pts::points_item<K, kGather, true>(U.T, U.pcat, U.qx, U.qy, U.qz, U.qperm, U.npts, sp, item, Pd, Yd);
                else
                    pts::points_item<K, kSymmetric, true>(U.T, U.pcat, U.qx, U.qy, U.qz, U.qperm, U.npts, sp, item, Pd,
                                                         Yd);
            } else {
                batch_item<K>(U, sp, item, P, *reinterpret_cast<pts::BTile<false>*>(dyn2));
            }
        }
    } else {
        __shared__ pts::BTile<false> Y;
        for (;;) {
            const unsigned int item = pts::lane0_atomic_add(&sp.counters[2], 1u);
            if (item >= nitems) break;
            batch_item<K>(U_of(units, sp, item), sp, item, P, Y);
        }
    }
}

__global__ void __launch_bounds__(64) k_batch_finish(const BatchUnit* __restrict__ units, pts::PointSplit sp) {
    const int nslots = (int)min(sp.counters[1], (unsigned int)sp.cap_slots);
    for (int split = blockIdx.x; split < nslots; split += gridDim.x) {
        const BatchUnit& U = units[sp.slot_unit[split]];
        if (U.deriv) pts::points_finish<true>(U.pcat, U.qperm, U.npts, U.out, U.ldo, sp, split);
        else pts::points_finish<false>(U.pcat, U.qperm, U.npts, U.out, U.ldo, sp, split);
    }
}

__global__ void k_zero_u64(unsigned long long* p) { *p = 0ull; }

__global__ void k_zero_split(unsigned int* p) {
    if (threadIdx.x < 4) p[threadIdx.x] = 0u;
}

}  // namespace
}  // namespace pmx

using namespace pmx;

struct pmx_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    int kernel = 0;
    std::vector<pmx_ctx*> ctx;
    std::vector<pmx_batch_job> jobs;
    std::vector<std::vector<void*>> out_cols;     // copies of the callers' column pointers
    std::vector<const int32_t*> qperm;            // per job: Morton order of point samples
    std::vector<lmap::LinePlan> plan;             // per job: binned lines
    std::vector<const double*> lines_o;           // per job: packed line origins
    std::vector<BatchUnit> units;
    DevBuf units_dev, counter, split_buf;
    pts::PointSplit split{};                      // heavy point tiles (gather / symmetric)
    int split_budget = 4096, split_item = 2048;
    int64_t total_tiles = 0;
    size_t lds_bytes = 0;
    size_t lds_items = 0;                         // dynamic LDS of the split-item kernel
    bool deriv = false;                           // some unit holds derivative terms
    bool tilt = false;                            // some line unit is a nearly parallel set
    int ncu = 256;
    // launch shape (autotuned)
    int waves_per_cu = 16, bucket_points = 64, bucket_lines = 512, heavy_first = 1;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t ev[4] = {};
    double info[4] = {0, 0, 0, 0};
    // The snapshots' builds are independent: they run on kBuildStreams streams (default priority)
    // forked from and joined back into `stream`, so small builds overlap (parallel branches of
    // the captured graph).  build_order: jobs by decreasing particle count, dealt round-robin.
    static constexpr int kBuildStreams = 4;
    hipStream_t bstream[kBuildStreams] = {};
    hipEvent_t fork = nullptr, join[kBuildStreams] = {};
    std::vector<int> build_order;
};

namespace pmx {
void note_exception(pmx_batch* b, const char* msg) noexcept {
    if (!b) return;
    try {
        b->err = msg;
    } catch (...) {
    }
}
void note_exception(const pmx_batch* b, const char* msg) noexcept { note_exception(const_cast<pmx_batch*>(b), msg); }
}  // namespace pmx

namespace {

int bfail(pmx_batch* b, int code, const std::string& msg) {
    b->err = msg;
    return code;
}

#define PMXB_HIP(b, expr)                                                                   \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess) return bfail(b, PMX_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

int ctx_fail(pmx_batch* b, pmx_ctx* c, int rc, int job) {
    return bfail(b, rc, "job " + std::to_string(job) + ": " + std::string(pmx_last_error(c)));
}

// Split-tile storage shared by the point jobs (gather / symmetric); budget 0 when no job
// needs it or splitting is off.
int make_split(pmx_batch* b) {
    pts::PointSplit& sp = b->split;
    sp = pts::PointSplit{};
    bool any = false;
    for (const pmx_batch_job& J : b->jobs) any = any || (J.kind == PMX_JOB_POINTS && J.strategy != PMX_SCATTER);
    if (!any) return PMX_OK;
    sp.budget = b->split_budget;
    sp.item_max = std::max(b->split_item, 2 * b->bucket_points);
    sp.cap_items = kSplitItems;
    sp.cap_slots = kSplitSlots;
    auto up = [](size_t v) { return (v + 255) & ~(size_t)255; };
    const size_t b_cnt = 256, b_items = up(8 * (size_t)sp.cap_items), b_tile = up(4 * (size_t)sp.cap_slots),
                 b_unit = up(4 * (size_t)sp.cap_slots), b_near = up(4 * (size_t)sp.cap_slots * 64),
                 b_acc = 8 * (size_t)sp.cap_slots * pts::kAccMax * 64;
    if (ensure(b->ctx[0], b->split_buf, b_cnt + b_items + b_tile + b_unit + b_near + b_acc) != PMX_OK)
        return bfail(b, PMX_ERR_NOMEM, "batch: split storage allocation failed");
    char* p = b->split_buf.as<char>();
    sp.counters = reinterpret_cast<unsigned int*>(p);
    p += b_cnt;
    sp.items = reinterpret_cast<int2*>(p);
    p += b_items;
    sp.slot_tile = reinterpret_cast<int*>(p);
    p += b_tile;
    sp.slot_unit = reinterpret_cast<int*>(p);
    p += b_unit;
    sp.slot_near = reinterpret_cast<int*>(p);
    p += b_near;
    sp.slot_acc = reinterpret_cast<double*>(p);
    return PMX_OK;
}

// Output columns of a catalog: scalars, 3 per gradient, 1 per divergence, 3 per curl.
int out_columns(const pmx_catalog& cat) { return cat.n_scalar + 3 * cat.n_grad + cat.n_div + 3 * cat.n_curl; }

// A job's catalog as slot lists (validated at creation; a gradient slot of -1 is rho's).
FullCat full_catalog(const pmx_catalog& cat) {
    FullCat fc;
    for (int j = 0; j < cat.n_scalar; ++j) {
        fc.ss.push_back(cat.scalar_slots[j]);
        fc.sh.push_back(cat.scalar_shepard ? (cat.scalar_shepard[j] ? 1 : 0) : 1);
    }
    for (int j = 0; j < cat.n_grad; ++j) fc.gs.push_back(cat.grad_slots[j]);
    for (int j = 0; j < 3 * cat.n_div; ++j) fc.ds.push_back(cat.div_slots[j]);
    for (int j = 0; j < 3 * cat.n_curl; ++j) fc.cs.push_back(cat.curl_slots[j]);
    return fc;
}

// Work units from the jobs with the current launch shape; tile ranges in queue order.
int make_units(pmx_batch* b) {
    b->units.clear();
    if (b->split.counters) {
        b->split.budget = b->split_budget;
        b->split.item_max = std::max(b->split_item, 2 * b->bucket_points);
    }
    size_t lds = std::max<size_t>(sizeof(pts::BPos) + sizeof(pts::BTile<false>), sizeof(lmap::LineTileLDS) + 64);
    size_t lds_items = sizeof(pts::BTile<false>);
    b->deriv = false;
    b->tilt = false;
    std::vector<std::pair<double, BatchUnit>> list;
    for (size_t j = 0; j < b->jobs.size(); ++j) {
        const pmx_batch_job& J = b->jobs[j];
        pmx_ctx* c = b->ctx[j];
        const int64_t nP = J.kind == PMX_JOB_POINTS ? J.points.n : J.lines.n;
        const double weight = (double)c->tree.n;        // heavier snapshots first
        float* outbuf = c->outbuf.as<float>();
        auto unit = [&] {
            BatchUnit U{};
            std::memset(&U, 0, sizeof(U));
            U.kind = J.kind;
            U.strategy = J.strategy;
            U.out = outbuf;
            U.ldo = nP;
            return U;
        };
        if (J.kind == PMX_JOB_POINTS) {
            // the per-call split: chunks of the whole catalog, then their sub-catalogs
            for (const CatChunk& ch : make_chunks(full_catalog(J.catalog)))
                for (const pts::PointCat& cat : point_cats(ch)) {
                    BatchUnit U = unit();
                    U.pcat = cat;
                    U.deriv = cat.ng + cat.nd + cat.nc > 0;
                    U.T = tree32(c, b->bucket_points, run_cap(U.deriv));
                    U.qx = J.points.x;
                    U.qy = J.points.y;
                    U.qz = J.points.z;
                    U.qperm = b->qperm[j];
                    U.npts = nP;
                    U.ntiles = (nP + 63) / 64;
                    if (U.deriv) {
                        const size_t dl = pts::deriv_tile_lds(cat.nder, cat.nacc);
                        lds = std::max(lds, sizeof(pts::BPos) + dl);
                        lds_items = std::max(lds_items, dl);
                        b->deriv = true;
                    }
                    list.push_back({weight, U});
                }
            continue;
        }
        const int ns = J.catalog.n_scalar;
        const int per = lmap::kMaxCols;
        for (int base = 0; base < std::max(ns, 1); base += per) {
            BatchUnit U = unit();
            const int cnt = std::min(per, ns - base);
            U.T = tree32(c, 64, kMaxRunLines);
            U.T.bucket = b->bucket_lines;
            U.f = b->plan[j].f;
            U.f.near = (double)c->tune.tiled_near;
            U.lcat.n = cnt;
            for (int k = 0; k < cnt; ++k) {
                U.lcat.fcol[k] = 1 + J.catalog.scalar_slots[base + k];
                U.lcat.shep[k] = J.catalog.scalar_shepard ? (J.catalog.scalar_shepard[base + k] ? 1 : 0) : 1;
            }
            U.cstart = b->plan[j].cstart;
            U.lperm = b->plan[j].perm;
            U.sts = b->plan[j].sts;
            U.root = c->nodes64.as<Node64>() + 1;
            U.amax = b->plan[j].amax;
            U.tilt = b->plan[j].tilt ? 1 : 0;
            U.tv = b->plan[j].tv;
            if (U.tilt) b->tilt = true;
            U.o = b->lines_o[j];
            U.nl = nP;
            U.lut = c->lutf.as<float>() + (size_t)b->kernel * kLutF;
            U.inv_dq = 1.0 / c->lut_dq[b->kernel];
            U.out = outbuf + (int64_t)base * nP;
            U.ntiles = (int64_t)U.f.tx * U.f.ty;
            lds = std::max(lds, sizeof(lmap::LineTileLDS) + lmap::lines_tile_lds(cnt, 0, true));
            list.push_back({weight, U});
        }
    }
    if (b->heavy_first)
        std::stable_sort(list.begin(), list.end(), [](const auto& x, const auto& y) { return x.first > y.first; });
    int64_t t0 = 0;
    for (auto& e : list) {
        e.second.tile0 = t0;
        t0 += e.second.ntiles;
        b->units.push_back(e.second);
    }
    b->total_tiles = t0;
    b->lds_bytes = (lds + 15) & ~(size_t)15;
    b->lds_items = (lds_items + 15) & ~(size_t)15;
    pmx_ctx* c0 = b->ctx[0];
    if (ensure(c0, b->units_dev, sizeof(BatchUnit) * std::max<size_t>(b->units.size(), 1)) != PMX_OK)
        return bfail(b, PMX_ERR_NOMEM, "batch: unit table allocation failed");
    PMXB_HIP(b, hipMemcpy(b->units_dev.p, b->units.data(), sizeof(BatchUnit) * b->units.size(), hipMemcpyHostToDevice));
    return PMX_OK;
}

template <int K>
void launch_batch(pmx_batch* b, int64_t grid) {
    hipStream_t s = b->stream;
    const BatchUnit* U = b->units_dev.as<BatchUnit>();
    const pts::PointSplit& sp = b->split;
    auto kb = b->tilt ? (b->deriv ? k_batch<K, true, true> : k_batch<K, false, true>)
                      : (b->deriv ? k_batch<K, true, false> : k_batch<K, false, false>);
    auto ki = b->deriv ? k_batch_items<K, true> : k_batch_items<K, false>;
    const size_t lds_items = b->deriv ? b->lds_items : 0;
    hipLaunchKernelGGL(kb, dim3(grid), dim3(64), b->lds_bytes, s, U, (int)b->units.size(), b->total_tiles,
                       b->counter.as<unsigned long long>(), sp);
    if (sp.budget <= 0) return;
    hipLaunchKernelGGL(ki, dim3(b->ncu * 32), dim3(64), lds_items, s, U, sp);
    hipLaunchKernelGGL(k_batch_finish, dim3(std::min(sp.cap_slots, 4096)), dim3(64), 0, s, U, sp);
}

// Persistent kernel (then the split tiles' items and outputs) plus the copies into the
// callers' columns (stream-ordered, capturable).
int enqueue_interpolation(pmx_batch* b) {
    hipStream_t s = b->stream;
    const int nu = (int)b->units.size();
    hipLaunchKernelGGL(k_frame_scale, dim3((nu + 63) / 64), dim3(64), 0, s, b->units_dev.as<BatchUnit>(), nu);
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(1), 0, s, b->counter.as<unsigned long long>());
    if (b->split.budget > 0) hipLaunchKernelGGL(k_zero_split, dim3(1), dim3(64), 0, s, b->split.counters);
    const int64_t grid = std::max<int64_t>(1, std::min<int64_t>(b->total_tiles, (int64_t)b->ncu * b->waves_per_cu));
    switch (b->kernel) {
    case kM4: launch_batch<kM4>(b, grid); break;
    case kM5: launch_batch<kM5>(b, grid); break;
    case kM6: launch_batch<kM6>(b, grid); break;
    case kC2: launch_batch<kC2>(b, grid); break;
    case kC4: launch_batch<kC4>(b, grid); break;
    default: launch_batch<kC6>(b, grid); break;
    }
    PMXB_HIP(b, hipGetLastError());
    for (size_t j = 0; j < b->jobs.size(); ++j) {
        const pmx_batch_job& J = b->jobs[j];
        const int64_t nP = J.kind == PMX_JOB_POINTS ? J.points.n : J.lines.n;
        const float* src = b->ctx[j]->outbuf.as<float>();
        for (size_t k = 0; k < b->out_cols[j].size(); ++k)
            PMXB_HIP(b, hipMemcpyAsync(b->out_cols[j][k], src + (int64_t)k * nP, 4 * (size_t)nP, hipMemcpyDeviceToDevice, s));
    }
    return PMX_OK;
}

// Every job's capture-safe build: forked from the batch stream onto the build streams, joined
// back before anything that follows on the batch stream.
int enqueue_builds(pmx_batch* b) {
    constexpr int K = pmx_batch::kBuildStreams;
    PMXB_HIP(b, hipEventRecord(b->fork, b->stream));
    for (int k = 0; k < K; ++k) PMXB_HIP(b, hipStreamWaitEvent(b->bstream[k], b->fork, 0));
    for (size_t r = 0; r < b->build_order.size(); ++r) {
        const int j = b->build_order[r];
        pmx_ctx* c = b->ctx[j];
        hipStream_t own = c->stream;
        c->stream = b->bstream[r % K];
        const int rc = build_lbvh(c, &b->jobs[j].particles, kBuildAsync, nullptr, nullptr);
        c->stream = own;
        if (rc != PMX_OK) return ctx_fail(b, c, rc, j);
    }
    for (int k = 0; k < K; ++k) {
        PMXB_HIP(b, hipEventRecord(b->join[k], b->bstream[k]));
        PMXB_HIP(b, hipStreamWaitEvent(b->stream, b->join[k], 0));
    }
    return PMX_OK;
}

int check_nans(pmx_batch* b) {
    for (size_t j = 0; j < b->jobs.size(); ++j) {
        int nan = 0;
        PMXB_HIP(b, hipMemcpy(&nan, nan_count_device(b->ctx[j]), sizeof(int), hipMemcpyDeviceToHost));
        if (nan > 0)
            return bfail(b, PMX_ERR_ARGUMENT,
                         "job " + std::to_string(j) + ": InexactError: NaN particle coordinate (UInt64(floor(NaN)))");
    }
    return PMX_OK;
}

void drop_graph(pmx_batch* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

int time_interpolation(pmx_batch* b, double* ms) {
    PMXB_HIP(b, hipEventRecord(b->ev[2], b->stream));
    PMX_TRY(enqueue_interpolation(b));
    PMXB_HIP(b, hipEventRecord(b->ev[3], b->stream));
    PMXB_HIP(b, hipStreamSynchronize(b->stream));
    float t = 0;
    PMXB_HIP(b, hipEventElapsedTime(&t, b->ev[2], b->ev[3]));
    *ms = t;
    return PMX_OK;
}

}  // namespace

extern "C" {

int pmx_batch_create(pmx_batch** out, int32_t device_ordinal, int32_t njobs, const pmx_batch_job* jobs) {
    return guard(nullptr, [&]() -> int {
        if (!out) return PMX_ERR_ARGUMENT;
        *out = nullptr;
        pmx_batch* b = new (std::nothrow) pmx_batch();
        if (!b) return PMX_ERR_NOMEM;
        *out = b;
        if (njobs < 1 || !jobs) return bfail(b, PMX_ERR_ARGUMENT, "batch: needs at least one job");
        b->device = device_ordinal;
        PMXB_HIP(b, hipSetDevice(device_ordinal));
        PMXB_HIP(b, hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
        for (auto& e : b->ev) PMXB_HIP(b, hipEventCreate(&e));
        for (auto& bs : b->bstream) PMXB_HIP(b, hipStreamCreateWithFlags(&bs, hipStreamNonBlocking));
        PMXB_HIP(b, hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
        for (auto& e : b->join) PMXB_HIP(b, hipEventCreateWithFlags(&e, hipEventDisableTiming));
        hipDeviceProp_t prop;
        PMXB_HIP(b, hipGetDeviceProperties(&prop, device_ordinal));
        b->ncu = std::max(1, prop.multiProcessorCount);
        b->kernel = jobs[0].particles.kernel;
        for (int j = 0; j < njobs; ++j) {
            const pmx_batch_job& J = jobs[j];
            const std::string tag = "batch job " + std::to_string(j) + ": ";
            if (J.particles.kernel != b->kernel) return bfail(b, PMX_ERR_ARGUMENT, tag + "all jobs must use one SPH kernel");
            if (J.particles.mem != PMX_MEM_DEVICE) return bfail(b, PMX_ERR_ARGUMENT, tag + "particle columns must be device memory");
            if (J.kind != PMX_JOB_POINTS && J.kind != PMX_JOB_LINES) return bfail(b, PMX_ERR_ARGUMENT, tag + "unknown kind");
            if (J.kind == PMX_JOB_LINES && J.strategy != PMX_SCATTER)
                return bfail(b, PMX_ERR_ARGUMENT, tag + "LineSamples_interpolation only supports itpScatter.");
            if (J.kind == PMX_JOB_POINTS && (J.strategy < 0 || J.strategy > 2))
                return bfail(b, PMX_ERR_ARGUMENT, tag + "unknown interpolation strategy");
            const pmx_catalog& Cg = J.catalog;
            if (Cg.n_scalar < 0 || Cg.n_grad < 0 || Cg.n_div < 0 || Cg.n_curl < 0)
                return bfail(b, PMX_ERR_DIMENSION, tag + "catalog: negative count");
            if (J.kind == PMX_JOB_LINES && (Cg.n_grad || Cg.n_div || Cg.n_curl))
                return bfail(b, PMX_ERR_ARGUMENT,
                             tag + "LineSamples_interpolation: catalog must hold scalars only (InterpolationCatalog{3,N,0,0,0,N})");
            if (out_columns(Cg) < 1) return bfail(b, PMX_ERR_ARGUMENT, tag + "catalog: no quantities");
            const int nc = J.particles.ncol;
            auto bad = [nc](const int* v, int n, bool rho_ok) {
                if (n > 0 && !v) return true;
                for (int k = 0; k < n; ++k)
                    if (!((v[k] >= 0 && v[k] < nc) || (rho_ok && v[k] == -1))) return true;
                return false;
            };
            if (bad(Cg.scalar_slots, Cg.n_scalar, false))
                return bfail(b, PMX_ERR_DIMENSION, tag + "catalog: scalar slot out of range");
            if (bad(Cg.grad_slots, Cg.n_grad, true))
                return bfail(b, PMX_ERR_DIMENSION, tag + "catalog: gradient slot out of range");
            if (bad(Cg.div_slots, 3 * Cg.n_div, false))
                return bfail(b, PMX_ERR_DIMENSION, tag + "catalog: divergence slot out of range");
            if (bad(Cg.curl_slots, 3 * Cg.n_curl, false))
                return bfail(b, PMX_ERR_DIMENSION, tag + "catalog: curl slot out of range");
            const pmx_output& O = J.out;
            if (O.dtype != PMX_F32 || O.mem != PMX_MEM_DEVICE || !O.cols)
                return bfail(b, PMX_ERR_ARGUMENT, tag + "outputs must be FP32 device columns");
            if (J.kind == PMX_JOB_POINTS && (J.points.mem != PMX_MEM_DEVICE || J.points.n < 1))
                return bfail(b, PMX_ERR_ARGUMENT, tag + "points must be device memory, n >= 1");
            if (J.kind == PMX_JOB_LINES && (J.lines.mem != PMX_MEM_DEVICE || J.lines.n < 1))
                return bfail(b, PMX_ERR_ARGUMENT, tag + "lines must be device memory, n >= 1");
        }
        b->jobs.assign(jobs, jobs + njobs);
        b->build_order.resize((size_t)njobs);
        for (int j = 0; j < njobs; ++j) b->build_order[(size_t)j] = j;
        std::stable_sort(b->build_order.begin(), b->build_order.end(),
                         [&](int a, int c2) { return jobs[a].particles.n > jobs[c2].particles.n; });
        b->qperm.assign(njobs, nullptr);
        b->plan.resize(njobs);
        b->lines_o.assign(njobs, nullptr);
        b->out_cols.resize(njobs);
        for (int j = 0; j < njobs; ++j) {
            pmx_batch_job& J = b->jobs[j];
            b->out_cols[j].assign(J.out.cols, J.out.cols + out_columns(J.catalog));
            pmx_ctx* c = nullptr;
            const int rc = pmx_create(&c, device_ordinal);
            if (rc != PMX_OK) return bfail(b, rc, "batch: context creation failed");
            b->ctx.push_back(c);
            // plan: a full (synchronous) build sizes every buffer the captured builds reuse
            int r = build_lbvh(c, &J.particles, 0, nullptr, nullptr);
            if (r != PMX_OK) return ctx_fail(b, c, r, j);
            const int64_t nP = J.kind == PMX_JOB_POINTS ? J.points.n : J.lines.n;
            if (ensure(c, c->outbuf, 4 * (size_t)nP * out_columns(J.catalog)) != PMX_OK) return ctx_fail(b, c, PMX_ERR_NOMEM, j);
            if (J.kind == PMX_JOB_POINTS) {
                int32_t* perm = nullptr;
                r = sort_queries(c, J.points.x, J.points.y, J.points.z, nP, &perm);
                if (r != PMX_OK) return ctx_fail(b, c, r, j);
                b->qperm[j] = perm;
            } else {
                if (ensure(c, c->qbuf, 48 * (size_t)nP) != PMX_OK) return ctx_fail(b, c, PMX_ERR_NOMEM, j);
                double* q = c->qbuf.as<double>();
                const double* src[6] = {J.lines.ox, J.lines.oy, J.lines.oz, J.lines.dx, J.lines.dy, J.lines.dz};
                for (int k = 0; k < 6; ++k)
                    PMXB_HIP(b, hipMemcpyAsync(q + k * nP, src[k], 8 * (size_t)nP, hipMemcpyDeviceToDevice, c->stream));
                bool ok = false;
                r = lines_plan(c, q, q + 3 * nP, nP, &b->plan[j], &ok);
                if (r != PMX_OK) return ctx_fail(b, c, r, j);
                if (!ok) return bfail(b, PMX_ERR_ARGUMENT, "batch job " + std::to_string(j) +
                                                               ": lines must be one parallel or nearly parallel "
                                                               "unit-direction set");
                b->lines_o[j] = q;
            }
            PMXB_HIP(b, hipStreamSynchronize(c->stream));
        }
        if (ensure(b->ctx[0], b->counter, 64) != PMX_OK) return bfail(b, PMX_ERR_NOMEM, "batch: counter allocation failed");
        if (const char* v = std::getenv("PMX_SPLIT_BUDGET")) b->split_budget = std::max(0, std::atoi(v));
        if (const char* v = std::getenv("PMX_SPLIT_ITEM")) b->split_item = std::max(64, std::atoi(v));
        PMX_TRY(make_split(b));
        PMX_TRY(make_units(b));
        b->info[3] = (double)b->total_tiles;
        return PMX_OK;
    });
}

int pmx_batch_interpolate(pmx_batch* b) {
    return guard(b, [&]() -> int {
        if (!b || b->ctx.empty()) return PMX_ERR_ARGUMENT;
        PMXB_HIP(b, hipSetDevice(b->device));
        double ms = 0;
        PMX_TRY(time_interpolation(b, &ms));
        b->info[2] = ms;
        return PMX_OK;
    });
}

int pmx_batch_run(pmx_batch* b, int32_t use_graph) {
    return guard(b, [&]() -> int {
        if (!b || b->ctx.empty()) return PMX_ERR_ARGUMENT;
        PMXB_HIP(b, hipSetDevice(b->device));
        hipStream_t s = b->stream;
        if (use_graph) {
            if (!b->exec) {
                PMXB_HIP(b, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                int rc = enqueue_builds(b);
                if (rc == PMX_OK) rc = enqueue_interpolation(b);
                hipGraph_t g = nullptr;
                const hipError_t ec = hipStreamEndCapture(s, &g);
                if (rc != PMX_OK) {
                    if (g) (void)hipGraphDestroy(g);
                    return rc;
                }
                if (ec != hipSuccess) return bfail(b, PMX_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ec));
                b->graph = g;
                PMXB_HIP(b, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
            }
            PMXB_HIP(b, hipEventRecord(b->ev[0], s));
            PMXB_HIP(b, hipGraphLaunch(b->exec, s));
            PMXB_HIP(b, hipEventRecord(b->ev[1], s));
        } else {
            PMXB_HIP(b, hipEventRecord(b->ev[0], s));
            PMX_TRY(enqueue_builds(b));
            PMXB_HIP(b, hipEventRecord(b->ev[2], s));
            PMX_TRY(enqueue_interpolation(b));
            PMXB_HIP(b, hipEventRecord(b->ev[1], s));
        }
        PMXB_HIP(b, hipStreamSynchronize(s));
        float t = 0;
        PMXB_HIP(b, hipEventElapsedTime(&t, b->ev[0], b->ev[1]));
        b->info[0] = t;
        if (!use_graph) {
            float tb = 0, ti = 0;
            PMXB_HIP(b, hipEventElapsedTime(&tb, b->ev[0], b->ev[2]));
            PMXB_HIP(b, hipEventElapsedTime(&ti, b->ev[2], b->ev[1]));
            b->info[1] = tb;
            b->info[2] = ti;
        }
        return check_nans(b);
    });
}

int pmx_batch_tune(pmx_batch* b) {
    return guard(b, [&]() -> int {
        if (!b || b->ctx.empty()) return PMX_ERR_ARGUMENT;
        PMXB_HIP(b, hipSetDevice(b->device));
        // coordinate descent over the launch shape, each candidate timed once after a warm-up
        struct Knob {
            int* v;
            std::vector<int> cand;
        };
        Knob knobs[] = {{&b->waves_per_cu, {4, 8, 16, 32}},
                        {&b->bucket_lines, {128, 256, 512, 1024}},
                        {&b->bucket_points, {16, 32, 64}},
                        {&b->heavy_first, {0, 1}},
                        {&b->split_budget, {4096, 8192, 16384}}};
        double best_ms = 0;
        PMX_TRY(make_units(b));
        PMX_TRY(time_interpolation(b, &best_ms));
        PMX_TRY(time_interpolation(b, &best_ms));
        for (Knob& k : knobs) {
            const int keep = *k.v;
            int best = keep;
            for (int cand : k.cand) {
                if (cand == keep) continue;
                *k.v = cand;
                PMX_TRY(make_units(b));
                double ms = 0;
                PMX_TRY(time_interpolation(b, &ms));
                if (ms < best_ms) {
                    best_ms = ms;
                    best = cand;
                }
            }
            *k.v = best;
        }
        PMX_TRY(make_units(b));
        drop_graph(b);   // the captured graph holds the old launch shape
        b->info[2] = best_ms;
        return PMX_OK;
    });
}

int pmx_batch_info(const pmx_batch* b, double* vals, int32_t n) {
    return guard(b, [&]() -> int {
        if (!b || !vals) return PMX_ERR_ARGUMENT;
        const double all[9] = {b->info[0], b->info[1], b->info[2], b->info[3], (double)b->waves_per_cu,
                               (double)b->bucket_points, (double)b->bucket_lines, (double)b->heavy_first,
                               (double)b->split_budget};
        for (int i = 0; i < n && i < 9; ++i) vals[i] = all[i];
        return PMX_OK;
    });
}

const char* pmx_batch_last_error(const pmx_batch* b) { return b ? b->err.c_str() : "null batch"; }

void pmx_batch_destroy(pmx_batch* b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    for (auto& bs : b->bstream)
        if (bs) (void)hipStreamSynchronize(bs);
    drop_graph(b);
    if (b->units_dev.p) (void)hipFree(b->units_dev.p);
    if (b->counter.p) (void)hipFree(b->counter.p);
    if (b->split_buf.p) (void)hipFree(b->split_buf.p);
    for (pmx_ctx* c : b->ctx) pmx_destroy(c);
    for (auto& e : b->ev)
        if (e) (void)hipEventDestroy(e);
    for (auto& bs : b->bstream)
        if (bs) (void)hipStreamDestroy(bs);
    if (b->fork) (void)hipEventDestroy(b->fork);
    for (auto& e : b->join)
        if (e) (void)hipEventDestroy(e);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

}  // extern "C"
