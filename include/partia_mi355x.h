/*
 * partia_mi355x.h — C ABI of libpartia_mi355x.so, the MI355X (gfx950) implementation of
 * Partia.jl's Morton-code LBVH neighbour search and SPH kernel interpolation.
 *
 * The library replaces, behind Partia's backend dispatch, the methods a GPU extension adds
 * to Partia.PointSamples_interpolation / Partia.LineSamples_interpolation (reference:
 * Partia.jl/ext/CUDAExt/KernelInterpolation/grid_interpolation/PointSamples_interpolation.jl:12
 * and LineSamples_interpolation.jl:11) and the CPU LBVH build they call
 * (Partia.jl/src/julia/KernelInterpolation/grid_interpolation/initialize_interpolation.jl:49,
 * LinearBVH! at Partia.jl/src/julia/KernelInterpolation/interpolation_setup/InterpolationInput.jl:203).
 * Julia binds these symbols with `ccall` (julia/MI355XExt.jl of the package); the Python package binds
 * them with ctypes.
 *
 * Conventions
 *   - Every entry point returns PMX_OK (0) or a negative pmx_status and never throws or
 *     aborts across the ABI; pmx_last_error(ctx) describes the last failure.
 *   - Indices are 0-based.  The reference's 1-based values are index + 1.
 *   - Pointers are host pointers unless the struct field `mem` says PMX_MEM_DEVICE.
 *     Host buffers are borrowed for the duration of the call only.
 *   - A context owns one HIP stream and its device buffers; it is not re-entrant (one
 *     thread at a time per context).  Calls are synchronous on return.
 */
#ifndef PARTIA_MI355X_H
#define PARTIA_MI355X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum pmx_status {
    PMX_OK = 0,
    PMX_ERR_ARGUMENT = -1,   /* ≙ ArgumentError (bad strategy, n = 0, n > 2^30, NaN input)  */
    PMX_ERR_DIMENSION = -2,  /* ≙ DimensionMismatch (length / catalog slot mismatch)        */
    PMX_ERR_STATE = -3,      /* no LBVH built in this context yet                            */
    PMX_ERR_HIP = -4,        /* HIP runtime failure (≙ ErrorException)                       */
    PMX_ERR_NOMEM = -5,      /* device allocation failed                                     */
    PMX_ERR_CAPACITY = -6,   /* an output buffer supplied by the caller is too small         */
    PMX_ERR_FORMAT = -7,     /* not a readable Phantom dump (≙ ErrorException of read_phantom) */
    PMX_ERR_IO = -8,         /* the file cannot be opened or read (≙ SystemError)            */
    PMX_ERR_INTERNAL = -9    /* a C++ exception inside the library, turned into a status      */
} pmx_status;

/* Element type of particle / sample columns (InterpolationInput{3,T}: T ∈ {Float32, Float64}). */
enum { PMX_F32 = 0, PMX_F64 = 1 };
/* Where the buffers of a struct live. */
enum { PMX_MEM_HOST = 0, PMX_MEM_DEVICE = 1 };
/* SPH kernels (Partia.jl/src/julia/KernelInterpolation/kernel_function/kernels/). */
enum { PMX_M4 = 0, PMX_M5 = 1, PMX_M6 = 2, PMX_C2 = 3, PMX_C4 = 4, PMX_C6 = 5 };
/* itpGather / itpScatter / itpSymmetric (interpolation_setup/InterpolationStrategy.jl:3-5). */
enum { PMX_GATHER = 0, PMX_SCATTER = 1, PMX_SYMMETRIC = 2 };
/* Arithmetic of the interpolation kernels.
 *   PMX_PREC_F64  : plain FP64, one thread per sample, stackless escape traversal — the
 *                   reference GPU design restated (ext/CUDAExt/.../PointSamples_interpolation.jl:32).
 *   PMX_PREC_F32C : FP32 arithmetic with compensated per-sample sums, Morton-ordered
 *                   sample tiles, wavefront-cooperative traversal and LDS-staged candidates;
 *                   neighbour decisions are exact FP64 decisions. */
enum { PMX_PREC_F64 = 0, PMX_PREC_F32C = 1 };
/* Flag OR'd into the precision argument of pmx_interpolate_lines / pmx_line_samples_interpolation:
 * the line-integrated weight is Inorm(q) / h^2, the dimensionally correct column kernel, instead
 * of the reference's Inorm(q) / h (line_integrated_kernel.jl:132-137).  Off by default (parity);
 * batches (pmx_batch_*) always use 1 / h. */
enum { PMX_LINE_INV_H2 = 0x100 };
/* pmx_build_lbvh flags */
enum { PMX_BUILD_WRITEBACK = 1 };   /* permute the caller's columns (host or device) into Morton order */

/* ≙ InterpolationInput{3,T,V,K,NCOLUMN} (InterpolationInput.jl:43-51). */
typedef struct pmx_particles {
    int32_t dtype;           /* PMX_F32 | PMX_F64, shared by all columns                    */
    int32_t mem;             /* PMX_MEM_HOST | PMX_MEM_DEVICE                               */
    int64_t n;               /* Npart                                                       */
    void *x, *y, *z;         /* coord                                                       */
    void *m, *h, *rho;       /* masses, smoothing lengths, densities                       */
    int32_t ncol;            /* NCOLUMN                                                     */
    void *const *quant;      /* ncol column pointers                                        */
    int32_t kernel;          /* smoothed_kernel: PMX_M4 .. PMX_C6                           */
} pmx_particles;

/* ≙ InterpolationCatalogConcise (InterpolationCatalog.jl:26-32), 0-based quant slots.
 * Output order (L = n_scalar + 3 n_grad + n_div + 3 n_curl) is the reference's:
 * scalars, gradients (x,y,z each), divergences, curls (x,y,z each). */
typedef struct pmx_catalog {
    int32_t n_scalar;
    const int32_t *scalar_slots;
    const uint8_t *scalar_shepard;  /* 0 only for the names :ρ / :rho (InterpolationCatalog.jl:222-225) */
    int32_t n_grad;
    const int32_t *grad_slots;      /* -1 selects the density gradient (reference slot 0) */
    int32_t n_div;
    const int32_t *div_slots;       /* 3 per entry */
    int32_t n_curl;
    const int32_t *curl_slots;      /* 3 per entry */
} pmx_catalog;

/* Sample points (≙ PointSamples coordinates, P/src/julia/Grids/grids/PointSamples.jl:25). */
typedef struct pmx_points {
    int32_t mem;                     /* PMX_MEM_HOST | PMX_MEM_DEVICE; coordinates are FP64 */
    int64_t n;
    const double *x, *y, *z;
} pmx_points;

/* Axes of a Cartesian structured grid (≙ StructuredGrid axes, P/src/julia/Grids/grids/
 * StructuredGrid.jl:271-518); a 2-D slice in 3-D has n[2] = 1. */
typedef struct pmx_axes {
    int32_t mem;                     /* PMX_MEM_HOST | PMX_MEM_DEVICE; axis values are FP64 */
    int64_t n[3];
    const double *a[3];
} pmx_axes;

/* Sample lines (≙ LineSamples origin/direction, P/src/julia/Grids/grids/LineSamples.jl:31). */
typedef struct pmx_lines {
    int32_t mem;
    int64_t n;
    const double *ox, *oy, *oz;      /* origins                                    */
    const double *dx, *dy, *dz;      /* unit directions                            */
} pmx_lines;

/* Output columns: L pointers of `dtype` elements (host or device per `mem`). */
typedef struct pmx_output {
    int32_t dtype;
    int32_t mem;
    void *const *cols;
} pmx_output;

/* --- context ---------------------------------------------------------------------- */
/* No reference counterpart: creates the device context (stream, buffers, kernel tables). */
int pmx_create(struct pmx_ctx **out, int32_t device_ordinal);
void pmx_destroy(struct pmx_ctx *ctx);
/* Text of the last error on this context (valid until the next call on it). */
const char *pmx_last_error(const struct pmx_ctx *ctx);
const char *pmx_version(void);

/* --- build ------------------------------------------------------------------------ */
/* ≙ LinearBVH!(input, Val(3)) (InterpolationInput.jl:203-224): FP64 bounding-box
 * normalisation and 21-bit quantisation, 63-bit Morton codes (MortonCode.jl:443-498,
 * 209-217, 275-305), stable sort + de-dup (sort_by_morton!, MortonCode.jl:181-195),
 * Karras radix tree with escape links (BinaryRadixTree.jl:81-281) and atomic AABB / hmax
 * refit (LinearBVH.jl:55-156).  The particle columns are gathered into Morton order on
 * the device and kept in the context for the interpolation calls.
 *   codes_out (n, nullable): de-duplicated sorted codes (enc.codes)
 *   order_out (n, nullable): 0-based original index of each sorted slot (enc.order - 1)
 *   flags & PMX_BUILD_WRITEBACK: the caller's x, y, z, m, h, rho and quant arrays (host or
 *   device memory) are overwritten in Morton order, as the reference's permute! does.
 * The call returns once the caller's columns are consumed and checked for NaN (a NaN
 * coordinate is reported here); without host outputs or write-back the rest of the build
 * runs on the context's stream after the call returns, ordered before every later call on
 * the context (pmx_last_timings waits for it). */
int pmx_build_lbvh(struct pmx_ctx *ctx, const pmx_particles *p, int32_t flags,
                   uint64_t *codes_out, int64_t *order_out);

/* --- interpolation (uses the context's current LBVH) ------------------------------ */
/* ≙ the per-sample loop of PointSamples_interpolation (PointSamples_interpolation.jl:34-185):
 * nearest-h (LinearBVHQueries.jl:118) for gather / symmetric, then
 * _general_quantity_interpolate_kernel (general_interpolation.jl:1-669). */
int pmx_interpolate_points(struct pmx_ctx *ctx, int32_t strategy, int32_t precision,
                           const pmx_catalog *cat, const pmx_points *pts, const pmx_output *out);

/* ≙ StructuredGrid_interpolation(backend, Cartesian, grid, ...) (StructuredGrid_interpolation.jl:46-64)
 * on the context's tree: the sample coordinates are generated on the device from the axes,
 * point i = ix + nx (iy + ny iz) taking (ax[ix], ay[iy], az[iz]) — the column-major order of
 * coordinate_grid(Cartesian, grid) (StructuredGrid.jl:126-141) and flatten (transform.jl:53-56);
 * the values are copies, so the samples equal the host flatten bit for bit.  Outputs have
 * nx ny nz entries in the same order (≙ restore_struct's reshape, transform.jl:107-116).
 * Otherwise as pmx_interpolate_points. */
int pmx_interpolate_cartesian_grid(struct pmx_ctx *ctx, int32_t strategy, int32_t precision,
                                   const pmx_catalog *cat, const pmx_axes *axes, const pmx_output *out);

/* ≙ LineSamples_interpolation (LineSamples_interpolation.jl:47-99) with
 * _line_integrated_quantities_interpolate_kernel (line_integrated_scalar_interpolation.jl:123-166).
 * Only PMX_SCATTER is accepted, as in the reference (LineSamples_interpolation.jl:48-52);
 * the catalog must hold scalars only. */
int pmx_interpolate_lines(struct pmx_ctx *ctx, int32_t strategy, int32_t precision,
                          const pmx_catalog *cat, const pmx_lines *lines, const pmx_output *out);

/* No counterpart in the reference.  For callers that cannot allocate device memory during the
 * call (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): the
 * PMX_PREC_F32C path of pmx_interpolate_lines on the context's current LBVH, with every work
 * buffer (staged lines, binning, sort keys, histograms, results before delivery) in a static
 * device array of about 128 MiB that each device holding the library keeps until the process
 * ends.  It calls no hipMalloc / hipFree and allocates no host memory.  Limits: at most 1048576
 * lines and 4 scalar columns (PMX_ERR_CAPACITY beyond), lines and outputs in device memory,
 * outputs PMX_F32 or PMX_F64; PMX_LINE_INV_H2 may be OR'd into precision as for
 * pmx_interpolate_lines.  Every call on one device shares the array: calls that overlap on a
 * device (other threads or contexts) race.  Synchronous on return. */
int pmx_interpolate_lines_noalloc(struct pmx_ctx *ctx, int32_t strategy, int32_t precision,
                                  const pmx_catalog *cat, const pmx_lines *lines, const pmx_output *out);

/* --- drivers: build + interpolate in one call ------------------------------------- */
/* ≙ Partia.PointSamples_interpolation(backend, grid, input, catalog, strategy)
 * (ext/CUDAExt/.../PointSamples_interpolation.jl:12-29): LBVH build (with write-back of
 * the Morton-ordered columns into the caller's arrays), interpolation, synchronise. */
int pmx_point_samples_interpolation(struct pmx_ctx *ctx, const pmx_particles *p,
                                    int32_t strategy, int32_t precision, const pmx_catalog *cat,
                                    const pmx_points *pts, const pmx_output *out);
/* ≙ Partia.LineSamples_interpolation (ext/CUDAExt/.../LineSamples_interpolation.jl:11-34). */
int pmx_line_samples_interpolation(struct pmx_ctx *ctx, const pmx_particles *p,
                                   int32_t strategy, int32_t precision, const pmx_catalog *cat,
                                   const pmx_lines *lines, const pmx_output *out);

/* --- inspection hooks (tests; ≙ fields of BinaryRadixTree / LinearBVH) ------------- */
/* Unified node ids as in the reference (internal 1..n-1, leaf k ↦ n-1+k, 0 = none);
 * arrays have 2n-1 entries, entry j describing unified id j+1 (BinaryRadixTree.jl:8-15). */
int pmx_export_tree(struct pmx_ctx *ctx, int32_t *left, int32_t *right, int32_t *parent,
                    int32_t *escape);
/* node AABB min/max (3 x (n-1) each, FP64) and node_hmax (n-1) (LinearBVH.jl:20-26). */
int pmx_export_nodes(struct pmx_ctx *ctx, double *node_min, double *node_max, double *node_hmax);
/* Morton-ordered particle columns held by the context (FP64): x,y,z,m,h,rho then quant. */
int pmx_export_particles(struct pmx_ctx *ctx, double *const *cols, int32_t ncols);
/* Nearest leaf (0-based sorted index) of each point (≙ LBVH_find_nearest, LinearBVHQueries.jl:79). */
int pmx_nearest(struct pmx_ctx *ctx, int32_t precision, const pmx_points *pts, int64_t *idx_out);
/* Accepted neighbour sets of each point for a strategy (≙ the leaves the traversal macros
 * hand to the accumulation, LinearBVHTraversalMacros.jl:26-371), CSR in increasing sorted
 * index.  counts (n) is always written; offsets (n+1) and idx (capacity entries) when
 * non-null.  ha (n, nullable, host FP64): query smoothing lengths for gather / symmetric;
 * null means nearest-h.  Returns PMX_ERR_CAPACITY if idx is too small (counts are valid). */
int pmx_point_neighbors(struct pmx_ctx *ctx, int32_t strategy, int32_t precision,
                        const pmx_points *pts, const double *ha, int64_t *counts,
                        int64_t *offsets, int32_t *idx, int64_t capacity);
/* Same for lines (LinearBVHTraversalMacros.jl:408-796). */
int pmx_line_neighbors(struct pmx_ctx *ctx, int32_t strategy, int32_t precision,
                       const pmx_lines *lines, const double *ha, int64_t *counts,
                       int64_t *offsets, int32_t *idx, int64_t capacity);

/* Visit statistics of the FP64 traversal (traffic model of the roofline report):
 * stats[5] = {nearest-h internal nodes, nearest-h leaves, internal nodes tested, leaves
 * tested, leaves accepted}, summed over all samples.  No reference counterpart (the
 * reference has no instrumentation); used by the benchmark only. */
int pmx_point_traversal_stats(struct pmx_ctx *ctx, int32_t strategy, const pmx_points *pts,
                              int64_t *stats);
int pmx_line_traversal_stats(struct pmx_ctx *ctx, const pmx_lines *lines, int64_t *stats);

/* Neighbour-set digest of the tiled line-map path (parallel or nearly parallel unit-direction
 * lines, the path pmx_interpolate_lines takes for PMX_PREC_F32C on such lines): for each line i (host
 * arrays of 3 n uint64) digest[i] = accepted count, digest[n + i] = sum of accepted sorted
 * particle indices, digest[2n + i] = sum of their squares (mod 2^64).  *used = 0 (digest
 * untouched) when the lines do not qualify.  No reference counterpart: a check that the
 * tiled path accepts exactly the reference's pairs (LineSamples neighbour sets,
 * line_integrated_scalar_interpolation.jl:123-166) at sizes where full lists are too big. */
int pmx_line_neighbor_digest(struct pmx_ctx *ctx, const pmx_lines *lines, uint64_t *digest, int32_t *used);

/* --- kernel functions (host-callable, evaluated on the device) --------------------- */
/* f(q), f'(q) of a kernel on the device in FP64 (kernels/{M4..C6}.jl, (::K)(q) and (::_dK)(q)). */
int pmx_kernel_eval(struct pmx_ctx *ctx, int32_t kernel, int64_t n, const double *q,
                    double *f, double *df);
/* The regenerated line-integrated table of a kernel: Q and Inorm, 4096 entries each
 * (≙ _K_Q / _K_Inorm of line_integrated_kernel_tables.jl).  Host only; no context. */
int pmx_line_kernel_table(int32_t kernel, double *Q, double *Inorm);

/* --- timing ------------------------------------------------------------------------ */
/* Device time (ms, HIP events on the context stream) of the phases of the last build /
 * interpolation call: [0] build total, [1] encode+sort, [2] tree+refit, [3] column gather,
 * [4] nearest-h (unused: included in [5]), [5] interpolation phase (sample ordering and
 * kernels), [6] last interpolation call total, [7] interpolation kernels only; and two
 * counts of the last FP32c gather / symmetric point pass: [8] heavy sample tiles split,
 * [9] subtree items they queued.  n <= 10. */
int pmx_last_timings(const struct pmx_ctx *ctx, double *ms, int32_t n);

/* --- batches of snapshots ------------------------------------------------------------ */
/* No counterpart in the reference, which interpolates one snapshot per call and rebuilds the
 * LBVH each time (initialize_interpolation.jl:49-65).  For callers with many independent
 * snapshots and samples of uneven size (cylindrical grids, slices, column maps): a batch keeps
 * one LBVH per job resident, builds all of them, then runs every job's sample tiles in one
 * persistent kernel fed by a device work queue; builds and kernel can be captured into one
 * hipGraph and replayed, and pmx_batch_tune picks launch shapes by timing them.
 * Point jobs take any catalog (scalars, gradients, divergences, curls), split as in
 * pmx_interpolate_points; line jobs take scalars only (as pmx_interpolate_lines).
 * Restrictions: the FP32-compensated path, one SPH kernel for all jobs,
 * particle columns / samples / outputs in device memory (FP32 outputs), line jobs must be one
 * parallel or nearly parallel unit-direction line set each (as the tiled path of
 * pmx_interpolate_lines takes them).  Job buffers are borrowed for the batch's lifetime; particle
 * values may change between runs (the line-map tolerances follow the rebuilt particle box), the
 * lines must not. */
typedef struct pmx_batch pmx_batch;
enum { PMX_JOB_POINTS = 0, PMX_JOB_LINES = 1 };
typedef struct pmx_batch_job {
    pmx_particles particles;   /* mem = PMX_MEM_DEVICE */
    int32_t kind;              /* PMX_JOB_POINTS or PMX_JOB_LINES */
    int32_t strategy;          /* points: PMX_GATHER / PMX_SCATTER / PMX_SYMMETRIC; lines: PMX_SCATTER */
    pmx_points points;         /* kind PMX_JOB_POINTS; mem = PMX_MEM_DEVICE */
    pmx_lines lines;           /* kind PMX_JOB_LINES; mem = PMX_MEM_DEVICE */
    pmx_catalog catalog;       /* points: any catalog; lines: scalars only */
    pmx_output out;            /* dtype PMX_F32, mem = PMX_MEM_DEVICE, one column per output in
                                  catalog order (n_scalar + 3 n_grad + n_div + 3 n_curl) */
} pmx_batch_job;
/* Plans the batch: one build per job (sizes every buffer), sample ordering, line binning. */
int pmx_batch_create(pmx_batch **out, int32_t device_ordinal, int32_t njobs, const pmx_batch_job *jobs);
/* Rebuilds every job's LBVH from its (possibly updated) columns and interpolates all jobs;
 * use_graph = 1 replays one captured hipGraph of the whole sequence (captured on first use). */
int pmx_batch_run(pmx_batch *b, int32_t use_graph);
/* The persistent interpolation kernel alone, on the current trees. */
int pmx_batch_interpolate(pmx_batch *b);
/* Times candidate launch shapes (waves per CU, leaf-run sizes, job order, heavy-tile split
 * budget) and keeps the fastest. */
int pmx_batch_tune(pmx_batch *b);
/* [0] last run, [1] builds of the last eager run, [2] last interpolation, [3] tiles; then the
 * chosen waves per CU, point leaf run, line leaf run, heavy-first order, split budget (leaf
 * tests before a gather / symmetric point tile hands its remaining subtrees to all
 * wavefronts; 0 = off).  n <= 9. */
int pmx_batch_info(const pmx_batch *b, double *vals, int32_t n);
const char *pmx_batch_last_error(const pmx_batch *b);
void pmx_batch_destroy(pmx_batch *b);

#ifdef __cplusplus
}
#endif
#endif
