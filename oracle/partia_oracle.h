/*
 * partia_oracle — CPU FP64 restatement of Partia.jl's LBVH neighbour search and SPH
 * interpolation path.  TEST INFRASTRUCTURE ONLY: this library is the checker for the
 * HIP implementation and the CPU baseline that bench.py times beside it.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The product
 * library (libpartia_mi355x.so) never links or calls it.
 *
 * Every function follows a reference file named in its comment (paths relative to the
 * reference repository; "P/" = Partia.jl/, "NS/" = P/src/julia/NeighborSearch/,
 * "KI/" = P/src/julia/KernelInterpolation/).  Arithmetic is IEEE binary64 with
 * -ffp-contract=off, in the operation order the Julia source evaluates (Julia does not
 * contract a*b+c unless muladd/fma is written explicitly).
 *
 * Parity pinning: see the header of partia_oracle.c for which functions are pinned by
 * which check.
 */
#ifndef PARTIA_ORACLE_H
#define PARTIA_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_M4 = 0, ORC_M5 = 1, ORC_M6 = 2, ORC_C2 = 3, ORC_C4 = 4, ORC_C6 = 5 };
enum { ORC_GATHER = 0, ORC_SCATTER = 1, ORC_SYMMETRIC = 2 };

typedef struct orc_lbvh {
    int64_t n;
    int32_t root;
    uint64_t *codes;   /* n, sorted and de-duplicated (MortonCode.jl:181-195)           */
    int64_t *order;    /* n, 1-based original particle index of each sorted slot         */
    int32_t *left, *right, *escape, *parent; /* 2n entries, indexed by unified id (0 unused) */
    double *lx, *ly, *lz, *lh;               /* leaf coordinates / smoothing lengths      */
    double *nminx, *nminy, *nminz, *nmaxx, *nmaxy, *nmaxz, *nhmax; /* n entries, internal id */
} orc_lbvh;

typedef struct orc_catalog {      /* InterpolationCatalogConcise (InterpolationCatalog.jl:26-32) */
    int n_scalar; const int32_t *scalar_slots; const uint8_t *scalar_shepard; /* 0-based slots */
    int n_grad;   const int32_t *grad_slots;   /* -1 = density gradient (reference slot 0)   */
    int n_div;    const int32_t *div_slots;    /* 3 per divergence                           */
    int n_curl;   const int32_t *curl_slots;   /* 3 per curl                                 */
} orc_catalog;

/* ---- Morton codes (MortonCode.jl) ---- */
uint64_t orc_expand_bits3(uint64_t x);
uint64_t orc_compact_bits3(uint64_t x);
uint64_t orc_encode3(uint64_t ix, uint64_t iy, uint64_t iz);
void orc_quantize_axis(int64_t n, const double *v, uint64_t *out);
void orc_morton_codes(int64_t n, const double *x, const double *y, const double *z, uint64_t *codes);

/* ---- LBVH build (InterpolationInput.jl:203-224 LinearBVH!) ---- */
orc_lbvh *orc_linear_bvh(int64_t n, double *x, double *y, double *z, double *m, double *h,
                         double *rho, int ncol, double **quant, int nthreads);
void orc_lbvh_free(orc_lbvh *t);
int orc_split_position(const uint64_t *codes, int64_t first, int64_t last);
void orc_find_range(const uint64_t *codes, int64_t n, int64_t i, int64_t *lo, int64_t *hi);

/* ---- SPH kernels (kernel_function/) ---- */
double orc_jl_pow(double x, int64_t n);
double orc_kernel_f(int k, double q);
double orc_kernel_df(int k, double q);
double orc_kernel_norm3(int k);
double orc_kernel_support(int k);
double orc_W(int k, double r, double h);
void orc_gradW(int k, double dx, double dy, double dz, double h, double *g);
void orc_line_table(int k, double *Q, double *Inorm);  /* 4096 entries each */
double orc_line_W(int k, double r, double h);

/* ---- queries (LinearBVHQueries.jl, traversal macros) ---- */
void orc_nearest(const orc_lbvh *t, int64_t npts, const double *qx, const double *qy,
                 const double *qz, int64_t *idx_out, double *h_out, int nthreads);
int64_t orc_point_neighbors(const orc_lbvh *t, int k, int strategy, int64_t npts,
                            const double *qx, const double *qy, const double *qz,
                            const double *ha, int64_t *offsets, int32_t *idx, int64_t cap);
int64_t orc_line_neighbors(const orc_lbvh *t, int k, int strategy, int64_t nl,
                           const double *ox, const double *oy, const double *oz,
                           const double *dx, const double *dy, const double *dz,
                           const double *ha, int64_t *offsets, int32_t *idx, int64_t cap);
void orc_point_visits(const orc_lbvh *t, int k, int strategy, int64_t npts, const double *qx,
                      const double *qy, const double *qz, int64_t *n_internal, int64_t *n_leaf,
                      int64_t *n_accept, int nthreads);
void orc_line_visits(const orc_lbvh *t, int k, int64_t nl, const double *ox, const double *oy,
                     const double *oz, const double *dx, const double *dy, const double *dz,
                     int64_t *n_internal, int64_t *n_leaf, int64_t *n_accept, int nthreads);

/* ---- interpolation drivers (grid_interpolation/) ---- */
int orc_point_samples(const orc_lbvh *t, const double *m, const double *rho, int ncol,
                      const double *const *quant, int k, int strategy, const orc_catalog *cat,
                      int64_t npts, const double *qx, const double *qy, const double *qz,
                      double *const *out, int nthreads);
int orc_line_samples(const orc_lbvh *t, const double *m, const double *rho, int ncol,
                     const double *const *quant, int k, int strategy, const double *ha,
                     int nscal, const int32_t *slots, const uint8_t *shepard, int64_t nl,
                     const double *ox, const double *oy, const double *oz, const double *dx,
                     const double *dy, const double *dz, double *const *out, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
