/*
 * partia_phantom.h — C ABI of the Phantom dump reader in libpartia_mi355x.so: the input side
 * of the interpolation path (ParticleIO.jl's read_phantom + build_input), streaming particle
 * columns from the file straight into host or device (HBM) buffers.
 *
 * Reference: ParticleIO.jl/src/julia/IO/phantomIO/read_phantom.jl (format parsing :20-225,
 * post-processing :273-321) and ParticleIO.jl/src/julia/Particles/add_quantities.jl:294-347
 * (add_rho!).  A Julia front end would keep ParticleDataFrame and call these entry points for the
 * file structure and the column bytes; the Python package mirrors it in particleio.py.
 *
 * Conventions are those of partia_mi355x.h (status codes, 0-based indices, `mem` selects host or
 * device memory for the destination).  A reader is single-threaded and holds the open file until
 * pmx_phantom_close.
 */
#ifndef PARTIA_PHANTOM_H
#define PARTIA_PHANTOM_H

#include "partia_mi355x.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Element types of dump columns and header variables, beside PMX_F32 / PMX_F64. */
enum { PMX_I8 = 10, PMX_I16 = 11, PMX_I32 = 12, PMX_I64 = 13 };

typedef struct pmx_phantom pmx_phantom;

/* Open and parse a dump: capture pattern (default integer / real sizes), file identifier,
 * global header and the array-block directory; no column data is read.
 * ≙ _read_capture_pattern, _read_file_identifier, _read_global_header, _read_array_blocks
 *   (read_phantom.jl:20-81, 83-99, 101-148, 150-207).  Errors: PMX_ERR_IO (open / stat),
 *   PMX_ERR_FORMAT (capture pattern or Fortran record tags mismatch, truncated file),
 *   PMX_ERR_DIMENSION (columns of unequal length in one data frame).  On error *out is NULL and
 *   pmx_phantom_open_error() describes the failure. */
int pmx_phantom_open(pmx_phantom **out, const char *path);
const char *pmx_phantom_open_error(void);   /* thread-local; valid until the next open */
void pmx_phantom_close(pmx_phantom *ph);
const char *pmx_phantom_last_error(const pmx_phantom *ph);

/* vals[0] bytes of the default integer (4), vals[1] bytes of the default real (4 or 8),
 * vals[2] array blocks, vals[3] columns, vals[4] header variables; n <= 5. */
int pmx_phantom_info(const pmx_phantom *ph, int64_t *vals, int32_t n);
/* The 100-character file identifier, whitespace-stripped, NUL-terminated in buf[cap]. */
int pmx_phantom_file_identifier(const pmx_phantom *ph, char *buf, int32_t cap);
/* Particles of array block `block` (0-based; block 1 holds the sink particles). */
int pmx_phantom_block_size(const pmx_phantom *ph, int32_t block, int64_t *n);
/* Column `col` (file order, a later column of the same name in the same data frame replaces the
 * earlier one in place, as df[!, tag] = data does): its tag (<= 16 characters, name[17]), its
 * array block and its element type. */
int pmx_phantom_column_info(const pmx_phantom *ph, int32_t col, char *name, int32_t *block, int32_t *dtype);
/* Header variable i (name[24]: a duplicate tag within one type block is renamed tag_2, tag_3, ...;
 * a later type block replaces an earlier variable of the same name).  Integers are returned in
 * *ival, reals in *fval (FP32 values widened exactly). */
int pmx_phantom_header_var(const pmx_phantom *ph, int32_t i, char *name, int32_t *dtype, double *fval,
                           int64_t *ival);
/* Read column `col` converted to out_dtype (numeric conversion, round to nearest; a real column
 * cannot be read as an integer type) into dst, host or device memory per `mem`.  keep: NULL for
 * every row, else nkeep strictly increasing 0-based row indices (host memory) — the rows a
 * filtered or grouped data frame keeps (read_phantom.jl:278-320).  Device destinations are
 * filled through pinned staging buffers and asynchronous copies, overlapped with the file reads;
 * the call returns after the last copy. */
int pmx_phantom_read_column(pmx_phantom *ph, int32_t col, int32_t out_dtype, int32_t mem, void *dst,
                            const int64_t *keep, int64_t nkeep);

/* add_rho! (add_quantities.jl:331-347) through _density (:294-320), all in TF = dtype (the mass
 * type; m and hfact are TF values passed widened): with c = m hfact^dim and invh = 1 / h_i,
 * rho_i = c invh invh (dim 2) or c invh invh invh (dim 3), each product rounded in TF.  h and
 * rho are TF arrays in `mem` memory; for mem = PMX_MEM_DEVICE the work runs on the current
 * device and the call returns when it is done. */
int pmx_phantom_density(int32_t dtype, int32_t dim, int32_t mem, int64_t n, const void *h, double m, double hfact,
                        void *rho);

#ifdef __cplusplus
}
#endif

#endif /* PARTIA_PHANTOM_H */
