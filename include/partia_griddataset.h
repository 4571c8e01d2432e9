/*
 * partia_griddataset.h — C ABI of libpartia_griddataset.so: HDF5 files of interpolated grids in
 * the layout of Partia's GridDataset writer and reader, so that a file written here is read by
 * Partia's read_GridDataset and a file written by Partia is read here.
 *
 * Replaces write_GridDataset (Partia.jl/src/julia/IO/gridsIO/write_grids.jl:29-74 for PointSamples,
 * :98-143 for StructuredGrid), the metadata of write_GridBundle (write_grids.jl:181-215) is filled
 * in by the caller (the Python mirror does), and read_GridDataset (read_grids.jl:29-86).
 *
 * File layout (HDF5.jl conventions of the reference's writer):
 *   /params/<key>       one scalar dataset per parameter: String (fixed-length UTF-8, NUL-
 *                       terminated), Int (64-bit integer), Bool (8-bit bitfield) or the grid's float
 *   /data/names         1-D variable-length UTF-8 strings, one per grid
 *   /data/coord/<x|y|z> PointSamples: the shared coordinate vectors (npts each)
 *   /data/axes/<x|y|z>  StructuredGrid: the shared axis vectors
 *   /data/grids/<i>     grid i (1-based name): npts values, or the column-major array of the axes'
 *                       lengths stored with its dimensions reversed (Julia's column-major order)
 *
 * Conventions: every entry point returns PMX_GD_OK (0) or a negative status and never throws across
 * the ABI; pmx_gd_last_error() describes the last failure of the calling thread.  All buffers are
 * host memory.
 */
#ifndef PARTIA_GRIDDATASET_H
#define PARTIA_GRIDDATASET_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    PMX_GD_OK = 0,
    PMX_GD_ERR_ARGUMENT = -1,   /* ≙ ArgumentError (schema errors, unshared vectors, bad values) */
    PMX_GD_ERR_IO = -8,         /* the file cannot be created, opened or written                  */
    PMX_GD_ERR_INTERNAL = -9    /* a C++ exception inside the library                              */
};
/* Grid types (params grid_type "PointSamples" / "StructuredGrid"). */
enum { PMX_GD_POINTSAMPLES = 0, PMX_GD_STRUCTURED = 1 };
/* Element types of vectors and grids (the GridDataset's TF). */
enum { PMX_GD_F32 = 0, PMX_GD_F64 = 1 };
/* Parameter kinds (Dict{Symbol, Union{String, Int, Bool, TF}}). */
enum { PMX_GD_STRING = 0, PMX_GD_INT = 1, PMX_GD_BOOL = 2, PMX_GD_FLOAT = 3 };

typedef struct pmx_gd_param {
    const char *key;
    int32_t kind;          /* PMX_GD_STRING | PMX_GD_INT | PMX_GD_BOOL | PMX_GD_FLOAT           */
    const char *s;         /* STRING                                                             */
    int64_t i;             /* INT; BOOL (0 / 1)                                                  */
    double f;              /* FLOAT (written in the dataset's element type)                     */
} pmx_gd_param;

typedef struct pmx_gd_desc {
    int32_t grid_type;     /* PMX_GD_POINTSAMPLES | PMX_GD_STRUCTURED                            */
    int32_t dtype;         /* PMX_GD_F32 | PMX_GD_F64                                            */
    int32_t D;             /* 1..3 vectors (x, y, z)                                             */
    int64_t n[3];          /* PointSamples: npts in n[0]; StructuredGrid: the axes' lengths      */
    const void *vec[3];    /* the shared coordinate (PointSamples) or axis (StructuredGrid) vectors */
    int32_t L;             /* number of grids (>= 1)                                             */
    const char *const *names;  /* L names                                                        */
    const void *const *grids;  /* L grids, each npts (PointSamples) or prod(n) values, column-major */
    int32_t nparams;
    const pmx_gd_param *params;
} pmx_gd_desc;

/* ≙ write_GridDataset(gd, filename) (write_grids.jl:29-74, 98-143): creates or truncates `filename`. */
int pmx_gd_write(const char *filename, const pmx_gd_desc *d);

/* ≙ read_GridDataset(filename) (read_grids.jl:29-86), in steps: open and check the schema (missing
 * /params or /data, params without dtype / grid_type, an unsupported dtype or grid_type, names of a
 * length other than the number of grids are ArgumentErrors with the reference's texts), then
 * query and copy. */
typedef struct pmx_gd_file pmx_gd_file;
typedef struct pmx_gd_info {
    int32_t grid_type;
    int32_t dtype;
    int32_t D;
    int64_t n[3];          /* vector lengths in x, y, z order                                    */
    int32_t L;
    int32_t nparams;
} pmx_gd_info;
int pmx_gd_open(const char *filename, pmx_gd_file **out);
int pmx_gd_get_info(const pmx_gd_file *f, pmx_gd_info *info);
/* Name of grid i (0-based) into buf (cap bytes, NUL-terminated); *len = its length in bytes. */
int pmx_gd_name(const pmx_gd_file *f, int32_t i, char *buf, size_t cap, size_t *len);
/* Parameter j (0-based, keys in HDF5 name order): key into key_buf, kind, and the value (a string
 * into str_buf; an Int / Bool into *ival; a float into *fval).  Lengths in *key_len / *str_len. */
int pmx_gd_get_param(const pmx_gd_file *f, int32_t j, char *key_buf, size_t key_cap, size_t *key_len, int32_t *kind,
                     char *str_buf, size_t str_cap, size_t *str_len, int64_t *ival, double *fval);
/* Vector d (0-based: x, y, z) and grid i (0-based, dataset "<i + 1>") in the file's dtype. */
int pmx_gd_read_vector(const pmx_gd_file *f, int32_t d, void *out);
int pmx_gd_read_grid(const pmx_gd_file *f, int32_t i, void *out);
void pmx_gd_close(pmx_gd_file *f);
const char *pmx_gd_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
