// pmx_griddataset.cpp — HDF5 files of interpolated grids in the layout of Partia's GridDataset
// writer / reader (Partia.jl/src/julia/IO/gridsIO/write_grids.jl, read_grids.jl), on the HDF5 C
// library.  Host code only; the layout and the datatypes follow HDF5.jl's conventions for the
// values the reference writes:
//   * a String is a scalar dataset of a fixed-length, NUL-terminated UTF-8 string of its byte
//     length (at least 1); a Vector{String} a 1-D dataset of variable-length UTF-8 strings;
//   * Int is a 64-bit integer, Bool an 8-bit bitfield, Float32 / Float64 the native floats;
//   * an array is stored with its dimensions reversed, so that its column-major elements are the
//     dataset's row-major elements.
#include <hdf5.h>

#include <cstdio>
#include <cstring>
#include <memory>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/partia_griddataset.h"

namespace {

thread_local std::string g_err;

struct Fail {
    int code;
    std::string msg;
};

// Owner of one HDF5 identifier, closed with its own close function.
struct Hid {
    hid_t id = -1;
    herr_t (*close)(hid_t) = nullptr;
    Hid() = default;
    Hid(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
    Hid(const Hid&) = delete;
    Hid& operator=(const Hid&) = delete;
    Hid(Hid&& o) noexcept : id(o.id), close(o.close) { o.id = -1; }
    Hid& operator=(Hid&& o) noexcept {
        if (this != &o) {
            if (id >= 0 && close) close(id);
            id = o.id;
            close = o.close;
            o.id = -1;
        }
        return *this;
    }
    ~Hid() {
        if (id >= 0 && close) close(id);
    }
    operator hid_t() const { return id; }
};

void check(bool ok, int code, const std::string& msg) {
    if (!ok) throw Fail{code, msg};
}
hid_t must(hid_t id, const std::string& what) {
    check(id >= 0, PMX_GD_ERR_IO, "HDF5: " + what);
    return id;
}

template <class F>
int guard(F&& body) {
    try {
        static const bool quiet = [] {
            H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);   // errors become statuses, not stderr text
            return true;
        }();
        (void)quiet;
        return body();
    } catch (const Fail& f) {
        g_err = f.msg;
        return f.code;
    } catch (const std::bad_alloc&) {
        g_err = "out of host memory";
        return PMX_GD_ERR_INTERNAL;
    } catch (const std::exception& e) {
        g_err = e.what();
        return PMX_GD_ERR_INTERNAL;
    } catch (...) {
        g_err = "unknown exception";
        return PMX_GD_ERR_INTERNAL;
    }
}

const char* const kCoordinate[3] = {"x", "y", "z"};   // _COORDINATE, write_grids.jl:1

hid_t float_type(int dtype) { return dtype == PMX_GD_F32 ? H5T_NATIVE_FLOAT : H5T_NATIVE_DOUBLE; }

// Fixed-length UTF-8 string type of n bytes (HDF5.jl's datatype(::String)).
Hid fixed_string(size_t n) {
    Hid t(must(H5Tcopy(H5T_C_S1), "string type"), H5Tclose);
    must(H5Tset_size(t, n < 1 ? 1 : n), "string size");
    must(H5Tset_cset(t, H5T_CSET_UTF8), "string cset");
    must(H5Tset_strpad(t, H5T_STR_NULLTERM), "string padding");
    return t;
}

Hid vlen_string() {
    Hid t(must(H5Tcopy(H5T_C_S1), "string type"), H5Tclose);
    must(H5Tset_size(t, H5T_VARIABLE), "variable string size");
    must(H5Tset_cset(t, H5T_CSET_UTF8), "string cset");
    return t;
}

void write_dataset(hid_t parent, const char* name, hid_t type, hid_t space, const void* buf) {
    Hid ds(must(H5Dcreate2(parent, name, type, space, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
                std::string("create dataset ") + name),
           H5Dclose);
    must(H5Dwrite(ds, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf), std::string("write dataset ") + name);
}

void write_vector(hid_t parent, const char* name, int dtype, int64_t n, const void* v) {
    const hsize_t dims[1] = {(hsize_t)n};
    Hid sp(must(H5Screate_simple(1, dims, nullptr), "dataspace"), H5Sclose);
    write_dataset(parent, name, float_type(dtype), sp, v);
}

Hid group(hid_t parent, const char* name) {
    return Hid(must(H5Gcreate2(parent, name, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), std::string("create group ") + name),
               H5Gclose);
}

}  // namespace

struct pmx_gd_file {
    Hid file;
    int grid_type = 0, dtype = 0, D = 0, L = 0;
    int64_t n[3] = {0, 0, 0};
    std::vector<std::string> vec_names;     // dataset names of the vectors, in x, y, z order
    std::vector<std::string> names;
    std::vector<std::string> param_keys;
};

extern "C" {

const char* pmx_gd_last_error(void) { return g_err.c_str(); }

int pmx_gd_write(const char* filename, const pmx_gd_desc* d) {
    return guard([&]() -> int {
        check(filename && d, PMX_GD_ERR_ARGUMENT, "null argument");
        check(d->grid_type == PMX_GD_POINTSAMPLES || d->grid_type == PMX_GD_STRUCTURED, PMX_GD_ERR_ARGUMENT,
              "grid_type must be PointSamples or StructuredGrid");
        check(d->dtype == PMX_GD_F32 || d->dtype == PMX_GD_F64, PMX_GD_ERR_ARGUMENT, "dtype must be Float32 or Float64");
        check(d->D >= 1 && d->D <= 3, PMX_GD_ERR_ARGUMENT, "D must be 1, 2 or 3");
        check(d->L >= 1 && d->names && d->grids, PMX_GD_ERR_ARGUMENT, "at least one grid with its name");
        check(d->nparams >= 0 && (d->nparams == 0 || d->params), PMX_GD_ERR_ARGUMENT, "null params");
        int64_t total = 1;
        for (int k = 0; k < d->D; ++k) {
            const int64_t nk = d->grid_type == PMX_GD_POINTSAMPLES ? d->n[0] : d->n[k];
            check(nk >= 0 && d->vec[k], PMX_GD_ERR_ARGUMENT, std::string("vector ") + kCoordinate[k] + ": null or negative length");
            if (d->grid_type == PMX_GD_STRUCTURED) total *= nk;
        }
        if (d->grid_type == PMX_GD_POINTSAMPLES) total = d->n[0];
        for (int i = 0; i < d->L; ++i)
            check(d->names[i] && (d->grids[i] || total == 0), PMX_GD_ERR_ARGUMENT, "grid " + std::to_string(i + 1) + ": null name or values");

        Hid f(H5Fcreate(filename, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
        check(f.id >= 0, PMX_GD_ERR_IO, std::string("cannot create ") + filename);
        // 1) params
        {
            Hid pg = group(f, "params");
            Hid scalar(must(H5Screate(H5S_SCALAR), "scalar dataspace"), H5Sclose);
            for (int j = 0; j < d->nparams; ++j) {
                const pmx_gd_param& p = d->params[j];
                check(p.key && *p.key, PMX_GD_ERR_ARGUMENT, "parameter " + std::to_string(j) + ": empty key");
                switch (p.kind) {
                case PMX_GD_STRING: {
                    check(p.s != nullptr, PMX_GD_ERR_ARGUMENT, std::string("parameter ") + p.key + ": null string");
                    const size_t len = std::strlen(p.s);
                    Hid t = fixed_string(len);
                    std::string buf(p.s, len);
                    buf.resize(len < 1 ? 1 : len, '\0');
                    write_dataset(pg, p.key, t, scalar, buf.data());
                    break;
                }
                case PMX_GD_INT:
                    write_dataset(pg, p.key, H5T_NATIVE_INT64, scalar, &p.i);
                    break;
                case PMX_GD_BOOL: {
                    const uint8_t b = p.i ? 1 : 0;
                    write_dataset(pg, p.key, H5T_NATIVE_B8, scalar, &b);
                    break;
                }
                case PMX_GD_FLOAT:
                    if (d->dtype == PMX_GD_F32) {
                        const float v = (float)p.f;
                        write_dataset(pg, p.key, H5T_NATIVE_FLOAT, scalar, &v);
                    } else {
                        write_dataset(pg, p.key, H5T_NATIVE_DOUBLE, scalar, &p.f);
                    }
                    break;
                default:
                    throw Fail{PMX_GD_ERR_ARGUMENT, std::string("parameter ") + p.key + ": unknown kind"};
                }
            }
        }
        // 2) data: names, shared vectors, grids
        Hid dg = group(f, "data");
        {
            Hid t = vlen_string();
            const hsize_t dims[1] = {(hsize_t)d->L};
            Hid sp(must(H5Screate_simple(1, dims, nullptr), "dataspace"), H5Sclose);
            write_dataset(dg, "names", t, sp, d->names);
        }
        {
            Hid vg = group(dg, d->grid_type == PMX_GD_POINTSAMPLES ? "coord" : "axes");
            for (int k = 0; k < d->D; ++k)
                write_vector(vg, kCoordinate[k], d->dtype, d->grid_type == PMX_GD_POINTSAMPLES ? d->n[0] : d->n[k], d->vec[k]);
        }
        {
            Hid gg = group(dg, "grids");
            hsize_t dims[3];
            int rank;
            if (d->grid_type == PMX_GD_POINTSAMPLES) {
                rank = 1;
                dims[0] = (hsize_t)d->n[0];
            } else {
                rank = d->D;
                for (int k = 0; k < d->D; ++k) dims[k] = (hsize_t)d->n[d->D - 1 - k];   // reversed: column-major
            }
            Hid sp(must(H5Screate_simple(rank, dims, nullptr), "dataspace"), H5Sclose);
            for (int i = 0; i < d->L; ++i)
                write_dataset(gg, std::to_string(i + 1).c_str(), float_type(d->dtype), sp, d->grids[i]);
        }
        must(H5Fflush(f, H5F_SCOPE_LOCAL), "flush");
        return PMX_GD_OK;
    });
}

namespace {

// Link names of a group, in the library's name order.
std::vector<std::string> keys_of(hid_t g) {
    std::vector<std::string> out;
    H5G_info_t info;
    must(H5Gget_info(g, &info), "group info");
    for (hsize_t i = 0; i < info.nlinks; ++i) {
        const ssize_t len = H5Lget_name_by_idx(g, ".", H5_INDEX_NAME, H5_ITER_INC, i, nullptr, 0, H5P_DEFAULT);
        must((hid_t)len, "link name");
        std::string s((size_t)len, '\0');
        H5Lget_name_by_idx(g, ".", H5_INDEX_NAME, H5_ITER_INC, i, &s[0], (size_t)len + 1, H5P_DEFAULT);
        out.push_back(s);
    }
    return out;
}

bool has_link(hid_t g, const char* name) { return H5Lexists(g, name, H5P_DEFAULT) > 0; }

// A scalar string dataset (fixed-length or variable-length) as a std::string.
std::string read_string(hid_t parent, const std::string& name) {
    Hid ds(must(H5Dopen2(parent, name.c_str(), H5P_DEFAULT), "open " + name), H5Dclose);
    Hid ft(must(H5Dget_type(ds), "type of " + name), H5Tclose);
    check(H5Tget_class(ft) == H5T_STRING, PMX_GD_ERR_ARGUMENT, "params: " + name + " is not a string");
    if (H5Tis_variable_str(ft) > 0) {
        Hid mt = vlen_string();
        char* p = nullptr;
        must(H5Dread(ds, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, &p), "read " + name);
        std::string s = p ? p : "";
        Hid sp(must(H5Dget_space(ds), "space"), H5Sclose);
        H5Dvlen_reclaim(mt, sp, H5P_DEFAULT, &p);
        return s;
    }
    const size_t n = H5Tget_size(ft);
    Hid mt = fixed_string(n);
    std::string s(n, '\0');
    must(H5Dread(ds, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, &s[0]), "read " + name);
    const size_t z = s.find('\0');
    if (z != std::string::npos) s.resize(z);
    return s;
}

int64_t dataset_length(hid_t parent, const std::string& name) {
    Hid ds(must(H5Dopen2(parent, name.c_str(), H5P_DEFAULT), "open " + name), H5Dclose);
    Hid sp(must(H5Dget_space(ds), "space of " + name), H5Sclose);
    const hssize_t n = H5Sget_simple_extent_npoints(sp);
    must((hid_t)n, "extent of " + name);
    return (int64_t)n;
}

void copy_out(const std::string& s, char* buf, size_t cap, size_t* len) {
    if (len) *len = s.size();
    if (buf && cap > 0) {
        const size_t k = s.size() < cap - 1 ? s.size() : cap - 1;
        std::memcpy(buf, s.data(), k);
        buf[k] = '\0';
    }
}

}  // namespace

int pmx_gd_open(const char* filename, pmx_gd_file** out) {
    return guard([&]() -> int {
        check(filename && out, PMX_GD_ERR_ARGUMENT, "null argument");
        *out = nullptr;
        auto* f = new pmx_gd_file();
        std::unique_ptr<pmx_gd_file> own(f);
        f->file = Hid(H5Fopen(filename, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
        check(f->file.id >= 0, PMX_GD_ERR_IO, std::string("cannot open ") + filename);
        check(has_link(f->file, "params"), PMX_GD_ERR_ARGUMENT, "Schema error: missing /params");
        check(has_link(f->file, "data"), PMX_GD_ERR_ARGUMENT, "Schema error: missing /data");
        Hid pg(must(H5Gopen2(f->file, "params", H5P_DEFAULT), "open /params"), H5Gclose);
        Hid dg(must(H5Gopen2(f->file, "data", H5P_DEFAULT), "open /data"), H5Gclose);
        f->param_keys = keys_of(pg);
        check(has_link(pg, "dtype"), PMX_GD_ERR_ARGUMENT, "params missing :dtype");
        check(has_link(pg, "grid_type"), PMX_GD_ERR_ARGUMENT, "params missing :grid_type");
        std::string dt = read_string(pg, "dtype");
        {   // strip
            const size_t a = dt.find_first_not_of(" \t\r\n"), b = dt.find_last_not_of(" \t\r\n");
            dt = a == std::string::npos ? std::string() : dt.substr(a, b - a + 1);
        }
        if (dt == "Float32") f->dtype = PMX_GD_F32;
        else if (dt == "Float64") f->dtype = PMX_GD_F64;
        else throw Fail{PMX_GD_ERR_ARGUMENT, "Unsupported dtype string in params[:dtype]: " + dt};
        const std::string gt = read_string(pg, "grid_type");
        std::string vgname;
        if (gt == "StructuredGrid") {
            f->grid_type = PMX_GD_STRUCTURED;
            vgname = "axes";
        } else if (gt == "PointSamples") {
            f->grid_type = PMX_GD_POINTSAMPLES;
            vgname = "coord";
        } else {
            throw Fail{PMX_GD_ERR_ARGUMENT, "Unsupported params[:grid_type] = " + gt};
        }
        // names
        check(has_link(dg, "names"), PMX_GD_ERR_ARGUMENT, "Schema error: missing data/names");
        {
            Hid ds(must(H5Dopen2(dg, "names", H5P_DEFAULT), "open data/names"), H5Dclose);
            Hid sp(must(H5Dget_space(ds), "space of data/names"), H5Sclose);
            const hssize_t n = H5Sget_simple_extent_npoints(sp);
            Hid ft(must(H5Dget_type(ds), "type of data/names"), H5Tclose);
            check(H5Tget_class(ft) == H5T_STRING, PMX_GD_ERR_ARGUMENT, "data/names are not strings");
            if (H5Tis_variable_str(ft) > 0) {
                Hid mt = vlen_string();
                std::vector<char*> p((size_t)n, nullptr);
                must(H5Dread(ds, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, p.data()), "read data/names");
                for (char* s : p) f->names.push_back(s ? s : "");
                H5Dvlen_reclaim(mt, sp, H5P_DEFAULT, p.data());
            } else {
                const size_t w = H5Tget_size(ft);
                Hid mt = fixed_string(w);
                std::vector<char> buf(w * (size_t)n);
                must(H5Dread(ds, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf.data()), "read data/names");
                for (hssize_t i = 0; i < n; ++i) {
                    std::string s(buf.data() + w * i, w);
                    const size_t z = s.find('\0');
                    if (z != std::string::npos) s.resize(z);
                    f->names.push_back(s);
                }
            }
        }
        // shared vectors (_read_shared_vectors, _ordered_keys: x, y, z when all three are there,
        // else the keys sorted)
        check(has_link(dg, vgname.c_str()), PMX_GD_ERR_ARGUMENT, "Schema error: missing data/" + vgname);
        {
            Hid vg(must(H5Gopen2(dg, vgname.c_str(), H5P_DEFAULT), "open data/" + vgname), H5Gclose);
            std::vector<std::string> ks = keys_of(vg);   // name order = sorted
            const bool xyz = has_link(vg, "x") && has_link(vg, "y") && has_link(vg, "z");
            if (xyz) {
                ks.assign(kCoordinate, kCoordinate + (ks.size() < 3 ? ks.size() : 3));
            }
            check(!ks.empty() && ks.size() <= 3, PMX_GD_ERR_ARGUMENT, "data/" + vgname + ": 1 to 3 vectors expected");
            f->D = (int)ks.size();
            f->vec_names = ks;
            for (int k = 0; k < f->D; ++k) f->n[k] = dataset_length(vg, ks[(size_t)k]);
        }
        // grids: "1" .. "L"
        check(has_link(dg, "grids"), PMX_GD_ERR_ARGUMENT, "Schema error: missing data/grids");
        {
            Hid gg(must(H5Gopen2(dg, "grids", H5P_DEFAULT), "open data/grids"), H5Gclose);
            H5G_info_t info;
            must(H5Gget_info(gg, &info), "group info");
            f->L = (int)info.nlinks;
            int64_t want = 1;
            if (f->grid_type == PMX_GD_POINTSAMPLES) want = f->n[0];
            else for (int k = 0; k < f->D; ++k) want *= f->n[k];
            for (int i = 0; i < f->L; ++i) {
                const std::string nm = std::to_string(i + 1);
                check(has_link(gg, nm.c_str()), PMX_GD_ERR_ARGUMENT, "Schema error: missing data/grids/" + nm);
                check(dataset_length(gg, nm) == want, PMX_GD_ERR_ARGUMENT,
                      "data/grids/" + nm + ": size does not match the shared vectors");
            }
        }
        check((int)f->names.size() == f->L, PMX_GD_ERR_ARGUMENT, "names length != number of grids");
        *out = own.release();
        return PMX_GD_OK;
    });
}

int pmx_gd_get_info(const pmx_gd_file* f, pmx_gd_info* info) {
    return guard([&]() -> int {
        check(f && info, PMX_GD_ERR_ARGUMENT, "null argument");
        info->grid_type = f->grid_type;
        info->dtype = f->dtype;
        info->D = f->D;
        for (int k = 0; k < 3; ++k) info->n[k] = k < f->D ? f->n[k] : 0;
        info->L = f->L;
        info->nparams = (int32_t)f->param_keys.size();
        return PMX_GD_OK;
    });
}

int pmx_gd_name(const pmx_gd_file* f, int32_t i, char* buf, size_t cap, size_t* len) {
    return guard([&]() -> int {
        check(f && i >= 0 && i < (int32_t)f->names.size(), PMX_GD_ERR_ARGUMENT, "grid index out of range");
        copy_out(f->names[(size_t)i], buf, cap, len);
        return PMX_GD_OK;
    });
}

int pmx_gd_get_param(const pmx_gd_file* f, int32_t j, char* key_buf, size_t key_cap, size_t* key_len, int32_t* kind,
                     char* str_buf, size_t str_cap, size_t* str_len, int64_t* ival, double* fval) {
    return guard([&]() -> int {
        check(f && j >= 0 && j < (int32_t)f->param_keys.size() && kind, PMX_GD_ERR_ARGUMENT, "parameter index out of range");
        const std::string& key = f->param_keys[(size_t)j];
        copy_out(key, key_buf, key_cap, key_len);
        Hid pg(must(H5Gopen2(f->file, "params", H5P_DEFAULT), "open /params"), H5Gclose);
        Hid ds(must(H5Dopen2(pg, key.c_str(), H5P_DEFAULT), "open params/" + key), H5Dclose);
        Hid ft(must(H5Dget_type(ds), "type of params/" + key), H5Tclose);
        const H5T_class_t cls = H5Tget_class(ft);
        if (cls == H5T_STRING) {
            *kind = PMX_GD_STRING;
            copy_out(read_string(pg, key), str_buf, str_cap, str_len);
        } else if (cls == H5T_INTEGER) {
            *kind = PMX_GD_INT;
            int64_t v = 0;
            must(H5Dread(ds, H5T_NATIVE_INT64, H5S_ALL, H5S_ALL, H5P_DEFAULT, &v), "read params/" + key);
            if (ival) *ival = v;
        } else if (cls == H5T_BITFIELD || cls == H5T_ENUM) {
            *kind = PMX_GD_BOOL;
            uint8_t v = 0;
            Hid mt(must(H5Tcopy(cls == H5T_BITFIELD ? H5T_NATIVE_B8 : H5T_NATIVE_UINT8), "bool type"), H5Tclose);
            if (cls == H5T_ENUM) {
                // an enum stored by another writer: read its integer value
                must(H5Dread(ds, H5T_NATIVE_UINT8, H5S_ALL, H5S_ALL, H5P_DEFAULT, &v), "read params/" + key);
            } else {
                must(H5Dread(ds, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, &v), "read params/" + key);
            }
            if (ival) *ival = v ? 1 : 0;
        } else if (cls == H5T_FLOAT) {
            *kind = PMX_GD_FLOAT;
            double v = 0;
            must(H5Dread(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, &v), "read params/" + key);
            if (fval) *fval = v;
        } else {
            throw Fail{PMX_GD_ERR_ARGUMENT, "params/" + key + ": unsupported datatype"};
        }
        return PMX_GD_OK;
    });
}

int pmx_gd_read_vector(const pmx_gd_file* f, int32_t d, void* out) {
    return guard([&]() -> int {
        check(f && out && d >= 0 && d < f->D, PMX_GD_ERR_ARGUMENT, "vector index out of range");
        Hid vg(must(H5Gopen2(f->file, f->grid_type == PMX_GD_POINTSAMPLES ? "data/coord" : "data/axes", H5P_DEFAULT),
                    "open vectors"),
               H5Gclose);
        Hid ds(must(H5Dopen2(vg, f->vec_names[(size_t)d].c_str(), H5P_DEFAULT), "open vector"), H5Dclose);
        must(H5Dread(ds, float_type(f->dtype), H5S_ALL, H5S_ALL, H5P_DEFAULT, out), "read vector");
        return PMX_GD_OK;
    });
}

int pmx_gd_read_grid(const pmx_gd_file* f, int32_t i, void* out) {
    return guard([&]() -> int {
        check(f && out && i >= 0 && i < f->L, PMX_GD_ERR_ARGUMENT, "grid index out of range");
        Hid gg(must(H5Gopen2(f->file, "data/grids", H5P_DEFAULT), "open data/grids"), H5Gclose);
        Hid ds(must(H5Dopen2(gg, std::to_string(i + 1).c_str(), H5P_DEFAULT), "open grid"), H5Dclose);
        must(H5Dread(ds, float_type(f->dtype), H5S_ALL, H5S_ALL, H5P_DEFAULT, out), "read grid");
        return PMX_GD_OK;
    });
}

void pmx_gd_close(pmx_gd_file* f) { delete f; }

}  // extern "C"
