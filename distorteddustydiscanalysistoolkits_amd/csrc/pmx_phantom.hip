// pmx_phantom.hip — Phantom native binary dump reader (the input side of the interpolation path).
//
// A dump is a sequence of Fortran unformatted records, each framed by a 4-byte tag before and
// after its payload (read_phantom.jl:83-93).  The reader parses the structure once at open —
// capture pattern, file identifier, the eight typed global-header blocks and the array-block
// directory — and records the file offset of every column's payload; column bytes are read only
// when a column is requested, with positioned reads (pread) of large chunks.  A device
// destination is filled through two pinned staging buffers: while one chunk's host-to-device
// copy runs on the reader's stream, the next chunk is read and converted on the host.
//
// Behaviour follows the reference reader, including its failure modes:
//   * the capture pattern tries a 4-byte default integer, and on mismatch rewinds 8 bytes and
//     reads an 8-byte integer there (read_phantom.jl:33-43; the rewind lands on the record's
//     leading tag, so dumps written with 8-byte default integers are rejected, as the reference
//     rejects them);
//   * a default real is FP32 unless the FP32 check fails, then the 8 bytes are re-read as FP64;
//   * header tags are whitespace-stripped, duplicates within one type block are renamed
//     tag_2, tag_3, ... and a later block's variable replaces an earlier one of the same name;
//   * array block 2 (index 1) holds the sink particles; every other block adds its columns to
//     the main data frame, a repeated tag replacing the column in place.
#include <hip/hip_runtime.h>

#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cctype>
#include <cerrno>
#include <cstdint>
#include <cstring>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

#include "../../include/partia_phantom.h"
#include "pmx_internal.h"

namespace {

struct PhColumn {
    std::string name;
    int block;          // array block (0-based)
    int dtype;          // PMX_I8 .. PMX_I64, PMX_F32, PMX_F64
    int64_t n;          // rows
    int64_t offset;     // file offset of the payload
};

struct PhVar {
    std::string name;
    int dtype;
    double fval;
    int64_t ival;
};

int dtype_bytes(int t) {
    switch (t) {
    case PMX_I8: return 1;
    case PMX_I16: return 2;
    case PMX_I32: case PMX_F32: return 4;
    default: return 8;
    }
}

bool is_real(int t) { return t == PMX_F32 || t == PMX_F64; }

std::string strip(const char* p, size_t n) {
    size_t a = 0, b = n;
    while (a < b && std::isspace((unsigned char)p[a])) ++a;
    while (b > a && std::isspace((unsigned char)p[b - 1])) --b;
    return std::string(p + a, b - a);
}

thread_local std::string g_open_error;

}  // namespace

struct pmx_phantom {
    int fd = -1;
    int64_t size = 0;
    std::string err;
    int int_bytes = 4, real_bytes = 4;
    std::string ident;
    std::vector<int64_t> block_n;
    std::vector<PhColumn> cols;
    std::vector<PhVar> vars;
    // device streaming state (created at the first device read)
    int device = -1;
    hipStream_t stream = nullptr;
    hipEvent_t done[2] = {nullptr, nullptr};
    void* raw = nullptr;          // pinned: file bytes of one chunk
    void* out[2] = {nullptr, nullptr};   // pinned: converted chunk, double-buffered
    size_t raw_bytes = 0, out_bytes = 0;
};

namespace pmx {
void note_exception(pmx_phantom* ph, const char* msg) noexcept {
    if (!ph) return;
    try {
        ph->err = msg;
    } catch (...) {
    }
}
void note_exception(const pmx_phantom* ph, const char* msg) noexcept {
    note_exception(const_cast<pmx_phantom*>(ph), msg);
}
}  // namespace pmx

namespace {

constexpr int64_t kChunkRows = 1 << 22;   // rows per streamed chunk (<= 32 MB of FP64)

int fail(pmx_phantom* ph, int code, const std::string& msg) {
    ph->err = msg;
    return code;
}

bool pread_all(int fd, char* d, int64_t n, int64_t at) {
    int64_t done = 0;
    while (done < n) {
        const ssize_t r = ::pread(fd, d + done, (size_t)(n - done), (off_t)(at + done));
        if (r < 0 && errno == EINTR) continue;
        if (r <= 0) return false;
        done += r;
    }
    return true;
}

constexpr int64_t kSplitBytes = 4 << 20;   // reads of at least twice this are split over threads

// Positioned reads over the file with a cursor (the reference reads sequentially from a stream).
// A large read is split over up to 16 threads: one thread copying from the page cache reaches
// only a fraction of the host's memory bandwidth.
struct Cursor {
    int fd;
    int64_t pos;
    int64_t size;
    bool read(void* dst, int64_t n) {
        if (n < 0 || pos < 0 || pos + n > size) return false;
        char* d = static_cast<char*>(dst);
        const int hw = (int)std::max(1u, std::thread::hardware_concurrency());
        const int parts = (int)std::min<int64_t>(std::min(16, hw), n / kSplitBytes);
        if (parts < 2) {
            if (!pread_all(fd, d, n, pos)) return false;
        } else {
            std::vector<std::thread> th;
            std::vector<char> ok(parts, 0);
            const int64_t step = (n + parts - 1) / parts;
            for (int t = 0; t < parts; ++t) {
                const int64_t a = t * step, b = std::min(n, a + step);
                th.emplace_back([&, t, a, b] { ok[t] = pread_all(fd, d + a, b - a, pos + a); });
            }
            for (auto& x : th) x.join();
            for (char v : ok)
                if (!v) return false;
        }
        pos += n;
        return true;
    }
    bool skip(int64_t n) {
        if (n < 0 || pos + n > size) return false;
        pos += n;
        return true;
    }
};

// One Fortran record of known payload size: tag, payload, tag; the tags must be equal
// (read_phantom.jl:83-93).  With dst == nullptr the payload is skipped and its offset returned.
bool fortran_record(Cursor& c, void* dst, int64_t bytes, int64_t* payload_at = nullptr) {
    uint32_t t0 = 0, t1 = 0;
    if (!c.read(&t0, 4)) return false;
    if (payload_at) *payload_at = c.pos;
    if (dst) {
        if (!c.read(dst, bytes)) return false;
    } else if (!c.skip(bytes)) {
        return false;
    }
    if (!c.read(&t1, 4)) return false;
    return t0 == t1;
}

int parse(pmx_phantom* ph) {
    Cursor c{ph->fd, 0, ph->size};
    const char* kCapture = "CapturePatternError: i1 mismatch. Is this a Phantom data file?";
    // capture pattern (read_phantom.jl:20-81)
    uint32_t start_tag = 0, end_tag = 0;
    if (!c.read(&start_tag, 4)) return fail(ph, PMX_ERR_FORMAT, kCapture);
    int32_t i1 = 0;
    if (!c.read(&i1, 4)) return fail(ph, PMX_ERR_FORMAT, kCapture);
    if (i1 != 60769) {
        c.pos -= 8;
        int64_t i1l = 0;
        if (!c.read(&i1l, 8) || i1l != 60769) return fail(ph, PMX_ERR_FORMAT, kCapture);
        ph->int_bytes = 8;
    }
    auto read_int = [&](int64_t* v) {
        if (ph->int_bytes == 4) {
            int32_t t;
            if (!c.read(&t, 4)) return false;
            *v = t;
        } else if (!c.read(v, 8)) {
            return false;
        }
        return true;
    };
    // real 1 == integer 2 == 060878, first as FP32
    float r1f = 0.0f;
    int64_t i2 = 0;
    if (!c.read(&r1f, 4) || !read_int(&i2)) return fail(ph, PMX_ERR_FORMAT, kCapture);
    if (i2 != 60878 || !((float)i2 == r1f)) {
        c.pos -= 8;
        double r1d = 0.0;
        if (!c.read(&r1d, 8) || !read_int(&i2)) return fail(ph, PMX_ERR_FORMAT, kCapture);
        if (i2 != 60878 || !((double)(float)i2 == r1d)) return fail(ph, PMX_ERR_FORMAT, kCapture);
        ph->real_bytes = 8;
    }
    int64_t iversion = 0, i3 = 0;
    if (!read_int(&iversion) || !read_int(&i3))
        return fail(ph, PMX_ERR_FORMAT, "CapturePatternError: i3 mismatch. Is this a Phantom data file?");
    if (i3 != 690706) return fail(ph, PMX_ERR_FORMAT, "CapturePatternError: i3 mismatch. Is this a Phantom data file?");
    if (!c.read(&end_tag, 4) || start_tag != end_tag)
        return fail(ph, PMX_ERR_FORMAT, "CapturePatternError: Fortran tags mismatch. Is this a Phantom data file?");

    // file identifier (read_phantom.jl:95-99)
    char ident[100];
    if (!fortran_record(c, ident, 100)) return fail(ph, PMX_ERR_FORMAT, "Fortran tags mismatch.");
    ph->ident = strip(ident, 100);

    // global header: eight typed blocks (read_phantom.jl:101-148)
    const int types[8] = {ph->int_bytes == 4 ? PMX_I32 : PMX_I64, PMX_I8, PMX_I16, PMX_I32, PMX_I64,
                          ph->real_bytes == 4 ? PMX_F32 : PMX_F64, PMX_F32, PMX_F64};
    for (int t = 0; t < 8; ++t) {
        int32_t nvars = 0;
        if (!fortran_record(c, &nvars, 4)) return fail(ph, PMX_ERR_FORMAT, "Fortran tags mismatch.");
        if (nvars <= 0) continue;
        const int tb = dtype_bytes(types[t]);
        if ((int64_t)nvars * (16 + tb) > ph->size) return fail(ph, PMX_ERR_FORMAT, "Fortran tags mismatch.");
        std::vector<char> keys((size_t)nvars * 16);
        std::vector<char> data((size_t)nvars * tb);
        if (!fortran_record(c, keys.data(), (int64_t)keys.size()) ||
            !fortran_record(c, data.data(), (int64_t)data.size()))
            return fail(ph, PMX_ERR_FORMAT, "Fortran tags mismatch.");
        std::vector<std::pair<std::string, int>> seen;     // original tag -> occurrences
        for (int j = 0; j < nvars; ++j) {
            std::string key = strip(&keys[(size_t)j * 16], 16);
            auto it = std::find_if(seen.begin(), seen.end(), [&](const auto& s) { return s.first == key; });
            if (it == seen.end()) {
                seen.emplace_back(key, 1);
            } else {
                it->second += 1;
                key += "_" + std::to_string(it->second);
            }
            PhVar v{key, types[t], 0.0, 0};
            const char* p = &data[(size_t)j * tb];
            switch (types[t]) {
            case PMX_I8: { int8_t x; std::memcpy(&x, p, 1); v.ival = x; break; }
            case PMX_I16: { int16_t x; std::memcpy(&x, p, 2); v.ival = x; break; }
            case PMX_I32: { int32_t x; std::memcpy(&x, p, 4); v.ival = x; break; }
            case PMX_I64: { int64_t x; std::memcpy(&x, p, 8); v.ival = x; break; }
            case PMX_F32: { float x; std::memcpy(&x, p, 4); v.fval = x; break; }
            default: { double x; std::memcpy(&x, p, 8); v.fval = x; break; }
            }
            auto old = std::find_if(ph->vars.begin(), ph->vars.end(), [&](const PhVar& u) { return u.name == v.name; });
            if (old == ph->vars.end()) ph->vars.push_back(v);
            else *old = v;
        }
    }

    // array blocks (read_phantom.jl:164-207): the directory, then each block's columns
    int32_t nblocks = 0;
    if (!fortran_record(c, &nblocks, 4)) return fail(ph, PMX_ERR_FORMAT, "Fortran tags mismatch.");
    if (nblocks < 0 || (int64_t)nblocks * 48 > ph->size) return fail(ph, PMX_ERR_FORMAT, "Fortran tags mismatch in array blocks.");
    std::vector<int64_t> n(nblocks);
    std::vector<int32_t> nums((size_t)nblocks * 8);
    for (int b = 0; b < nblocks; ++b) {
        uint32_t t0 = 0, t1 = 0;
        if (!c.read(&t0, 4) || !c.read(&n[b], 8) || !c.read(&nums[(size_t)b * 8], 32) || !c.read(&t1, 4) || t0 != t1)
            return fail(ph, PMX_ERR_FORMAT, "Fortran tags mismatch in array blocks.");
    }
    ph->block_n = n;
    const int atypes[8] = {types[0], PMX_I8, PMX_I16, PMX_I32, PMX_I64, types[5], PMX_F32, PMX_F64};
    int64_t main_rows = -1;       // rows of the main data frame once it has a column
    for (int b = 0; b < nblocks; ++b) {
        for (int t = 0; t < 8; ++t) {
            for (int k = 0; k < nums[(size_t)b * 8 + t]; ++k) {
                char tag[16];
                if (!fortran_record(c, tag, 16)) return fail(ph, PMX_ERR_FORMAT, "Fortran tags mismatch.");
                const int64_t bytes = (int64_t)dtype_bytes(atypes[t]) * std::max<int64_t>(0, n[b]);
                if (n[b] < 0 || bytes / dtype_bytes(atypes[t]) != n[b])
                    return fail(ph, PMX_ERR_FORMAT, "Fortran tags mismatch.");
                int64_t at = 0;
                if (!fortran_record(c, nullptr, bytes, &at)) return fail(ph, PMX_ERR_FORMAT, "Fortran tags mismatch.");
                PhColumn col{strip(tag, 16), b, atypes[t], n[b], at};
                const bool sinks = b == 1;
                if (!sinks) {
                    if (main_rows >= 0 && main_rows != n[b])
                        return fail(ph, PMX_ERR_DIMENSION,
                                    "DimensionMismatch: column " + col.name + " has " + std::to_string(n[b]) +
                                        " rows, the data frame " + std::to_string(main_rows));
                    main_rows = n[b];
                }
                auto same = std::find_if(ph->cols.begin(), ph->cols.end(), [&](const PhColumn& u) {
                    return u.name == col.name && (u.block == 1) == sinks;
                });
                if (same == ph->cols.end()) ph->cols.push_back(col);
                else *same = col;
            }
        }
    }
    return PMX_OK;
}

// Numeric conversion of one chunk, keeping the rows listed in keep[k0 .. k1) (row numbers are
// relative to the chunk's first row r0), or every row when keep is null.
template <class S, class D>
int64_t convert(const void* src, int64_t rows, void* dst, const int64_t* keep, int64_t k0, int64_t k1, int64_t r0) {
    const S* s = static_cast<const S*>(src);
    D* d = static_cast<D*>(dst);
    if (!keep) {
        for (int64_t i = 0; i < rows; ++i) d[i] = static_cast<D>(s[i]);
        return rows;
    }
    for (int64_t k = k0; k < k1; ++k) d[k - k0] = static_cast<D>(s[keep[k] - r0]);
    return k1 - k0;
}

template <class S>
int64_t convert_to(int out, const void* src, int64_t rows, void* dst, const int64_t* keep, int64_t k0, int64_t k1,
                   int64_t r0) {
    switch (out) {
    case PMX_I8: return convert<S, int8_t>(src, rows, dst, keep, k0, k1, r0);
    case PMX_I16: return convert<S, int16_t>(src, rows, dst, keep, k0, k1, r0);
    case PMX_I32: return convert<S, int32_t>(src, rows, dst, keep, k0, k1, r0);
    case PMX_I64: return convert<S, int64_t>(src, rows, dst, keep, k0, k1, r0);
    case PMX_F32: return convert<S, float>(src, rows, dst, keep, k0, k1, r0);
    default: return convert<S, double>(src, rows, dst, keep, k0, k1, r0);
    }
}

int64_t convert_chunk(int in, int out, const void* src, int64_t rows, void* dst, const int64_t* keep, int64_t k0,
                      int64_t k1, int64_t r0) {
    switch (in) {
    case PMX_I8: return convert_to<int8_t>(out, src, rows, dst, keep, k0, k1, r0);
    case PMX_I16: return convert_to<int16_t>(out, src, rows, dst, keep, k0, k1, r0);
    case PMX_I32: return convert_to<int32_t>(out, src, rows, dst, keep, k0, k1, r0);
    case PMX_I64: return convert_to<int64_t>(out, src, rows, dst, keep, k0, k1, r0);
    case PMX_F32: return convert_to<float>(out, src, rows, dst, keep, k0, k1, r0);
    default: return convert_to<double>(out, src, rows, dst, keep, k0, k1, r0);
    }
}

#define PH_HIP(ph, expr)                                                                   \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess) return fail(ph, PMX_ERR_HIP, std::string(#expr ": ") + hipGetErrorString(e_)); \
    } while (0)

int ensure_device_state(pmx_phantom* ph) {
    int dev = 0;
    PH_HIP(ph, hipGetDevice(&dev));
    if (ph->stream && ph->device != dev) return fail(ph, PMX_ERR_ARGUMENT, "a reader streams to one device");
    if (!ph->stream) {
        ph->device = dev;
        PH_HIP(ph, hipStreamCreateWithFlags(&ph->stream, hipStreamNonBlocking));
        for (int k = 0; k < 2; ++k) PH_HIP(ph, hipEventCreateWithFlags(&ph->done[k], hipEventDisableTiming));
        ph->raw_bytes = (size_t)kChunkRows * 8;
        ph->out_bytes = (size_t)kChunkRows * 8;
        PH_HIP(ph, hipHostMalloc(&ph->raw, ph->raw_bytes, hipHostMallocDefault));
        for (int k = 0; k < 2; ++k) PH_HIP(ph, hipHostMalloc(&ph->out[k], ph->out_bytes, hipHostMallocDefault));
    }
    return PMX_OK;
}

template <class T>
__global__ void k_density(const T* __restrict__ h, int64_t n, T c, int dim, T* __restrict__ rho) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const T invh = T(1) / h[i];
        T invhd = invh * invh;
        if (dim == 3) invhd = invhd * invh;
        rho[i] = c * invhd;
    }
}

template <class T>
void density_host(const T* h, int64_t n, T c, int dim, T* rho) {
    for (int64_t i = 0; i < n; ++i) {
        const T invh = T(1) / h[i];
        T invhd = invh * invh;
        if (dim == 3) invhd = invhd * invh;
        rho[i] = c * invhd;
    }
}

}  // namespace

using pmx::guard;

extern "C" {

int pmx_phantom_open(pmx_phantom** out, const char* path) {
    return guard(&g_open_error, [&]() -> int {
        if (!out) return PMX_ERR_ARGUMENT;
        *out = nullptr;
        g_open_error.clear();
        if (!path) {
            g_open_error = "path is NULL";
            return PMX_ERR_ARGUMENT;
        }
        const int fd = ::open(path, O_RDONLY | O_CLOEXEC);
        if (fd < 0) {
            g_open_error = std::string("SystemError: opening file \"") + path + "\": " + std::strerror(errno);
            return PMX_ERR_IO;
        }
        struct stat st;
        if (::fstat(fd, &st) != 0) {
            g_open_error = std::string("SystemError: stat \"") + path + "\": " + std::strerror(errno);
            ::close(fd);
            return PMX_ERR_IO;
        }
        pmx_phantom* ph = new (std::nothrow) pmx_phantom();
        if (!ph) {
            ::close(fd);
            g_open_error = "out of host memory";
            return PMX_ERR_NOMEM;
        }
        ph->fd = fd;
        ph->size = (int64_t)st.st_size;
        const int rc = parse(ph);
        if (rc != PMX_OK) {
            g_open_error = ph->err;
            pmx_phantom_close(ph);
            return rc;
        }
        *out = ph;
        return PMX_OK;
    });
}

const char* pmx_phantom_open_error(void) { return g_open_error.c_str(); }

void pmx_phantom_close(pmx_phantom* ph) {
    if (!ph) return;
    if (ph->stream) {
        (void)hipStreamSynchronize(ph->stream);
        for (int k = 0; k < 2; ++k)
            if (ph->done[k]) (void)hipEventDestroy(ph->done[k]);
        (void)hipStreamDestroy(ph->stream);
    }
    if (ph->raw) (void)hipHostFree(ph->raw);
    for (int k = 0; k < 2; ++k)
        if (ph->out[k]) (void)hipHostFree(ph->out[k]);
    if (ph->fd >= 0) ::close(ph->fd);
    delete ph;
}

const char* pmx_phantom_last_error(const pmx_phantom* ph) { return ph ? ph->err.c_str() : "reader is NULL"; }

int pmx_phantom_info(const pmx_phantom* ph, int64_t* vals, int32_t n) {
    return guard(ph, [&]() -> int {
        if (!ph || !vals || n < 0) return PMX_ERR_ARGUMENT;
        const int64_t v[5] = {ph->int_bytes, ph->real_bytes, (int64_t)ph->block_n.size(), (int64_t)ph->cols.size(),
                              (int64_t)ph->vars.size()};
        for (int i = 0; i < n && i < 5; ++i) vals[i] = v[i];
        return PMX_OK;
    });
}

int pmx_phantom_file_identifier(const pmx_phantom* ph, char* buf, int32_t cap) {
    return guard(ph, [&]() -> int {
        if (!ph || !buf || cap < 1) return PMX_ERR_ARGUMENT;
        const size_t k = std::min((size_t)cap - 1, ph->ident.size());
        std::memcpy(buf, ph->ident.data(), k);
        buf[k] = 0;
        return PMX_OK;
    });
}

int pmx_phantom_block_size(const pmx_phantom* ph, int32_t block, int64_t* n) {
    return guard(ph, [&]() -> int {
        if (!ph || !n || block < 0 || block >= (int32_t)ph->block_n.size()) return PMX_ERR_ARGUMENT;
        *n = ph->block_n[block];
        return PMX_OK;
    });
}

int pmx_phantom_column_info(const pmx_phantom* ph, int32_t col, char* name, int32_t* block, int32_t* dtype) {
    return guard(ph, [&]() -> int {
        if (!ph || col < 0 || col >= (int32_t)ph->cols.size()) return PMX_ERR_ARGUMENT;
        const PhColumn& c = ph->cols[col];
        if (name) {
            const size_t k = std::min<size_t>(16, c.name.size());
            std::memcpy(name, c.name.data(), k);
            name[k] = 0;
        }
        if (block) *block = c.block;
        if (dtype) *dtype = c.dtype;
        return PMX_OK;
    });
}

int pmx_phantom_header_var(const pmx_phantom* ph, int32_t i, char* name, int32_t* dtype, double* fval, int64_t* ival) {
    return guard(ph, [&]() -> int {
        if (!ph || i < 0 || i >= (int32_t)ph->vars.size()) return PMX_ERR_ARGUMENT;
        const PhVar& v = ph->vars[i];
        if (name) {
            const size_t k = std::min<size_t>(23, v.name.size());
            std::memcpy(name, v.name.data(), k);
            name[k] = 0;
        }
        if (dtype) *dtype = v.dtype;
        if (fval) *fval = v.fval;
        if (ival) *ival = v.ival;
        return PMX_OK;
    });
}

int pmx_phantom_read_column(pmx_phantom* ph, int32_t col, int32_t out_dtype, int32_t mem, void* dst,
                            const int64_t* keep, int64_t nkeep) {
    return guard(ph, [&]() -> int {
        if (!ph) return PMX_ERR_ARGUMENT;
        ph->err.clear();
        if (col < 0 || col >= (int32_t)ph->cols.size()) return fail(ph, PMX_ERR_ARGUMENT, "column index out of range");
        const PhColumn& c = ph->cols[col];
        const bool out_ok = out_dtype == PMX_F32 || out_dtype == PMX_F64 || (out_dtype >= PMX_I8 && out_dtype <= PMX_I64);
        if (!out_ok) return fail(ph, PMX_ERR_ARGUMENT, "unknown output element type");
        if (is_real(c.dtype) && !is_real(out_dtype))
            return fail(ph, PMX_ERR_ARGUMENT, "InexactError: a real column cannot be read as an integer type");
        if (mem != PMX_MEM_HOST && mem != PMX_MEM_DEVICE) return fail(ph, PMX_ERR_ARGUMENT, "mem must be host or device");
        if (keep) {
            if (nkeep < 0) return fail(ph, PMX_ERR_ARGUMENT, "nkeep < 0");
            for (int64_t k = 0; k < nkeep; ++k)
                if (keep[k] < 0 || keep[k] >= c.n || (k > 0 && keep[k] <= keep[k - 1]))
                    return fail(ph, PMX_ERR_ARGUMENT, "keep must hold strictly increasing rows of the column");
        }
        const int64_t rows_out = keep ? nkeep : c.n;
        if (rows_out > 0 && !dst) return fail(ph, PMX_ERR_ARGUMENT, "destination is NULL");
        if (rows_out == 0) return PMX_OK;
        const int ib = dtype_bytes(c.dtype), ob = dtype_bytes(out_dtype);
        const bool device = mem == PMX_MEM_DEVICE;
        if (device) {
            const int rc = ensure_device_state(ph);
            if (rc != PMX_OK) return rc;
        }
        std::vector<char> host_raw;
        if (!device) host_raw.resize((size_t)std::min<int64_t>(c.n, kChunkRows) * ib);
        Cursor cur{ph->fd, 0, ph->size};
        int64_t written = 0, k = 0, chunk = 0;
        // rows to read: every row, or from the first to the last kept row
        const int64_t first = keep ? keep[0] : 0, last = keep ? keep[nkeep - 1] + 1 : c.n;
        for (int64_t r0 = first; r0 < last; r0 += kChunkRows, ++chunk) {
            const int64_t r1 = std::min(last, r0 + kChunkRows);
            int64_t k1 = k;
            if (keep)
                while (k1 < nkeep && keep[k1] < r1) ++k1;
            if (keep && k1 == k) continue;
            // with a sparse keep list, read only the span of this chunk's kept rows
            const int64_t s0 = keep ? keep[k] : r0, s1 = keep ? keep[k1 - 1] + 1 : r1;
            // every row in the stored type: the file bytes go straight to their destination
            const bool direct = !keep && c.dtype == out_dtype;
            cur.pos = c.offset + s0 * ib;
            if (!device) {
                void* raw = direct ? static_cast<char*>(dst) + written * ob : host_raw.data();
                if (!cur.read(raw, (s1 - s0) * ib)) return fail(ph, PMX_ERR_IO, "read of column " + c.name + " failed");
                written += direct ? s1 - s0
                                  : convert_chunk(c.dtype, out_dtype, raw, s1 - s0, static_cast<char*>(dst) + written * ob,
                                                  keep, k, k1, s0);
            } else {
                const int b = (int)(chunk & 1);
                PH_HIP(ph, hipEventSynchronize(ph->done[b]));      // the copy that last used out[b] is done
                void* raw = direct ? ph->out[b] : ph->raw;
                if (!cur.read(raw, (s1 - s0) * ib)) return fail(ph, PMX_ERR_IO, "read of column " + c.name + " failed");
                const int64_t m = direct ? s1 - s0 : convert_chunk(c.dtype, out_dtype, raw, s1 - s0, ph->out[b], keep, k, k1, s0);
                PH_HIP(ph, hipMemcpyAsync(static_cast<char*>(dst) + written * ob, ph->out[b], (size_t)(m * ob),
                                          hipMemcpyHostToDevice, ph->stream));
                PH_HIP(ph, hipEventRecord(ph->done[b], ph->stream));
                written += m;
            }
            k = k1;
        }
        if (device) PH_HIP(ph, hipStreamSynchronize(ph->stream));
        if (written != rows_out) return fail(ph, PMX_ERR_IO, "short read of column " + c.name);
        return PMX_OK;
    });
}

int pmx_phantom_density(int32_t dtype, int32_t dim, int32_t mem, int64_t n, const void* h, double m, double hfact,
                        void* rho) {
    return guard(nullptr, [&]() -> int {
        if ((dtype != PMX_F32 && dtype != PMX_F64) || (dim != 2 && dim != 3) || n < 0 ||
            (mem != PMX_MEM_HOST && mem != PMX_MEM_DEVICE))
            return PMX_ERR_ARGUMENT;
        if (n == 0) return PMX_OK;
        if (!h || !rho) return PMX_ERR_ARGUMENT;
        // c = m * hfact^dim, left to right in TF (_density: m * hfact * hfact [* hfact])
        if (dtype == PMX_F64) {
            double c = m * hfact * hfact;
            if (dim == 3) c = c * hfact;
            if (mem == PMX_MEM_HOST) {
                density_host(static_cast<const double*>(h), n, c, dim, static_cast<double*>(rho));
                return PMX_OK;
            }
            const unsigned grid = (unsigned)std::min<int64_t>((n + 255) / 256, 1 << 16);
            hipLaunchKernelGGL(k_density<double>, dim3(grid), dim3(256), 0, 0, static_cast<const double*>(h), n, c, dim,
                               static_cast<double*>(rho));
        } else {
            const float mf = (float)m, hf = (float)hfact;
            float c = mf * hf * hf;
            if (dim == 3) c = c * hf;
            if (mem == PMX_MEM_HOST) {
                density_host(static_cast<const float*>(h), n, c, dim, static_cast<float*>(rho));
                return PMX_OK;
            }
            const unsigned grid = (unsigned)std::min<int64_t>((n + 255) / 256, 1 << 16);
            hipLaunchKernelGGL(k_density<float>, dim3(grid), dim3(256), 0, 0, static_cast<const float*>(h), n, c, dim,
                               static_cast<float*>(rho));
        }
        if (hipGetLastError() != hipSuccess) return PMX_ERR_HIP;
        if (hipDeviceSynchronize() != hipSuccess) return PMX_ERR_HIP;
        return PMX_OK;
    });
}

}  // extern "C"
