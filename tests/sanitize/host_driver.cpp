// Exercises host code of libpartia_mi355x under AddressSanitizer and UndefinedBehaviorSanitizer
// (no device is needed): the line-integrated kernel tables (pmx_lut.cpp) and the Phantom dump
// parser (pmx_phantom.hip) on the dump files named on the command line, reading every column of
// the ones that parse into host memory.  Exit status 0 when every call returns; a sanitizer
// report aborts with a non-zero status.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../include/partia_phantom.h"

int main(int argc, char** argv) {
    std::vector<double> Q(4096), I(4096);
    for (int k = 0; k < 6; ++k)
        if (pmx_line_kernel_table(k, Q.data(), I.data()) != PMX_OK) return 1;
    if (pmx_line_kernel_table(9, Q.data(), I.data()) != PMX_ERR_ARGUMENT) return 1;
    int parsed = 0, rejected = 0;
    for (int a = 1; a < argc; ++a) {
        pmx_phantom* ph = nullptr;
        if (pmx_phantom_open(&ph, argv[a]) != PMX_OK) {
            ++rejected;
            continue;
        }
        ++parsed;
        int64_t info[5];
        pmx_phantom_info(ph, info, 5);
        char name[32], ident[128];
        pmx_phantom_file_identifier(ph, ident, sizeof ident);
        for (int64_t v = 0; v < info[4]; ++v) {
            int32_t dt; double fv; int64_t iv;
            pmx_phantom_header_var(ph, (int32_t)v, name, &dt, &fv, &iv);
        }
        for (int64_t c = 0; c < info[3]; ++c) {
            int32_t blk = 0, dt = 0;
            pmx_phantom_column_info(ph, (int32_t)c, name, &blk, &dt);
            int64_t rows = 0;
            pmx_phantom_block_size(ph, blk, &rows);
            if (rows < 0 || rows > (1 << 20)) continue;
            std::vector<double> out((size_t)rows + 1);
            (void)pmx_phantom_read_column(ph, (int32_t)c, PMX_F64, PMX_MEM_HOST, out.data(), nullptr, 0);
        }
        pmx_phantom_close(ph);
    }
    std::printf("host driver ok: %d parsed, %d rejected\n", parsed, rejected);
    return 0;
}
