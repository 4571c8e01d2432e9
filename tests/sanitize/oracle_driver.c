/* Exercises the oracle (oracle/partia_oracle.c) under AddressSanitizer and
 * UndefinedBehaviorSanitizer: builds, queries and interpolations on small discs, duplicate and
 * single-particle sets, every kernel and strategy.  Exit status 0 when every call returns; a
 * sanitizer report aborts with a non-zero status. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "../../oracle/partia_oracle.h"

static uint64_t state = 88172645463325252ull;
static double urand(void) {
    state ^= state << 13; state ^= state >> 7; state ^= state << 17;
    return (double)(state >> 11) * (1.0 / 9007199254740992.0);
}

static int run_set(int64_t n, int dup, int kernel) {
    double *c[6 + 4];
    for (int j = 0; j < 10; ++j) c[j] = malloc(sizeof(double) * (size_t)n);
    for (int64_t i = 0; i < n; ++i) {
        const double r = 1.0 + 9.0 * urand(), phi = 6.283185307179586 * urand();
        c[0][i] = dup && i < n / 2 ? 0.5 : r * cos(phi);
        c[1][i] = dup && i < n / 2 ? -0.25 : r * sin(phi);
        c[2][i] = 0.1 * (urand() - 0.5);
        c[3][i] = 1.0 / (double)n;
        c[4][i] = 0.3 + 0.5 * urand();
        c[5][i] = 1.0 + urand();
        for (int q = 0; q < 4; ++q) c[6 + q][i] = urand() - 0.5;
    }
    double *quant[4] = {c[6], c[7], c[8], c[9]};
    orc_lbvh *t = orc_linear_bvh(n, c[0], c[1], c[2], c[3], c[4], c[5], 4, quant, 2);
    if (!t) return 1;
    const int64_t np = 300;
    double qx[300], qy[300], qz[300], ha[300];
    for (int i = 0; i < np; ++i) { qx[i] = 20 * urand() - 10; qy[i] = 20 * urand() - 10; qz[i] = 0.2 * urand() - 0.1; }
    int64_t idx[300];
    orc_nearest(t, np, qx, qy, qz, idx, ha, 2);
    for (int s = 0; s < 3; ++s) {
        int64_t *off = malloc(sizeof(int64_t) * (np + 1));
        const int64_t tot = orc_point_neighbors(t, kernel, s, np, qx, qy, qz, NULL, NULL, NULL, 0);
        int32_t *nb = malloc(sizeof(int32_t) * (size_t)(tot + 1));
        orc_point_neighbors(t, kernel, s, np, qx, qy, qz, NULL, off, nb, tot);
        free(off); free(nb);
        const int32_t ss[2] = {0, 3}, gs[1] = {-1}, ds[3] = {1, 2, 3}, cs[3] = {1, 2, 3};
        const uint8_t sh[2] = {0, 1};
        orc_catalog cat = {2, ss, sh, 1, gs, 1, ds, 1, cs};
        double *out[2 + 3 + 1 + 3];
        for (int j = 0; j < 9; ++j) out[j] = malloc(sizeof(double) * np);
        if (orc_point_samples(t, c[3], c[5], 4, (const double *const *)quant, kernel, s, &cat, np, qx, qy, qz,
                              out, 2) != 0)
            return 1;
        for (int j = 0; j < 9; ++j) free(out[j]);
    }
    double ox[300], oy[300], oz[300], dx[300], dy[300], dz[300];
    for (int i = 0; i < np; ++i) {
        ox[i] = qx[i]; oy[i] = qy[i]; oz[i] = 0.0;
        const double a = 0.2 * (urand() - 0.5), b = 0.2 * (urand() - 0.5), nn = sqrt(1 + a * a + b * b);
        dx[i] = a / nn; dy[i] = b / nn; dz[i] = 1 / nn;
    }
    const int32_t ls[2] = {0, 3};
    const uint8_t lsh[2] = {0, 1};
    double *lout[2] = {malloc(sizeof(double) * np), malloc(sizeof(double) * np)};
    if (orc_line_samples(t, c[3], c[5], 4, (const double *const *)quant, kernel, ORC_SCATTER, NULL, 2, ls, lsh, np, ox,
                         oy, oz, dx, dy, dz, lout, 2) != 0)
        return 1;
    int64_t *lo = malloc(sizeof(int64_t) * (np + 1));
    const int64_t ltot = orc_line_neighbors(t, kernel, ORC_SCATTER, np, ox, oy, oz, dx, dy, dz, NULL, NULL, NULL, 0);
    int32_t *lnb = malloc(sizeof(int32_t) * (size_t)(ltot + 1));
    orc_line_neighbors(t, kernel, ORC_SCATTER, np, ox, oy, oz, dx, dy, dz, NULL, lo, lnb, ltot);
    int64_t vi[300], vl[300], va[300];
    orc_point_visits(t, kernel, ORC_SYMMETRIC, np, qx, qy, qz, vi, vl, va, 2);
    orc_line_visits(t, kernel, np, ox, oy, oz, dx, dy, dz, vi, vl, va, 2);
    free(lo); free(lnb); free(lout[0]); free(lout[1]);
    orc_lbvh_free(t);
    for (int j = 0; j < 10; ++j) free(c[j]);
    return 0;
}

int main(void) {
    double Q[4096], I[4096];
    for (int k = 0; k < 6; ++k) {
        orc_line_table(k, Q, I);
        for (int i = 0; i <= 40; ++i) (void)orc_kernel_df(k, 0.1 * i);
    }
    const int64_t sizes[] = {1, 2, 3, 17, 1000, 5000};
    for (int s = 0; s < 6; ++s)
        for (int k = 0; k < 6; ++k)
            if (run_set(sizes[s], s % 2, k) != 0) {
                fprintf(stderr, "oracle call failed: n=%lld kernel=%d\n", (long long)sizes[s], k);
                return 1;
            }
    printf("oracle driver ok\n");
    return 0;
}
