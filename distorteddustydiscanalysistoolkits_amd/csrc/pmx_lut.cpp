// pmx_lut.cpp — line-integrated SPH kernel tables, regenerated from their definition.
//
// The reference ships, per kernel, Q (4096 points on [0, support]) and
// Inorm(q) = sigma_3 * I(q), I(q) = integral over the whole line of f(sqrt(q^2 + s^2)) ds
// (Partia.jl/src/julia/KernelInterpolation/table/line_integrated_kernel_tables.jl, used by
// _lin_lut in line_integrated_kernel.jl:13-52).  Here I(q) is computed in long double:
// for q > 0 the substitution s = q sinh(t) (r = q cosh t) makes every polynomial piece
// of f an entire function of t, split at the kernel breakpoints and integrated with
// composite 10-point Gauss-Legendre on sub-intervals of length <= 0.25; q = 0 integrates
// the polynomial pieces of f(s) directly.
#include <cmath>

#include "pmx_internal.h"

namespace pmx {
namespace {

long double shape(int k, long double q) {
    switch (k) {
    case kM4:
        if (q < 1) return 0.25L * powl(2 - q, 3) - powl(1 - q, 3);
        if (q < 2) return 0.25L * powl(2 - q, 3);
        return 0;
    case kM5:
        if (q < 0.5L) return powl(2.5L - q, 4) - 5 * powl(1.5L - q, 4) + 10 * powl(0.5L - q, 4);
        if (q < 1.5L) return powl(2.5L - q, 4) - 5 * powl(1.5L - q, 4);
        if (q < 2.5L) return powl(2.5L - q, 4);
        return 0;
    case kM6:
        if (q < 1) return powl(3 - q, 5) - 6 * powl(2 - q, 5) + 15 * powl(1 - q, 5);
        if (q < 2) return powl(3 - q, 5) - 6 * powl(2 - q, 5);
        if (q < 3) return powl(3 - q, 5);
        return 0;
    case kC2: return q < 2 ? powl(1 - q / 2, 4) * (2 * q + 1) : 0;
    case kC4: return q < 2 ? powl(1 - q / 2, 6) * (35.0L / 12 * q * q + 3 * q + 1) : 0;
    default: return q < 2 ? powl(1 - q / 2, 8) * (4 * q * q * q + 6.25L * q * q + 4 * q + 1) : 0;
    }
}

int breakpoints(int k, long double* b) {
    switch (k) {
    case kM4: b[0] = 1; return 1;
    case kM5: b[0] = 0.5L; b[1] = 1.5L; return 2;
    case kM6: b[0] = 1; b[1] = 2; return 2;
    default: return 0;
    }
}

struct GL {
    long double x[10], w[10];
    GL() {
        const int n = 10;
        for (int i = 0; i < n; ++i) {
            long double z = cosl(3.14159265358979323846264338327950288L * (i + 0.75L) / (n + 0.5L));
            long double pp = 1;
            for (int it = 0; it < 60; ++it) {
                long double p1 = 1, p2 = 0;
                for (int j = 1; j <= n; ++j) {
                    long double p3 = p2;
                    p2 = p1;
                    p1 = ((2.0L * j - 1) * z * p2 - (j - 1.0L) * p3) / j;
                }
                pp = n * (z * p1 - p2) / (z * z - 1);
                long double dz = p1 / pp;
                z -= dz;
                if (fabsl(dz) < 1e-21L) break;
            }
            x[i] = z;
            w[i] = 2 / ((1 - z * z) * pp * pp);
        }
    }
};

long double integrate_piece(int k, long double q, long double a, long double c, const GL& g,
                            bool substituted) {
    int nsub = (int)ceill((c - a) / 0.25L);
    if (nsub < 1) nsub = 1;
    long double acc = 0;
    for (int s = 0; s < nsub; ++s) {
        long double lo = a + (c - a) * s / nsub, hi = a + (c - a) * (s + 1) / nsub;
        long double hm = (hi - lo) / 2, mid = (hi + lo) / 2;
        for (int i = 0; i < 10; ++i) {
            long double t = mid + hm * g.x[i];
            if (substituted) {
                long double ch = coshl(t);
                acc += g.w[i] * hm * shape(k, q * ch) * q * ch;
            } else {
                acc += g.w[i] * hm * shape(k, t);
            }
        }
    }
    return acc;
}

long double line_integral(int k, long double q, const GL& g) {
    const long double qmax = support(k);
    if (q >= qmax) return 0;
    long double b[4];
    int nb = breakpoints(k, b);
    long double edges[8];
    int ne = 0;
    bool sub = q > 0;
    edges[ne++] = 0;
    for (int i = 0; i < nb; ++i) {
        if (!sub) edges[ne++] = b[i];
        else if (b[i] > q) edges[ne++] = acoshl(b[i] / q);
    }
    edges[ne++] = sub ? acoshl(qmax / q) : qmax;
    long double total = 0;
    for (int e = 0; e + 1 < ne; ++e) total += integrate_piece(k, q, edges[e], edges[e + 1], g, sub);
    return 2 * total;
}

}  // namespace

void line_table(int kernel, double* Q, double* Inorm) {
    static const GL g;
    const double qmax = support(kernel);
    const long double sigma = (long double)norm3<double>(kernel);
    for (int i = 0; i < kLutN; ++i) {
        Q[i] = (double)i * qmax / 4095.0;
        Inorm[i] = (double)(sigma * line_integral(kernel, (long double)Q[i], g));
    }
}

}  // namespace pmx
