// pmx_device.h — device-side building blocks shared by the build and interpolation
// kernels of libpartia_mi355x: SPH kernel shapes, the line-integrated table lookup and the
// distance predicates of the LBVH traversal.
//
// FP64 functions evaluate in the operation order of the reference Julia code (the whole
// library is compiled with -ffp-contract=off, so a*b+c is never fused unless fma() is
// written), which makes the plain FP64 path reproduce the reference's FP64 arithmetic.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pmx {

constexpr int kM4 = 0, kM5 = 1, kM6 = 2, kC2 = 3, kC4 = 4, kC6 = 5;
constexpr int kGather = 0, kScatter = 1, kSymmetric = 2;
constexpr int kLutN = 4096;
// Stride of the FP32 copies of the tables: each kernel's 4096 values (the last is Inorm(K) = 0)
// are followed by zeros, so a lookup at an index clamped to kLutN - 1 reads lut[i + 1] = 0.
constexpr int kLutF = kLutN + 64;

// KernelFunctionValid(K, T) (e.g. M4_spline.jl:62, M5_spline.jl:66, M6_spline.jl:66).
__host__ __device__ constexpr double support(int k) {
    return k == kM5 ? 2.5 : (k == kM6 ? 3.0 : 2.0);
}

// KernelFunctionnorm(K, Val(3), T) evaluated in T, e.g. T(1)/T(π) (M4_spline.jl:97).
template <typename T>
__host__ __device__ inline T norm3(int k) {
    const T pi = T(3.141592653589793);
    switch (k) {
    case kM4: return T(1) / pi;
    case kM5: return T(0.05) / pi;
    case kM6: return T(1) / (T(120) * pi);
    case kC2: return T(21) / (T(16) * pi);
    case kC4: return T(495) / (T(256) * pi);
    default: return T(1365) / (T(512) * pi);
    }
}

// Julia Base `^(x::Float64, n::Integer)` (Base.Math.pow_body): compensated power by
// squaring; two_mul via fma, muladd taken as fused.  n == 2 / 3 are x*x and x*x*x
// (Base.literal_pow).
__device__ inline double jl_pow(double x, int n) {
    if (n == 2) return x * x;
    if (n == 3) return x * x * x;
    double y = 1.0, xnlo = 0.0, ynlo = 0.0;
    while (n > 1) {
        if (n & 1) {
            double err = fma(y, xnlo, x * ynlo);
            double yh = x * y;
            double yl = fma(x, y, -yh);
            y = yh;
            ynlo = yl + err;
        }
        double err = x * 2.0 * xnlo;
        double xh = x * x;
        double xl = fma(x, x, -xh);
        x = xh;
        xnlo = xl + err;
        n >>= 1;
    }
    double err = fma(y, xnlo, x * ynlo);
    return (isfinite(x) && isfinite(err)) ? fma(x, y, err) : x * y;
}

// Kernel shape f(q) in FP64 (M4_spline.jl:9-20, M5_spline.jl:9-22, M6_spline.jl:9-22,
// C2_Wendland.jl:9-18, C4_Wendland.jl:9-18, C6_Wendland.jl:9-18).
template <int K>
__device__ inline double shape_f64(double q) {
    if (q < 0.0) return __builtin_nan("");
    if (K == kM4) {
        if (q < 1.0) { double a = 2.0 - q, b = 1.0 - q; return 0.25 * (a * a * a) - b * b * b; }
        if (q < 2.0) { double a = 2.0 - q; return 0.25 * (a * a * a); }
        return 0.0;
    } else if (K == kM5) {
        if (q < 0.5) return (jl_pow(2.5 - q, 4) - 5.0 * jl_pow(1.5 - q, 4)) + 10.0 * jl_pow(0.5 - q, 4);
        if (q < 1.5) return jl_pow(2.5 - q, 4) - 5.0 * jl_pow(1.5 - q, 4);
        if (q < 2.5) return jl_pow(2.5 - q, 4);
        return 0.0;
    } else if (K == kM6) {
        if (q < 1.0) return (jl_pow(3.0 - q, 5) - 6.0 * jl_pow(2.0 - q, 5)) + 15.0 * jl_pow(1.0 - q, 5);
        if (q < 2.0) return jl_pow(3.0 - q, 5) - 6.0 * jl_pow(2.0 - q, 5);
        if (q < 3.0) return jl_pow(3.0 - q, 5);
        return 0.0;
    } else if (K == kC2) {
        if (q < 2.0) return jl_pow(1.0 - 0.5 * q, 4) * (2.0 * q + 1.0);
        return 0.0;
    } else if (K == kC4) {
        if (q < 2.0) return jl_pow(1.0 - 0.5 * q, 6) * (((35.0 / 12.0) * (q * q) + 3.0 * q) + 1.0);
        return 0.0;
    } else {
        if (q < 2.0)
            return jl_pow(1.0 - 0.5 * q, 8) * (((4.0 * (q * q * q) + 6.25 * (q * q)) + 4.0 * q) + 1.0);
        return 0.0;
    }
}

// Kernel derivative f'(q) in FP64 (M4_spline.jl:22-33, C4_Wendland.jl:20-32 and siblings).
template <int K>
__device__ inline double dshape_f64(double q) {
    if (q < 0.0) return __builtin_nan("");
    if (K == kM4) {
        if (q < 1.0) { double a = 2.0 - q, b = 1.0 - q; return -0.75 * (a * a) + 3.0 * (b * b); }
        if (q < 2.0) { double a = 2.0 - q; return -0.75 * (a * a); }
        return 0.0;
    } else if (K == kM5) {
        if (q < 0.5) return (-4.0 * jl_pow(2.5 - q, 3) + 20.0 * jl_pow(1.5 - q, 3)) - 40.0 * jl_pow(0.5 - q, 3);
        if (q < 1.5) return -4.0 * jl_pow(2.5 - q, 3) + 20.0 * jl_pow(1.5 - q, 3);
        if (q < 2.5) return -4.0 * jl_pow(2.5 - q, 3);
        return 0.0;
    } else if (K == kM6) {
        if (q < 1.0) return (-5.0 * jl_pow(3.0 - q, 4) + 30.0 * jl_pow(2.0 - q, 4)) - 75.0 * jl_pow(1.0 - q, 4);
        if (q < 2.0) return -5.0 * jl_pow(3.0 - q, 4) + 30.0 * jl_pow(2.0 - q, 4);
        if (q < 3.0) return -5.0 * jl_pow(3.0 - q, 4);
        return 0.0;
    } else if (K == kC2) {
        if (q < 2.0) return jl_pow(1.0 - 0.5 * q, 4) * 2.0 - jl_pow(1.0 - 0.5 * q, 3) * (4.0 * q + 2.0);
        return 0.0;
    } else if (K == kC4) {
        if (q < 2.0)
            return jl_pow(1.0 - 0.5 * q, 6) * ((35.0 / 6.0) * q + 3.0) -
                   jl_pow(1.0 - 0.5 * q, 5) * (((35.0 / 4.0) * (q * q) + 9.0 * q) + 3.0);
        return 0.0;
    } else {
        if (q < 2.0)
            return jl_pow(1.0 - 0.5 * q, 8) * ((12.0 * (q * q) + 12.5 * q) + 4.0) -
                   jl_pow(1.0 - 0.5 * q, 7) * (((16.0 * (q * q * q) + 25.0 * (q * q)) + 16.0 * q) + 4.0);
        return 0.0;
    }
}

// Smoothed_kernel_function(K, r, h, Val(3)), kernel.jl:152-156.
template <int K>
__device__ inline double W_f64(double r, double h) {
    double invh = 1.0 / h;
    double inv_hD = invh * invh * invh;
    double q = r * invh;
    return inv_hD * (norm3<double>(K) * shape_f64<K>(q));
}

// Smoothed_gradient_kernel_function(K, dx, dy, dz, h), kernel.jl:284-296 and 371-376.
template <int K>
__device__ inline void gradW_f64(double dx, double dy, double dz, double h, double g[3]) {
    double invh = 1.0 / h;
    double inv_hDp1 = invh * invh * invh * invh;
    double r2 = dx * dx + dy * dy + dz * dz;
    double gx = 0.0, gy = 0.0, gz = 0.0;
    if (r2 != 0.0) {
        double r = sqrt(r2);
        double invr = 1.0 / r;
        double q = r / h;
        double coeff = dshape_f64<K>(q) * norm3<double>(K) * invr;
        gx = dx * coeff;
        gy = dy * coeff;
        gz = dz * coeff;
    }
    g[0] = inv_hDp1 * gx;
    g[1] = inv_hDp1 * gy;
    g[2] = inv_hDp1 * gz;
}

// _lin_lut (line_integrated_kernel.jl:13-19) on the Inorm table: linear interpolation on
// a uniform grid with the 1-based index clamped to 1..N-1.
__device__ inline double lin_lut_f64(double q, double dq, const double* __restrict__ I) {
    double idxf = q / dq + 1.0;
    int i = (int)idxf;
    i = i < 1 ? 1 : (i > kLutN - 1 ? kLutN - 1 : i);
    double t = idxf - (double)i;
    return I[i - 1] * (1.0 - t) + I[i] * t;
}

// line_integrated_kernel_function(K, r, h), line_integrated_kernel.jl:90-94, 132-137
// (physical value invh * Inorm(q): the reference's 1/h prefactor).
__device__ inline double lineW_f64(int k, double r, double h, double dq, const double* __restrict__ I) {
    double invh = 1.0 / h;
    double q = r * invh;
    double v = (q >= support(k)) ? 0.0 : lin_lut_f64(q, dq, I);
    return invh * v;
}

// Julia's max for floats without NaN handling differences on the values used here.
__device__ inline double jl_max(double a, double b) {
    if (a != a || b != b) return a + b;
    if (a > b) return a;
    if (b > a) return b;
    return signbit(a) ? b : a;
}
__device__ inline double jl_min(double a, double b) {
    if (a != a || b != b) return a + b;
    if (a < b) return a;
    if (b < a) return b;
    return signbit(a) ? a : b;
}

// Node of the FP64 tree, indexed by internal unified id (1..n-1).
struct __align__(16) Node64 {
    double mn[3];
    double mx[3];
    double hmax;
    int32_t left;
    int32_t escape;
};

// Leaf record (Morton order): position and smoothing length.
struct __align__(32) Leaf64 {
    double x, y, z, h;
};

// _squared_distance_point_aabb (LinearBVHTraversalMacros.jl:816-835).
__device__ inline double d2_point_aabb(const double p[3], const Node64& nd) {
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        double a = nd.mn[k], b = nd.mx[k], x = p[k];
        double t = x < a ? (a - x) : (x > b ? (x - b) : 0.0);
        s += t * t;
    }
    return s;
}

// _squared_distance_point_coords (:800-814).
__device__ inline double d2_point_leaf(const double p[3], const Leaf64& L) {
    double s = 0.0, d;
    d = p[0] - L.x; s += d * d;
    d = p[1] - L.y; s += d * d;
    d = p[2] - L.z; s += d * d;
    return s;
}

// _squared_distance_line_coords (:859-876).
__device__ inline double d2_line_leaf(const double o[3], const double dv[3], const Leaf64& L) {
    double D2 = 0.0, Dm = 0.0, d;
    d = L.x - o[0]; D2 += d * d; Dm += d * dv[0];
    d = L.y - o[1]; D2 += d * d; Dm += d * dv[1];
    d = L.z - o[2]; D2 += d * d; Dm += d * dv[2];
    return D2 - Dm * Dm;
}

// _line_intersects_aabb (:878-916) and _squared_distance_line_aabb_lower_bound (:918-953).
__device__ inline double d2_line_aabb_lb(const double o[3], const double dv[3], const Node64& nd) {
    double tmin = -__builtin_inf(), tmax = __builtin_inf();
    bool hit = true;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        if (!hit) break;
        double a = nd.mn[k], b = nd.mx[k];
        if (dv[k] == 0.0) {
            if (o[k] < a || o[k] > b) hit = false;
        } else {
            double t0 = (a - o[k]) / dv[k], t1 = (b - o[k]) / dv[k];
            if (t0 > t1) { double s = t0; t0 = t1; t1 = s; }
            tmin = jl_max(tmin, t0);
            tmax = jl_min(tmax, t1);
            if (tmin > tmax) hit = false;
        }
    }
    if (hit) return 0.0;
    double c[3], r2 = 0.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) c[k] = (nd.mn[k] + nd.mx[k]) * 0.5;
#pragma unroll
    for (int k = 0; k < 3; ++k) { double hh = (nd.mx[k] - nd.mn[k]) * 0.5; r2 += hh * hh; }
    double D2 = 0.0, Dm = 0.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) { double d = c[k] - o[k]; D2 += d * d; Dm += d * dv[k]; }
    double dc2 = D2 - Dm * Dm;
    if (dc2 <= r2) return 0.0;
    double dl = sqrt(dc2) - sqrt(r2);
    return dl * dl;
}

// Radius rule of the traversal macros: gather uses radius2, scatter (K h)^2, symmetric
// max(radius2, (K h)^2) (LinearBVHTraversalMacros.jl:26-371).
template <int S>
__device__ inline double radius2_rule(double Kv, double radius2, double h) {
    if (S == kGather) return radius2;
    double r = Kv * h;
    if (S == kScatter) return r * r;
    return jl_max(radius2, r * r);
}


// neighbouring floats towards -inf / +inf (finite, non-NaN input)
__device__ inline float next_down(float f) {
    if (f == 0.0f) return -__uint_as_float(1u);
    const uint32_t u = __float_as_uint(f);
    return __uint_as_float(f > 0.0f ? u - 1u : u + 1u);
}
__device__ inline float next_up(float f) { return -next_down(-f); }
__device__ inline float f32_down(double d) {
    float f = (float)d;
    if ((double)f > d) f = next_down(f);
    return f;
}
__device__ inline float f32_up(double d) {
    float f = (float)d;
    if ((double)f < d) f = next_up(f);
    return f;
}

}  // namespace pmx
