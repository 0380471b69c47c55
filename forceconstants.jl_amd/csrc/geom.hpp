// Scalar FP64 geometry and LJ closed forms used by the assembly kernels.
//
// Every expression is written in the evaluation order of the reference's Julia
// source (cited per function) and the files that include this header are compiled
// with -ffp-contract=off, so each value is the same IEEE result the reference
// computes, not just the same mathematics.
#pragma once

#include <hip/hip_runtime.h>

namespace fc3 {
namespace geom {

// nearest_mirror! (helper_funcs.jl:7-18): at most one shift per component.
__device__ __forceinline__ void mirror(double r[3], const double L[3])
{
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        if (r[c] > L[c] / 2) r[c] -= L[c];
        else if (r[c] < -L[c] / 2) r[c] += L[c];
    }
}

// sum(x -> x^2, r): ((x*x + y*y) + z*z)
__device__ __forceinline__ double sumsq(const double r[3])
{
    return (r[0] * r[0] + r[1] * r[1]) + r[2] * r[2];
}

// x^n for n >= 2.  Julia's Float64^Int is a compensated power by squaring (the
// result is correctly rounded except in rare ties); this carries the running
// product and the running square as double-double values (Dekker products via
// fma) and rounds once.  n = 2, 3 are plain products as in Julia's literal_pow.
__device__ __forceinline__ double ipow(double x, int n)
{
    if (n == 2) return x * x;
    if (n == 3) return x * x * x;
    double yh = 1.0, yl = 0.0, xh = x, xl = 0.0;
    while (n > 1) {
        if (n & 1) {
            const double ph = yh * xh;
            double pl = fma(yh, xh, -ph);
            pl = pl + (yh * xl + yl * xh);
            yh = ph;
            yl = pl;
        }
        const double sh = xh * xh;
        double sl = fma(xh, xh, -sh);
        sl = sl + 2.0 * xh * xl;
        xh = sh;
        xl = sl;
        n >>= 1;
    }
    const double ph = yh * xh;
    double pl = fma(yh, xh, -ph);
    pl = pl + (yh * xl + yl * xh);
    return ph + pl;
}

// LJ radial derivatives, LennardJones.jl:38-52 (n-ary Julia products fold left;
// 14(x) is 14*x).  s12 = sigma^12, s6 = sigma^6 are precomputed by the caller.
struct LJRadial {
    double d1, d2, d3;
};

__device__ __forceinline__ LJRadial lj_radial(double s12, double s6, double eps, double r,
                                              bool need_d3)
{
    LJRadial o;
    o.d1 = (-1.0 * eps) * (48.0 * (s12 / ipow(r, 13)) - 24.0 * (s6 / ipow(r, 7)));
    o.d2 = (4.0 * eps) * (156.0 * (s12 / ipow(r, 14)) - 42.0 * (s6 / ipow(r, 8)));
    o.d3 = need_d3 ? (-4.0 * eps) * (156.0 * (14.0 * (s12 / ipow(r, 15))) -
                                     42.0 * (8.0 * (s6 / ipow(r, 9))))
                   : 0.0;
    return o;
}

// phi2 (second_order_analytical.jl:44-51)
__device__ __forceinline__ double phi2(const LJRadial& f, double dist, const double r[3], int a,
                                      int b)
{
    const double t = ((r[a] * r[b]) / (dist * dist)) * (f.d2 - (f.d1 / dist));
    return (a == b) ? t + (f.d1 / dist) : t;
}

// phi3 (third_order_analytical.jl:216-224) with the triple product in argument
// order (x, y, z).
__device__ __forceinline__ double phi3(const LJRadial& f, double dist, const double r[3], int x,
                                      int y, int z)
{
    const double t1 = (((r[x] * r[y]) * r[z]) / ((dist * dist) * dist)) *
                      ((f.d3 - ((3.0 * f.d2) / dist)) + ((3.0 * f.d1) / (dist * dist)));
    const double dl = ((y == z ? r[x] : 0.0) + (z == x ? r[y] : 0.0)) + (x == y ? r[z] : 0.0);
    const double t2 = (dl / (dist * dist)) * (f.d2 - (f.d1 / dist));
    return t1 + t2;
}

}  // namespace geom
}  // namespace fc3
