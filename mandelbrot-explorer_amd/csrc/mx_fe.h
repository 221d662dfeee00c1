// FloatExp for the device and the host: value = m * 2^e, 0.5 <= |m| < 1 or exact zero.
//
// Same semantics as the engine's FloatExp (src/engine/floatexp.h:17-107): every
// operation uses the same IEEE operations in the same order (frexp / ldexp /
// correctly-rounded * / + sqrt), so with FP contraction disabled a gfx950 build
// produces the same bits as the CPU engine.  One representation change: the
// exponent is int32 instead of int64.  Exponents on this path stay within a few
// times 1e4 (the depth in bits plus the derivative growth), far inside int32, so
// the results are identical; the host rejects frames whose depth could approach
// 2^30.
//
// FeF is the FP32-mantissa variant (0.5 <= |m| < 1, int32 exponent) used by the
// mixed-precision form.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define MX_HD __host__ __device__ __forceinline__

struct Fe {
    double m;
    int e;
};

MX_HD Fe fe_make(double m, int e) {
    Fe r;
    r.m = m;
    r.e = e;
    return r;
}

MX_HD Fe fe_from(double x) {
    if (x == 0.0) return fe_make(0.0, 0);
    int ex;
    double mm = frexp(x, &ex);
    return fe_make(mm, ex);
}

MX_HD double fe_to_double(Fe a) {
    if (a.m == 0.0) return 0.0;
    if (a.e > 1100) return a.m > 0 ? __builtin_inf() : -__builtin_inf();
    if (a.e < -1100) return 0.0;
    return ldexp(a.m, a.e);
}

MX_HD Fe fe_renorm(double m, int e) {
    if (m == 0.0) return fe_make(0.0, 0);
    int ex;
    double mm = frexp(m, &ex);
    return fe_make(mm, e + ex);
}

// fe_renorm of a value that cannot be 0: the product or quotient of two non-zero
// mantissas (|m| >= 0.5 or a stored constant such as 1), the square root of a positive
// one.  frexp alone; identical results, without fe_renorm's zero test.
MX_HD Fe fe_renorm_nz(double m, int e) {
    int ex;
    double mm = frexp(m, &ex);
    return fe_make(mm, e + ex);
}

MX_HD Fe fe_mul(Fe a, Fe b) {
    if (a.m == 0.0 || b.m == 0.0) return fe_make(0.0, 0);
    return fe_renorm_nz(a.m * b.m, a.e + b.e);
}

MX_HD Fe fe_mul_d(Fe a, double b) {
    if (a.m == 0.0 || b == 0.0) return fe_make(0.0, 0);
    return fe_renorm(a.m * b, a.e);
}

MX_HD Fe fe_add(Fe a, Fe b) {
    if (a.m == 0.0) return b;
    if (b.m == 0.0) return a;
    int d = a.e - b.e;
    if (d > 54) return a;
    if (d < -54) return b;
    if (d >= 0) return fe_renorm(a.m + ldexp(b.m, -d), a.e);
    return fe_renorm(b.m + ldexp(a.m, d), b.e);
}

MX_HD Fe fe_neg(Fe a) { return fe_make(-a.m, a.e); }
MX_HD Fe fe_sub(Fe a, Fe b) { return fe_add(a, fe_neg(b)); }

MX_HD Fe fe_div(Fe a, Fe b) {
    if (a.m == 0.0) return fe_make(0.0, 0);
    return fe_renorm_nz(a.m / b.m, a.e - b.e);
}

MX_HD bool fe_abs_less(Fe a, Fe b) {
    if (a.m == 0.0) return b.m != 0.0;
    if (b.m == 0.0) return false;
    if (a.e != b.e) return a.e < b.e;
    return fabs(a.m) < fabs(b.m);
}

MX_HD Fe fe_sqrt(Fe a) {
    if (a.m <= 0.0) return fe_make(0.0, 0);
    int e = a.e;
    double m = a.m;
    if (e & 1) {
        m *= 2.0;
        --e;
    }
    return fe_renorm_nz(sqrt(m), e / 2);
}

// Exact 2^d for d in [-1022, 1023] built from the exponent bits.
MX_HD double fe_pow2(int d) {
    uint64_t bits = (uint64_t)(int64_t)(d + 1023) << 52;
    return __builtin_bit_cast(double, bits);
}
