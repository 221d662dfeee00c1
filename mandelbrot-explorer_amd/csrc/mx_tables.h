// Host construction of the mixed form's FP32 copies of one reference: the FP32 orbit
// shadow and the FP32-floatexp BLA coefficients.  Shared by the reference upload and
// by host builds of the kernel code.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>

namespace mx {

// zf32[k] = (float)zf[k mod 2 len] for k < 2 nent (a periodic orbit repeats its first
// entries after the last)
inline void fe32_orbit(const double* zf, size_t len, size_t nent, float* zf32) {
    for (size_t k = 0; k < 2 * nent; ++k) zf32[k] = (float)zf[k % (2 * len)];
}

// A and B of each BLA entry get one exponent each (that of their larger component)
// and FP32 mantissas of magnitude below 1
inline void fe32_coefs(const double* coef, int64_t total, float4* cf, int2* ce) {
    for (int64_t q = 0; q < total; ++q) {
        const double* c = coef + 4 * q;
        int ea = 0, eb = 0;
        const double a = std::max(std::fabs(c[0]), std::fabs(c[1])), bb = std::max(std::fabs(c[2]), std::fabs(c[3]));
        if (a != 0.0 && std::isfinite(a)) (void)std::frexp(a, &ea);
        if (bb != 0.0 && std::isfinite(bb)) (void)std::frexp(bb, &eb);
        cf[q] = make_float4((float)std::ldexp(c[0], -ea), (float)std::ldexp(c[1], -ea), (float)std::ldexp(c[2], -eb), (float)std::ldexp(c[3], -eb));
        ce[q] = make_int2(ea, eb);
    }
}

} // namespace mx
