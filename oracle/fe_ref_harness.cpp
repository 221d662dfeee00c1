// ORACLE harness (test infrastructure): exposes the reference's own floatexp.h,
// compiled unchanged from src/engine/floatexp.h, through a C ABI so the oracle's
// FloatExp restatement can be checked against it bit for bit.
#include "floatexp.h"

extern "C" __attribute__((visibility("default"))) void fe_ref_op(int op, int n, const double* am, const long long* ae, const double* bm, const long long* be, double* rm, long long* re) {
    for (int k = 0; k < n; ++k) {
        FloatExp a{am[k], (int64_t)ae[k]}, b{bm[k], (int64_t)be[k]}, r{0.0, 0};
        switch (op) {
        case 0: r = fe_mul(a, b); break;
        case 1: r = fe_mul_d(a, bm[k]); break;
        case 2: r = fe_add(a, b); break;
        case 3: r = fe_sub(a, b); break;
        case 4: r = fe_div(a, b); break;
        case 5: r = fe_sqrt(a); break;
        case 6: r = fe_from(am[k]); break;
        case 7: r = FloatExp{fe_to_double(a), 0}; break;
        case 8: r = FloatExp{(double)fe_abs_less(a, b), 0}; break;
        }
        rm[k] = r.m;
        re[k] = (long long)r.e;
    }
}
