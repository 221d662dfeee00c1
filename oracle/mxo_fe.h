/* ORACLE — test infrastructure only. Nothing in the product may include this file.
 *
 * FloatExp restatement: value = m * 2^e with 0.5 <= |m| < 1 (or exact zero {0,0}),
 * double mantissa, int64 exponent.  Follows src/engine/floatexp.h:17-107 operation
 * for operation (same branches, same libm calls), so a CPU build with
 * -ffp-contract=off reproduces the reference's bits.
 *
 * Pinned by tests/test_oracle_fe.py: exhaustive random checks of every operation
 * against exact rational arithmetic (Python fractions) and, where it builds, against
 * the reference header compiled unchanged from its own source (oracle/_ref).
 */
#ifndef MXO_FE_H
#define MXO_FE_H

#include <math.h>
#include <stdint.h>

typedef struct {
    double m;  /* mantissa, 0.5 <= |m| < 1, or 0 */
    int64_t e; /* base-2 exponent */
} mxo_fe;

static inline mxo_fe mxo_fe_make(double m, int64_t e) {
    mxo_fe r;
    r.m = m;
    r.e = e;
    return r;
}

/* floatexp.h:23-28 */
static inline mxo_fe mxo_fe_from(double x) {
    if (x == 0.0) return mxo_fe_make(0.0, 0);
    int ex;
    double mm = frexp(x, &ex);
    return mxo_fe_make(mm, (int64_t)ex);
}

/* floatexp.h:31-37: clamp the exponent before ldexp */
static inline double mxo_fe_to_double(mxo_fe a) {
    if (a.m == 0.0) return 0.0;
    if (a.e > 1100) return a.m > 0 ? INFINITY : -INFINITY;
    if (a.e < -1100) return 0.0;
    return ldexp(a.m, (int)a.e);
}

/* floatexp.h:39-44 */
static inline mxo_fe mxo_fe_renorm(double m, int64_t e) {
    if (m == 0.0) return mxo_fe_make(0.0, 0);
    int ex;
    double mm = frexp(m, &ex);
    return mxo_fe_make(mm, e + ex);
}

/* floatexp.h:47-51 */
static inline mxo_fe mxo_fe_mul(mxo_fe a, mxo_fe b) {
    if (a.m == 0.0 || b.m == 0.0) return mxo_fe_make(0.0, 0);
    return mxo_fe_renorm(a.m * b.m, a.e + b.e);
}

/* floatexp.h:53-56 */
static inline mxo_fe mxo_fe_mul_d(mxo_fe a, double b) {
    if (a.m == 0.0 || b == 0.0) return mxo_fe_make(0.0, 0);
    return mxo_fe_renorm(a.m * b, a.e);
}

/* floatexp.h:59-61 */
static inline mxo_fe mxo_fe_scale2(mxo_fe a, int64_t k) {
    return a.m == 0.0 ? a : mxo_fe_make(a.m, a.e + k);
}

/* floatexp.h:63-73: an exponent gap above 54 returns the larger operand */
static inline mxo_fe mxo_fe_add(mxo_fe a, mxo_fe b) {
    if (a.m == 0.0) return b;
    if (b.m == 0.0) return a;
    int64_t d = a.e - b.e;
    if (d > 54) return a;
    if (d < -54) return b;
    if (d >= 0) return mxo_fe_renorm(a.m + ldexp(b.m, -(int)d), a.e);
    return mxo_fe_renorm(b.m + ldexp(a.m, (int)d), b.e);
}

static inline mxo_fe mxo_fe_neg(mxo_fe a) { return mxo_fe_make(-a.m, a.e); }
static inline mxo_fe mxo_fe_sub(mxo_fe a, mxo_fe b) { return mxo_fe_add(a, mxo_fe_neg(b)); }

/* floatexp.h:82-85 */
static inline mxo_fe mxo_fe_div(mxo_fe a, mxo_fe b) {
    if (a.m == 0.0) return mxo_fe_make(0.0, 0);
    return mxo_fe_renorm(a.m / b.m, a.e - b.e);
}

/* floatexp.h:89-94: |a| < |b| */
static inline int mxo_fe_abs_less(mxo_fe a, mxo_fe b) {
    if (a.m == 0.0) return b.m != 0.0;
    if (b.m == 0.0) return 0;
    if (a.e != b.e) return a.e < b.e;
    return fabs(a.m) < fabs(b.m);
}

/* floatexp.h:97-107: split the exponent by parity */
static inline mxo_fe mxo_fe_sqrt(mxo_fe a) {
    if (a.m <= 0.0) return mxo_fe_make(0.0, 0);
    int64_t e = a.e;
    double m = a.m;
    if (e & 1) {
        m *= 2.0;
        --e;
    }
    return mxo_fe_renorm(sqrt(m), e / 2);
}

#endif
