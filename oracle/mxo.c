/* ORACLE — test infrastructure only.
 *
 * CPU restatement of the reference engine's per-pixel perturbation path
 * (src/engine/mandel_perturbation.cpp, "MP.cpp" below) used as the checker for the
 * HIP kernels.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it; the product never links it.
 *
 * Contents (each function cites the reference lines it follows):
 *   - frame grid and reference point in MPFR            MP.cpp:1403-1424, 4020-4074
 *   - reference orbit (Z_{i+1} = Z_i^2 + C) in MPFR      MP.cpp:4014-4187, 4189-4271
 *   - BLA table build                                    MP.cpp:2613-2709
 *   - floatexp BLA apply (tryBLAfe)                      MP.cpp:2883-2970
 *   - cold underflow helpers                             MP.cpp:2770-2876
 *   - per-pixel rescaled perturbation (pixelRescaled)    MP.cpp:3287-3618
 *   - plain-double perturbation loop + tryBLA + SA seed  MP.cpp:2291-2584, 2714-2755
 *   - brute-force MPFR pixel (accuratePointCompute)      MP.cpp:719-894
 *
 * Where each piece is pinned (tests/test_oracle_*.py):
 *   FloatExp ops       exact rational arithmetic on random inputs; the reference header
 *                      itself compiled into oracle/_ref where g++ builds it
 *   orbit              shadows equal the correctly-truncated MPFR value at 4x precision
 *   BLA tables         one merged entry == composition of its single steps (A,B) in
 *                      exact rationals; radius formula on hand-computed cases
 *   pixel loops        invariant z = X_m + S*w tracked against the MPFR brute force;
 *                      escape counts/classes vs the reference's golden frames
 *                      (tests/golden f32 frames) and vs the MPFR brute force on small frames
 *   brute force        closed forms: c=i Misiurewicz frame all-exterior, c in main
 *                      cardioid interior, c=0.25+e escape count vs sqrt law
 */
#include <gmp.h>
#include <mpfr.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

#include "mxo_fe.h"

#define MXO_EXPORT __attribute__((visibility("default")))

/* ------------------------------------------------------------------------- */
/* exact conversions                                                          */
/* ------------------------------------------------------------------------- */

/* mpf_to_fe / mpf_get_ld (MP.cpp:69-74, float_math.cpp:15-19): truncate the
 * high-precision value toward zero to a 53-bit mantissa in [0.5,1). */
static mxo_fe mpfr_to_fe_trunc(mpfr_srcptr x) {
    if (mpfr_zero_p(x)) return mxo_fe_make(0.0, 0);
    long ex;
    double d = mpfr_get_d_2exp(&ex, x, MPFR_RNDZ);
    /* mpfr_get_d_2exp returns |d| in [0.5,1) */
    return mxo_fe_make(d, (int64_t)ex);
}

/* The engine's mpf numbers set to `prec` bits (mpf_set_default_prec) carry
 * ceil((max(53,prec)+127)/64) limbs, and mpf arithmetic keeps one more: frame-level
 * work here runs MPFR at that many bits, so grid, reference point and orbit track
 * what the engine holds rather than a coarser rounding of it. */
static int mpf_bits(int prec) {
    const int p = prec < 53 ? 53 : prec;
    return ((p + 127) / 64 + 1) * 64;
}

static double mpfr_to_ld_trunc(mpfr_srcptr x) {
    if (mpfr_zero_p(x)) return 0.0;
    long ex;
    double d = mpfr_get_d_2exp(&ex, x, MPFR_RNDZ);
    return ldexp(d, (int)ex);
}

/* ------------------------------------------------------------------------- */
/* frame grid (MP.cpp:1403-1424)                                              */
/* ------------------------------------------------------------------------- */

/* The frame-level grid (mpf mode) follows the engine's own GMP mpf arithmetic
 * operation for operation (MP.cpp:1403-1424 with mpf_set_str, mpf_div, ...), so that
 * the grid and reference points equal the engine's to the last bit; mpf truncates
 * where MPFR would round, and long decimal centres convert differently.  Every mpf
 * result (at most prec+1 limbs) is then held exactly in MPFR at mpf_bits(prec).  The
 * brute force keeps an MPFR grid at its own working precision. */
typedef struct {
    mpfr_t c0_re, c0_im, dx, dy;
    mpfr_prec_t prec;
    int as_mpf;
    mpf_t fc0_re, fc0_im, fdx, fdy;
} mxo_grid;

/* mpf precision (bits argument of mpf_init2) holding the same limbs as an engine mpf
 * that mpf_bits() describes */
static mp_bitcnt_t mpf_prec_of(mpfr_prec_t pb) { return (mp_bitcnt_t)pb - 128; }

static void grid_init(mxo_grid* g, const char* cre, const char* cim, const char* scale, int W, int H, int full_h, int row_base, mpfr_prec_t prec, int as_mpf) {
    g->prec = prec;
    g->as_mpf = as_mpf;
    mpfr_inits2(prec, g->c0_re, g->c0_im, g->dx, g->dy, (mpfr_ptr)0);
    const int rows = full_h > 0 ? full_h : H;
    if (as_mpf) {
        const mp_bitcnt_t fp = mpf_prec_of(prec);
        mpf_t c_re, c_im, sc, dw, dh;
        mpf_init2(c_re, fp);
        mpf_init2(c_im, fp);
        mpf_init2(sc, fp);
        mpf_init2(dw, fp);
        mpf_init2(dh, fp);
        mpf_init2(g->fc0_re, fp);
        mpf_init2(g->fc0_im, fp);
        mpf_init2(g->fdx, fp);
        mpf_init2(g->fdy, fp);
        mpf_set_str(c_re, cre, 10);
        mpf_set_str(c_im, cim, 10);
        mpf_set_str(sc, scale, 10);
        mpf_set_ui(dw, 2);
        mpf_div(dw, dw, sc);
        mpf_set(dh, dw);
        mpf_div_ui(dh, dh, W);
        mpf_mul_ui(dh, dh, rows);
        mpf_sub(g->fc0_re, c_re, dw);
        mpf_sub(g->fc0_im, c_im, dh);
        mpf_div_ui(g->fdx, dw, W - 1);
        mpf_mul_ui(g->fdx, g->fdx, 2);
        mpf_div_ui(g->fdy, dh, rows - 1);
        mpf_mul_ui(g->fdy, g->fdy, 2);
        if (row_base > 0) {
            mpf_t off;
            mpf_init2(off, fp);
            mpf_mul_ui(off, g->fdy, row_base);
            mpf_add(g->fc0_im, g->fc0_im, off);
            mpf_clear(off);
        }
        mpfr_set_f(g->c0_re, g->fc0_re, MPFR_RNDN);
        mpfr_set_f(g->c0_im, g->fc0_im, MPFR_RNDN);
        mpfr_set_f(g->dx, g->fdx, MPFR_RNDN);
        mpfr_set_f(g->dy, g->fdy, MPFR_RNDN);
        mpf_clear(c_re);
        mpf_clear(c_im);
        mpf_clear(sc);
        mpf_clear(dw);
        mpf_clear(dh);
        return;
    }
    mpfr_t c_re, c_im, sc, dw, dh;
    mpfr_inits2(prec, c_re, c_im, sc, dw, dh, (mpfr_ptr)0);
    mpfr_set_str(c_re, cre, 10, MPFR_RNDN);
    mpfr_set_str(c_im, cim, 10, MPFR_RNDN);
    mpfr_set_str(sc, scale, 10, MPFR_RNDN);
    mpfr_ui_div(dw, 2, sc, MPFR_RNDN);   /* dw = 2/scale */
    mpfr_div_ui(dh, dw, W, MPFR_RNDN);   /* dh = dw/W*rows */
    mpfr_mul_ui(dh, dh, rows, MPFR_RNDN);
    mpfr_sub(g->c0_re, c_re, dw, MPFR_RNDN);
    mpfr_sub(g->c0_im, c_im, dh, MPFR_RNDN);
    mpfr_div_ui(g->dx, dw, W - 1, MPFR_RNDN);
    mpfr_mul_ui(g->dx, g->dx, 2, MPFR_RNDN);
    mpfr_div_ui(g->dy, dh, rows - 1, MPFR_RNDN);
    mpfr_mul_ui(g->dy, g->dy, 2, MPFR_RNDN);
    if (row_base > 0) {
        mpfr_t off;
        mpfr_init2(off, prec);
        mpfr_mul_ui(off, g->dy, row_base, MPFR_RNDN);
        mpfr_add(g->c0_im, g->c0_im, off, MPFR_RNDN);
        mpfr_clear(off);
    }
    mpfr_clears(c_re, c_im, sc, dw, dh, (mpfr_ptr)0);
}

static void grid_clear(mxo_grid* g) {
    mpfr_clears(g->c0_re, g->c0_im, g->dx, g->dy, (mpfr_ptr)0);
    if (g->as_mpf) {
        mpf_clear(g->fc0_re);
        mpf_clear(g->fc0_im);
        mpf_clear(g->fdx);
        mpf_clear(g->fdy);
    }
}

/* c of pixel (row i, col j) with sub-sample offsets (u,v); MP.cpp:896-915 (mpf mode:
 * the reference pixel c0 + d*k of setPixelReference / createRef, MP.cpp:4020-4040) */
static void grid_pixel_c(const mxo_grid* g, int i, int j, int u, int v, int sub, mpfr_ptr cre, mpfr_ptr cim) {
    if (g->as_mpf && u == 0 && v == 0) {
        mpf_t t, a, b;
        mpf_init2(t, mpf_prec_of(g->prec));
        mpf_init2(a, mpf_prec_of(g->prec));
        mpf_init2(b, mpf_prec_of(g->prec));
        mpf_mul_ui(t, g->fdx, j);
        mpf_add(a, g->fc0_re, t);
        mpf_mul_ui(t, g->fdy, i);
        mpf_add(b, g->fc0_im, t);
        mpfr_set_f(cre, a, MPFR_RNDN);
        mpfr_set_f(cim, b, MPFR_RNDN);
        mpf_clear(t);
        mpf_clear(a);
        mpf_clear(b);
        return;
    }
    mpfr_t t;
    mpfr_init2(t, g->prec);
    mpfr_mul_ui(t, g->dx, j, MPFR_RNDN);
    mpfr_add(cre, g->c0_re, t, MPFR_RNDN);
    mpfr_mul_si(t, g->dx, v, MPFR_RNDN);
    mpfr_div_ui(t, t, sub, MPFR_RNDN);
    mpfr_add(cre, cre, t, MPFR_RNDN);
    mpfr_mul_ui(t, g->dy, i, MPFR_RNDN);
    mpfr_add(cim, g->c0_im, t, MPFR_RNDN);
    mpfr_mul_si(t, g->dy, u, MPFR_RNDN);
    mpfr_div_ui(t, t, sub, MPFR_RNDN);
    mpfr_add(cim, cim, t, MPFR_RNDN);
    mpfr_clear(t);
}

/* The reference point of sample (i, j, u, v) as createRef forms it in mpf
 * (MP.cpp:4057-4074): c0 + dx j + (+-dx |v|) / sub, c0 + dy i + (+-dy |u|) / sub. */
static void grid_ref_c(const mxo_grid* g, int i, int j, int u, int v, int sub, mpfr_ptr cre, mpfr_ptr cim) {
    mpf_t t, a;
    mpf_init2(t, mpf_prec_of(g->prec));
    mpf_init2(a, mpf_prec_of(g->prec));
    const mpf_srcptr c0[2] = {g->fc0_re, g->fc0_im}, d[2] = {g->fdx, g->fdy};
    const int k[2] = {j, i}, off[2] = {v, u};
    for (int q = 0; q < 2; ++q) {
        mpf_set(a, c0[q]);
        mpf_mul_ui(t, d[q], (unsigned long)k[q]);
        mpf_add(a, a, t);
        mpf_mul_ui(t, d[q], (unsigned long)abs(off[q]));
        if (off[q] < 0) mpf_neg(t, t);
        mpf_div_ui(t, t, (unsigned long)sub);
        mpf_add(a, a, t);
        mpfr_set_f(q ? cim : cre, a, MPFR_RNDN);
    }
    mpf_clear(t);
    mpf_clear(a);
}

/* exact geometric centre c0 + d*(n-1)/2 (MP.cpp:4024-4036) */
static void grid_view_center(const mxo_grid* g, int W, int H, mpfr_ptr cre, mpfr_ptr cim) {
    if (g->as_mpf) {
        mpf_t t, a, b;
        mpf_init2(t, mpf_prec_of(g->prec));
        mpf_init2(a, mpf_prec_of(g->prec));
        mpf_init2(b, mpf_prec_of(g->prec));
        mpf_mul_ui(t, g->fdx, W - 1);
        mpf_div_ui(t, t, 2);
        mpf_add(a, g->fc0_re, t);
        mpf_mul_ui(t, g->fdy, H - 1);
        mpf_div_ui(t, t, 2);
        mpf_add(b, g->fc0_im, t);
        mpfr_set_f(cre, a, MPFR_RNDN);
        mpfr_set_f(cim, b, MPFR_RNDN);
        mpf_clear(t);
        mpf_clear(a);
        mpf_clear(b);
        return;
    }
    mpfr_t t;
    mpfr_init2(t, g->prec);
    mpfr_mul_ui(t, g->dx, W - 1, MPFR_RNDN);
    mpfr_div_ui(t, t, 2, MPFR_RNDN);
    mpfr_add(cre, g->c0_re, t, MPFR_RNDN);
    mpfr_mul_ui(t, g->dy, H - 1, MPFR_RNDN);
    mpfr_div_ui(t, t, 2, MPFR_RNDN);
    mpfr_add(cim, g->c0_im, t, MPFR_RNDN);
    mpfr_clear(t);
}

/* ------------------------------------------------------------------------- */
/* reference orbit (MP.cpp:4014-4187 createRef, 4189-4271 calCoefficient)     */
/* ------------------------------------------------------------------------- */

typedef struct {
    int len;      /* entries filled: 0..ref_it (ref_it+1 entries) */
    int ref_it;   /* escape index i (Z[i] escaped) or mxit */
    int escaped;  /* reference escaped before mxit */
    double* zf;   /* interleaved (re,im) double shadow, zf[2k..2k+1] = X_{k+1} */
    mxo_fe* zre_fe;
    mxo_fe* zim_fe;
    double* dfr;  /* reference derivative shadow (EDE) */
    double* dfi;
    double* z_m3; /* |Z|^2/1e6 (double path glitch metric) */
    mpfr_t last_re, last_im; /* full-precision tail Z[ref_it] */
    mpfr_t c_re, c_im;
    mpfr_prec_t prec;
} mxo_orbit;

/* Fixed-point orbit arithmetic of the floatexp path (src/engine/bigfixed.h; chosen by
 * createRef at MP.cpp:4084-4111, stepped by calCoefficient at MP.cpp:4197-4227).
 * Restated on signed integers: a value is the integer V standing for
 * V * 2^(-64(L-1)).  Sums are exact.  A product is the exact product of the
 * magnitudes minus the partial products a_i*b_j (limbs i, j) with i + j < L-4 that
 * the engine's short product leaves out for 16 <= L < 72 (bigfixed.h:75-100), rounded
 * half up at 2^(64(L-1)) (bigfixed.h:102-125).  The double shadow is the top nonzero
 * limb times 2^64 plus the next limb, each converted to double, as a FloatExp scaled
 * by 2^(64(top-L)) (bigfixed.h:305-319). */
typedef struct {
    int L;
    mpz_t a, b, p, u, half;
} mxo_fx;

static void fx_init(mxo_fx* F, int L) {
    F->L = L;
    mpz_inits(F->a, F->b, F->p, F->u, F->half, (mpz_ptr)0);
    mpz_setbit(F->half, (mp_bitcnt_t)64 * (L - 1) - 1);
}

static void fx_clear(mxo_fx* F) { mpz_clears(F->a, F->b, F->p, F->u, F->half, (mpz_ptr)0); }

static void fx_mul(mxo_fx* F, mpz_ptr r, mpz_srcptr x, mpz_srcptr y) {
    const int sg = mpz_sgn(x) * mpz_sgn(y), L = F->L;
    if (sg == 0) {
        mpz_set_ui(r, 0);
        return;
    }
    mpz_abs(F->a, x);
    mpz_abs(F->b, y);
    mpz_mul(F->p, F->a, F->b);
    if (L >= 16 && L < 72) {
        const int lo = L - 4;
        for (int i = 0; i < lo; ++i) {
            const mp_limb_t ai = mpz_getlimbn(F->a, i);
            if (!ai) continue;
            mpz_tdiv_r_2exp(F->u, F->b, (mp_bitcnt_t)64 * (lo - i)); /* limbs j < lo - i of b */
            mpz_mul_ui(F->u, F->u, ai);
            mpz_mul_2exp(F->u, F->u, (mp_bitcnt_t)64 * i);
            mpz_sub(F->p, F->p, F->u);
        }
    }
    mpz_add(F->p, F->p, F->half);
    mpz_fdiv_q_2exp(r, F->p, (mp_bitcnt_t)64 * (L - 1));
    if (sg < 0) mpz_neg(r, r);
}

static mxo_fe fx_to_fe(mpz_srcptr v, int L) {
    const size_t n = mpz_size(v);
    if (n == 0) return mxo_fe_make(0.0, 0);
    const int top = (int)n - 1;
    const double hi = (double)mpz_getlimbn(v, top);
    const double lo = top >= 1 ? (double)mpz_getlimbn(v, top - 1) : 0.0;
    mxo_fe r = mxo_fe_from(hi * 18446744073709551616.0 + lo);
    r.e += (int64_t)64 * (top - L);
    if (mpz_sgn(v) < 0) r.m = -r.m;
    return r;
}

/* truncation toward zero of x * 2^(64(L-1)) (bf_from_mpf, bigfixed.h:282-303) */
static void fx_from_mpfr(mpz_ptr v, mpfr_srcptr x, int L) {
    mpfr_t t;
    mpfr_init2(t, mpfr_get_prec(x));
    mpfr_mul_2ui(t, x, (unsigned long)64 * (L - 1), MPFR_RNDN); /* exact */
    mpfr_get_z(v, t, MPFR_RNDZ);
    mpfr_clear(t);
}

/* Serial full-precision orbit. Escape test on the double shadow with hypot, as
 * std::abs(Comp) > R (MP.cpp:2047-2049, 4272).  The derivative shadow follows the
 * floatexp branch of MP.cpp:4255-4263 when fe_deriv, else the double branch
 * MP.cpp:4265-4267.  fxL > 0 steps the orbit in the L-limb fixed point above (the
 * floatexp path's form); fxL == 0 in MPFR at `prec` bits, mpf's Karatsuba square. */
static void orbit_build(mxo_orbit* o, mpfr_srcptr cre, mpfr_srcptr cim, mpfr_prec_t prec, int mxit, double escape_radius, int deriv, int fe_deriv, int fxL) {
    o->prec = prec;
    o->zf = (double*)malloc(sizeof(double) * 2 * (size_t)(mxit + 1));
    o->zre_fe = (mxo_fe*)malloc(sizeof(mxo_fe) * (size_t)(mxit + 1));
    o->zim_fe = (mxo_fe*)malloc(sizeof(mxo_fe) * (size_t)(mxit + 1));
    o->dfr = (double*)calloc((size_t)(mxit + 1), sizeof(double));
    o->dfi = (double*)calloc((size_t)(mxit + 1), sizeof(double));
    o->z_m3 = (double*)malloc(sizeof(double) * (size_t)(mxit + 1));
    mpfr_inits2(prec, o->last_re, o->last_im, o->c_re, o->c_im, (mpfr_ptr)0);
    mpfr_set(o->c_re, cre, MPFR_RNDN);
    mpfr_set(o->c_im, cim, MPFR_RNDN);
    mpfr_t zr, zi, t1, t2, nzr, nzi;
    mpfr_inits2(prec, zr, zi, t1, t2, nzr, nzi, (mpfr_ptr)0);
    mpfr_set(zr, cre, MPFR_RNDN);
    mpfr_set(zi, cim, MPFR_RNDN);
    mxo_fx F;
    mpz_t vcr, vci, vzr, vzi, vs, vd, vp, vq;
    if (fxL > 0) {
        fx_init(&F, fxL);
        mpz_inits(vcr, vci, vzr, vzi, vs, vd, vp, vq, (mpz_ptr)0);
        fx_from_mpfr(vcr, cre, fxL);
        fx_from_mpfr(vci, cim, fxL);
        mpz_set(vzr, vcr);
        mpz_set(vzi, vci);
    }
    o->zf[0] = mpfr_to_ld_trunc(zr);
    o->zf[1] = mpfr_to_ld_trunc(zi);
    o->zre_fe[0] = mpfr_to_fe_trunc(zr);
    o->zim_fe[0] = mpfr_to_fe_trunc(zi);
    o->dfr[0] = 1.0;
    o->dfi[0] = 0.0;
    o->z_m3[0] = 0.0;
    mxo_fe der = mxo_fe_make(1.0, 0), dei = mxo_fe_make(0.0, 0);
    o->escaped = 0;
    o->ref_it = mxit;
    int i;
    for (i = 1; i <= mxit; ++i) {
        double dr, di;
        if (fxL > 0) {
            /* re = (a+b)(a-b) + c_re, im = 2ab + c_im (MP.cpp:4203-4209) */
            mpz_add(vs, vzr, vzi);
            mpz_sub(vd, vzr, vzi);
            fx_mul(&F, vp, vzr, vzi);
            fx_mul(&F, vq, vs, vd);
            mpz_add(vzr, vq, vcr);
            mpz_mul_2exp(vp, vp, 1);
            mpz_add(vzi, vp, vci);
            o->zre_fe[i] = fx_to_fe(vzr, fxL);
            o->zim_fe[i] = fx_to_fe(vzi, fxL);
            dr = mxo_fe_to_double(o->zre_fe[i]);
            di = mxo_fe_to_double(o->zim_fe[i]);
        } else {
            mpfr_add(t1, zr, zi, MPFR_RNDN); /* a+b */
            mpfr_sub(t2, zr, zi, MPFR_RNDN); /* a-b */
            mpfr_mul(nzi, zr, zi, MPFR_RNDN);
            mpfr_mul(nzr, t1, t2, MPFR_RNDN);
            mpfr_mul_2ui(nzi, nzi, 1, MPFR_RNDN);
            mpfr_add(zr, nzr, cre, MPFR_RNDN);
            mpfr_add(zi, nzi, cim, MPFR_RNDN);
            dr = mpfr_to_ld_trunc(zr);
            di = mpfr_to_ld_trunc(zi);
            o->zre_fe[i] = mpfr_to_fe_trunc(zr);
            o->zim_fe[i] = mpfr_to_fe_trunc(zi);
        }
        o->zf[2 * i] = dr;
        o->zf[2 * i + 1] = di;
        o->z_m3[i] = (dr * dr + di * di) / 1000000;
        if (deriv) {
            if (fe_deriv) {
                mxo_fe a = der, b = dei, ar = o->zre_fe[i - 1], ai = o->zim_fe[i - 1];
                mxo_fe re = mxo_fe_add(mxo_fe_scale2(mxo_fe_sub(mxo_fe_mul(a, ar), mxo_fe_mul(b, ai)), 1), mxo_fe_make(1.0, 0));
                mxo_fe im = mxo_fe_scale2(mxo_fe_add(mxo_fe_mul(a, ai), mxo_fe_mul(b, ar)), 1);
                der = re;
                dei = im;
                o->dfr[i] = mxo_fe_to_double(re);
                o->dfi[i] = mxo_fe_to_double(im);
            } else {
                double a = o->dfr[i - 1], b = o->dfi[i - 1], xr = o->zf[2 * (i - 1)], xi = o->zf[2 * (i - 1) + 1];
                o->dfr[i] = 2.0 * (a * xr - b * xi) + 1.0;
                o->dfi[i] = 2.0 * (a * xi + b * xr);
            }
        }
        if (hypot(dr, di) > escape_radius) {
            o->escaped = 1;
            o->ref_it = i;
            break;
        }
    }
    o->len = o->ref_it + 1;
    if (fxL > 0) {
        /* the fixed-point tail back in floating point (bf_to_mpf, bigfixed.h:268-279):
         * exact, the value has at most 64(L-1)+64 significant bits */
        mpfr_set_z(zr, vzr, MPFR_RNDN);
        mpfr_div_2ui(zr, zr, (unsigned long)64 * (fxL - 1), MPFR_RNDN);
        mpfr_set_z(zi, vzi, MPFR_RNDN);
        mpfr_div_2ui(zi, zi, (unsigned long)64 * (fxL - 1), MPFR_RNDN);
        mpz_clears(vcr, vci, vzr, vzi, vs, vd, vp, vq, (mpz_ptr)0);
        fx_clear(&F);
    }
    mpfr_set(o->last_re, zr, MPFR_RNDN);
    mpfr_set(o->last_im, zi, MPFR_RNDN);
    mpfr_clears(zr, zi, t1, t2, nzr, nzi, (mpfr_ptr)0);
}

static void orbit_free(mxo_orbit* o) {
    free(o->zf);
    free(o->zre_fe);
    free(o->zim_fe);
    free(o->dfr);
    free(o->dfi);
    free(o->z_m3);
    mpfr_clears(o->last_re, o->last_im, o->c_re, o->c_im, (mpfr_ptr)0);
}

/* refTailEscapes (MP.cpp:927-957): continue the orbit `extra` steps from its tail */
static int orbit_tail_escapes(const mxo_orbit* o, int extra, double esc2) {
    mpfr_t zr, zi, t1, t2;
    mpfr_inits2(o->prec, zr, zi, t1, t2, (mpfr_ptr)0);
    mpfr_set(zr, o->last_re, MPFR_RNDN);
    mpfr_set(zi, o->last_im, MPFR_RNDN);
    int esc = 0;
    for (int k = 0; k < extra && !esc; ++k) {
        mpfr_mul(t1, zr, zi, MPFR_RNDN);
        mpfr_mul(zr, zr, zr, MPFR_RNDN);
        mpfr_mul(t2, zi, zi, MPFR_RNDN);
        mpfr_sub(zr, zr, t2, MPFR_RNDN);
        mpfr_add(zr, zr, o->c_re, MPFR_RNDN);
        mpfr_mul_2ui(zi, t1, 1, MPFR_RNDN);
        mpfr_add(zi, zi, o->c_im, MPFR_RNDN);
        mpfr_mul(t1, zr, zr, MPFR_RNDN);
        mpfr_mul(t2, zi, zi, MPFR_RNDN);
        mpfr_add(t1, t1, t2, MPFR_RNDN);
        if (mpfr_get_d(t1, MPFR_RNDZ) > esc2) esc = 1;
    }
    mpfr_clears(zr, zi, t1, t2, (mpfr_ptr)0);
    return esc;
}

/* ------------------------------------------------------------------------- */
/* brute-force pixel (accuratePointCompute, MP.cpp:719-894)                    */
/* ------------------------------------------------------------------------- */

/* Returns the value the reference writes (smooth, EDE, or -2 interior) as double;
 * *n_out = escape index (i+1) or -1 (ran to mxit) / -2 (|d|^2 < 1e-9 early out). */
static double point_mpfr(mpfr_srcptr c_re, mpfr_srcptr c_im, mpfr_prec_t wp, int mxit, int ede, int track_d, double esc2, mpfr_srcptr dx, int* n_out) {
    mpfr_t zr, zi, dr, di, t1, t2, er, ei, e1, e2;
    mpfr_inits2(wp, zr, zi, dr, di, t1, t2, er, ei, e1, e2, (mpfr_ptr)0);
    mpfr_set(zr, c_re, MPFR_RNDN);
    mpfr_set(zi, c_im, MPFR_RNDN);
    mpfr_set_ui(dr, 2, MPFR_RNDN);
    mpfr_set_ui(di, 0, MPFR_RNDN);
    mpfr_set_ui(er, 1, MPFR_RNDN);
    mpfr_set_ui(ei, 0, MPFR_RNDN);
    const double LG = log(2);
    double res = -2.0;
    int nres = -1;
    int i = 1;
    while (i < mxit) {
        if (track_d) { /* d = 2 d z */
            mpfr_mul(t1, dr, zi, MPFR_RNDN);
            mpfr_mul(dr, dr, zr, MPFR_RNDN);
            mpfr_mul(t2, di, zi, MPFR_RNDN);
            mpfr_mul(di, di, zr, MPFR_RNDN);
            mpfr_sub(dr, dr, t2, MPFR_RNDN);
            mpfr_mul_2ui(dr, dr, 1, MPFR_RNDN);
            mpfr_add(di, di, t1, MPFR_RNDN);
            mpfr_mul_2ui(di, di, 1, MPFR_RNDN);
        }
        if (ede) { /* dc = 2 dc z + 1 */
            mpfr_mul(e1, er, zi, MPFR_RNDN);
            mpfr_mul(e2, ei, zr, MPFR_RNDN);
            mpfr_add(e2, e1, e2, MPFR_RNDN);
            mpfr_mul(e1, er, zr, MPFR_RNDN);
            mpfr_mul(t2, ei, zi, MPFR_RNDN);
            mpfr_sub(e1, e1, t2, MPFR_RNDN);
            mpfr_mul_2ui(er, e1, 1, MPFR_RNDN);
            mpfr_add_ui(er, er, 1, MPFR_RNDN);
            mpfr_mul_2ui(ei, e2, 1, MPFR_RNDN);
        }
        /* z = z^2 + c */
        mpfr_mul(t1, zr, zi, MPFR_RNDN);
        mpfr_mul(zr, zr, zr, MPFR_RNDN);
        mpfr_mul(zi, zi, zi, MPFR_RNDN);
        mpfr_sub(zr, zr, zi, MPFR_RNDN);
        mpfr_add(zr, zr, c_re, MPFR_RNDN);
        mpfr_mul_2ui(zi, t1, 1, MPFR_RNDN);
        mpfr_add(zi, zi, c_im, MPFR_RNDN);
        mpfr_mul(t1, zr, zr, MPFR_RNDN);
        mpfr_mul(t2, zi, zi, MPFR_RNDN);
        mpfr_add(t1, t1, t2, MPFR_RNDN);
        double rad = mpfr_get_d(t1, MPFR_RNDZ); /* mpf_get_d truncates */
        if (rad > esc2) {
            nres = i + 1;
            if (ede) {
                mpfr_mul(e1, er, er, MPFR_RNDN);
                mpfr_mul(e2, ei, ei, MPFR_RNDN);
                mpfr_add(e1, e1, e2, MPFR_RNDN);
                mpfr_sqrt(e1, e1, MPFR_RNDN);
                mpfr_abs(t1, dx, MPFR_RNDN);
                mpfr_mul(e1, e1, t1, MPFR_RNDN);
                double num = isfinite(rad) && rad > 0.0 ? sqrt(rad) * log(rad) : 0.0;
                mpfr_set_d(e2, num, MPFR_RNDN);
                if (!mpfr_zero_p(e1)) {
                    mpfr_div(e1, e2, e1, MPFR_RNDN);
                    res = mpfr_get_d(e1, MPFR_RNDZ);
                } else
                    res = 0.0;
            } else {
                res = (double)(i + 1) - log(log(rad) / 2 / LG) / LG;
            }
            break;
        }
        if (track_d) {
            mpfr_mul(t1, dr, dr, MPFR_RNDN);
            mpfr_mul(t2, di, di, MPFR_RNDN);
            mpfr_add(t1, t1, t2, MPFR_RNDN);
            double r = mpfr_get_d(t1, MPFR_RNDZ);
            if (r < 0.000000001) {
                nres = -2;
                break;
            }
        }
        ++i;
    }
    mpfr_clears(zr, zi, dr, di, t1, t2, er, ei, e1, e2, (mpfr_ptr)0);
    *n_out = nres;
    return res;
}

/* ------------------------------------------------------------------------- */
/* BLA table (MP.cpp:2613-2709)                                               */
/* ------------------------------------------------------------------------- */

#define MXO_MAXLEV 40

typedef struct {
    int nlev;
    int64_t off[MXO_MAXLEV];
    int64_t size[MXO_MAXLEV];
    int64_t total;
    double* r2;   /* per entry */
    int64_t* l;   /* per entry, skip length */
    double* coef; /* 4 per entry: ar ai br bi */
    double* der;  /* 4 per entry: cr ci er ei (ede only) */
    double rmax2;
    double dcmax2;
} mxo_bla;

/* zfr[s] = zf[2s], s = 1..reflen-1 */
static void bla_build(mxo_bla* b, const double* zf, const double* dfr, const double* dfi, int reflen, double dcmax, double eps, int ede) {
    memset(b, 0, sizeof(*b));
    if (reflen < 3) return;
    /* sizes: n0 = reflen-1, n_{p+1} = n_p/2 until one entry */
    int64_t n = reflen - 1, tot = 0;
    int nl = 0;
    for (;;) {
        b->off[nl] = tot;
        b->size[nl] = n;
        tot += n;
        ++nl;
        if (n <= 1) break;
        n = n / 2;
    }
    b->nlev = nl;
    b->total = tot;
    b->r2 = (double*)malloc(sizeof(double) * tot);
    b->l = (int64_t*)malloc(sizeof(int64_t) * tot);
    b->coef = (double*)malloc(sizeof(double) * 4 * tot);
    b->der = ede ? (double*)malloc(sizeof(double) * 4 * tot) : NULL;
    b->dcmax2 = dcmax * dcmax;
    b->rmax2 = 0.0;
    for (int s = 1; s < reflen; ++s) {
        double zr = zf[2 * s], zi = zf[2 * s + 1];
        double Zmag = sqrt(zr * zr + zi * zi);
        double ar = 2.0 * zr, ai = 2.0 * zi;
        double Amag = 2.0 * Zmag;
        double R = eps * (Zmag - dcmax) / (Amag + 1.0);
        double r2 = R > 0 ? R * R : 0.0;
        if (r2 > b->rmax2) b->rmax2 = r2;
        int64_t k = s - 1;
        b->r2[k] = r2;
        b->l[k] = 1;
        b->coef[4 * k + 0] = ar;
        b->coef[4 * k + 1] = ai;
        b->coef[4 * k + 2] = 1.0;
        b->coef[4 * k + 3] = 0.0;
        if (ede) {
            b->der[4 * k + 0] = 2.0 * dfr[s];
            b->der[4 * k + 1] = 2.0 * dfi[s];
            b->der[4 * k + 2] = 0.0;
            b->der[4 * k + 3] = 0.0;
        }
    }
    for (int p = 1; p < nl; ++p) {
        int64_t po = b->off[p - 1], no = b->off[p];
        for (int64_t i = 0; i < b->size[p]; ++i) {
            int64_t x = po + 2 * i, y = po + 2 * i + 1, z = no + i;
            const double* xc = b->coef + 4 * x;
            const double* yc = b->coef + 4 * y;
            double* zc = b->coef + 4 * z;
            zc[0] = yc[0] * xc[0] - yc[1] * xc[1];
            zc[1] = yc[0] * xc[1] + yc[1] * xc[0];
            zc[2] = yc[0] * xc[2] - yc[1] * xc[3] + yc[2];
            zc[3] = yc[0] * xc[3] + yc[1] * xc[2] + yc[3];
            if (ede) {
                const double* xd = b->der + 4 * x;
                const double* yd = b->der + 4 * y;
                double* zd = b->der + 4 * z;
                zd[0] = (yd[0] * xc[0] - yd[1] * xc[1]) + (yc[0] * xd[0] - yc[1] * xd[1]);
                zd[1] = (yd[0] * xc[1] + yd[1] * xc[0]) + (yc[0] * xd[1] + yc[1] * xd[0]);
                zd[2] = (yd[0] * xc[2] - yd[1] * xc[3]) + (yc[0] * xd[2] - yc[1] * xd[3]) + yd[2];
                zd[3] = (yd[0] * xc[3] + yd[1] * xc[2]) + (yc[0] * xd[3] + yc[1] * xd[2]) + yd[3];
            }
            double Rx = sqrt(b->r2[x]), Ry = sqrt(b->r2[y]);
            double Axmag = sqrt(xc[0] * xc[0] + xc[1] * xc[1]);
            double Bxmag = sqrt(xc[2] * xc[2] + xc[3] * xc[3]);
            double cand = Axmag > 0 ? (Ry - Bxmag * dcmax) / Axmag : 0.0;
            double R = (cand < Rx) ? cand : Rx; /* std::min(Rx, cand): returns Rx unless cand < Rx */
            b->r2[z] = R > 0 ? R * R : 0.0;
            b->l[z] = b->l[x] + b->l[y];
        }
    }
}

static void bla_free(mxo_bla* b) {
    free(b->r2);
    free(b->l);
    free(b->coef);
    free(b->der);
}

/* ------------------------------------------------------------------------- */
/* pixel context                                                              */
/* ------------------------------------------------------------------------- */

typedef struct {
    const double* zf;
    const mxo_fe* zre_fe;
    const mxo_fe* zim_fe;
    int mx_ref_it;
    const mxo_bla* bla; /* NULL => BLA off */
    int bla_minlevel;
    double esc2;
    int use_interior;
    double interior_eps2;
    int interior_confirm;
    int int_rep;
    int ede;
    mxo_fe dxfe;
    int per; /* > 0: periodic reference, X_m read at (m-1) mod per (MP.cpp:3308-3313) */
} mxo_ctx;

typedef struct {
    long long steps, bla_applies, bla_skipped, rebases, rescales, underflow_steps, deriv_underflows;
} mxo_stats;

static inline int tzcnt32(unsigned x) { return __builtin_ctz(x); }

/* blaReRescaleUnderflow (MP.cpp:2770-2781) */
static void bla_rerescale_underflow(mxo_fe* S, double* wr, double* wi, double nwr, double nwi) {
    if (nwr == 0.0 && nwi == 0.0) {
        *S = mxo_fe_make(1.0, 0);
        *wr = *wi = 0.0;
        return;
    }
    mxo_fe wmag = mxo_fe_sqrt(mxo_fe_add(mxo_fe_mul(mxo_fe_from(nwr), mxo_fe_from(nwr)), mxo_fe_mul(mxo_fe_from(nwi), mxo_fe_from(nwi))));
    *S = mxo_fe_mul(*S, wmag);
    double inv = 1.0 / mxo_fe_to_double(wmag);
    *wr = nwr * inv;
    *wi = nwi * inv;
}

static inline double pow2_exact(int64_t d) {
    uint64_t bits = (uint64_t)(d + 1023) << 52;
    double s;
    memcpy(&s, &bits, sizeof s);
    return s;
}

/* feRescaledStepUnderflow (MP.cpp:2794-2835) */
static void fe_step_underflow(mxo_fe Xr, mxo_fe Xi, mxo_fe dcr, mxo_fe dci, mxo_fe* S, double* wr, double* wi) {
    const double sm = S->m, sm2 = sm * sm;
    const int64_t se = S->e;
    int64_t e = INT64_MIN;
    const int has_w = *wr != 0.0 || *wi != 0.0;
    if (has_w && Xr.m != 0.0 && Xr.e + se > e) e = Xr.e + se;
    if (has_w && Xi.m != 0.0 && Xi.e + se > e) e = Xi.e + se;
    if (has_w && se + se > e) e = se + se;
    if (dcr.m != 0.0 && dcr.e > e) e = dcr.e;
    if (dci.m != 0.0 && dci.e > e) e = dci.e;
    if (e == INT64_MIN) {
        *S = mxo_fe_make(1.0, 0);
        *wr = *wi = 0.0;
        return;
    }
#define ALIGN_(mm, te) (((mm) == 0.0 || ((te) - e) < -54) ? 0.0 : (mm) * pow2_exact((te) - e))
    const double xs_r = ALIGN_(Xr.m * sm, Xr.e + se);
    const double xs_i = ALIGN_(Xi.m * sm, Xi.e + se);
    const double s2 = ALIGN_(sm2, se + se);
    const double cr = ALIGN_(dcr.m, dcr.e);
    const double ci = ALIGN_(dci.m, dci.e);
#undef ALIGN_
    const double w_r = *wr, w_i = *wi;
    const double nr = 2.0 * (xs_r * w_r - xs_i * w_i) + s2 * (w_r * w_r - w_i * w_i) + cr;
    const double ni = 2.0 * (xs_r * w_i + xs_i * w_r) + 2.0 * s2 * w_r * w_i + ci;
    const double mag2 = nr * nr + ni * ni;
    const double mag = mag2 > 0.0 ? sqrt(mag2) : hypot(nr, ni);
    if (mag == 0.0) {
        *S = mxo_fe_make(1.0, 0);
        *wr = *wi = 0.0;
        return;
    }
    *S = mxo_fe_renorm(mag, e);
    *wr = nr / mag;
    *wi = ni / mag;
}

/* feDerivativeStepUnderflow (MP.cpp:2842-2876) */
static void fe_deriv_underflow(mxo_fe zr, mxo_fe zi, mxo_fe* SJ, double* jr, double* ji, double* invSJd) {
    double im = 0.5 / SJ->m;
    int64_t ie = 1 - SJ->e;
    if (im >= 1.0) {
        im *= 0.5;
        ++ie;
    }
    mxo_fe invSJ = mxo_fe_make(im, ie);
    int64_t e = invSJ.e;
    if (zr.m != 0.0 && zr.e > e) e = zr.e;
    if (zi.m != 0.0 && zi.e > e) e = zi.e;
#define ALIGNF_(x) (((x).m == 0.0 || ((x).e - e) < -54) ? 0.0 : (x).m * pow2_exact((x).e - e))
    double azr = ALIGNF_(zr), azi = ALIGNF_(zi);
    double ar = 2.0 * (azr * *jr - azi * *ji) + ALIGNF_(invSJ);
    double ai = 2.0 * (azr * *ji + azi * *jr);
#undef ALIGNF_
    double mag = sqrt(ar * ar + ai * ai);
    if (mag == 0.0) {
        *SJ = mxo_fe_make(1.0, 0);
        *jr = *ji = 0.0;
        *invSJd = 1.0;
        return;
    }
    *SJ = mxo_fe_mul(*SJ, mxo_fe_renorm(mag, e));
    *jr = ar / mag;
    *ji = ai / mag;
    *invSJd = 1.0 / mxo_fe_to_double(*SJ);
}

/* tryBLAfe (MP.cpp:2883-2970); mx_ref_it here is the caller's reflen. */
static int try_bla_fe(const mxo_ctx* c, int s, mxo_fe* S, mxo_fe S2, double* wr, double* wi, double dr, double di, int mx_ref_it, double* outAB) {
    const mxo_bla* b = c->bla;
    if (s < 1 || !b || b->nlev == 0) return 0;
    int maxp = b->nlev - 1;
    int startp = (s == 1) ? maxp : tzcnt32((unsigned)(s - 1));
    if (startp > maxp) startp = maxp;
    if (startp < c->bla_minlevel) return 0;
    mxo_fe dz2 = mxo_fe_mul(S2, mxo_fe_from(*wr * *wr + *wi * *wi));
    if (b->rmax2 <= 0.0 || !mxo_fe_abs_less(dz2, mxo_fe_from(b->rmax2))) return 0;
    for (int p = startp; p >= c->bla_minlevel; --p) {
        int64_t i = (int64_t)((s - 1) >> p);
        if (i >= b->size[p]) continue;
        int64_t k = b->off[p] + i;
        double r2 = b->r2[k];
        if (r2 <= 0 || !mxo_fe_abs_less(dz2, mxo_fe_from(r2))) continue;
        int land = s + (int)b->l[k];
        if (land >= mx_ref_it - 1) continue;
        const double* bc = b->coef + 4 * k;
        double nwr = (bc[0] * *wr - bc[1] * *wi) + (bc[2] * dr - bc[3] * di);
        double nwi = (bc[0] * *wi + bc[1] * *wr) + (bc[2] * di + bc[3] * dr);
        const double BIG = 1e150;
        if (isfinite(nwr) && isfinite(nwi) && fabs(nwr) < BIG && fabs(nwi) < BIG) {
            double fr = c->zf[2 * land] + mxo_fe_to_double(mxo_fe_mul_d(*S, nwr));
            double fi = c->zf[2 * land + 1] + mxo_fe_to_double(mxo_fe_mul_d(*S, nwi));
            if (fr * fr + fi * fi > c->esc2) return 0;
            double wm2 = nwr * nwr + nwi * nwi;
            if (wm2 == 0.0) {
                /* a periodic reference keeps the nonzero unit delta; a full one rejects
                 * the unrepresentable landing (MP.cpp:2919-2930) */
                if (c->per > 0)
                    bla_rerescale_underflow(S, wr, wi, nwr, nwi);
                else
                    return 0;
            } else {
                double wmag = sqrt(wm2);
                *S = mxo_fe_mul(*S, mxo_fe_from(wmag));
                double inv = 1.0 / wmag;
                *wr = nwr * inv;
                *wi = nwi * inv;
            }
        } else {
            mxo_fe dzr = mxo_fe_mul_d(*S, *wr), dzi = mxo_fe_mul_d(*S, *wi);
            mxo_fe dcrfe = mxo_fe_mul_d(*S, dr), dcife = mxo_fe_mul_d(*S, di);
            mxo_fe nzr = mxo_fe_add(mxo_fe_sub(mxo_fe_mul_d(dzr, bc[0]), mxo_fe_mul_d(dzi, bc[1])), mxo_fe_sub(mxo_fe_mul_d(dcrfe, bc[2]), mxo_fe_mul_d(dcife, bc[3])));
            mxo_fe nzi = mxo_fe_add(mxo_fe_add(mxo_fe_mul_d(dzr, bc[1]), mxo_fe_mul_d(dzi, bc[0])), mxo_fe_add(mxo_fe_mul_d(dcrfe, bc[3]), mxo_fe_mul_d(dcife, bc[2])));
            double fr = c->zf[2 * land] + mxo_fe_to_double(nzr), fi = c->zf[2 * land + 1] + mxo_fe_to_double(nzi);
            if (fr * fr + fi * fi > c->esc2) return 0;
            mxo_fe Snew = mxo_fe_sqrt(mxo_fe_add(mxo_fe_mul(nzr, nzr), mxo_fe_mul(nzi, nzi)));
            if (Snew.m == 0.0) {
                *S = mxo_fe_make(1.0, 0);
                *wr = *wi = 0.0;
            } else {
                *S = Snew;
                *wr = mxo_fe_to_double(mxo_fe_div(nzr, *S));
                *wi = mxo_fe_to_double(mxo_fe_div(nzi, *S));
            }
        }
        if (outAB) {
            outAB[0] = bc[0];
            outAB[1] = bc[1];
            outAB[2] = bc[2];
            outAB[3] = bc[3];
        }
        return (int)b->l[k];
    }
    return 0;
}

/* pixelRescaled (MP.cpp:3287-3618), smooth (c_method 0) and EDE outputs.
 * Returns the double the reference casts to float; *n_out = escape iteration, or
 * -1 (reached mxit) / -2 (interior detector). */
static double pixel_rescaled(const mxo_ctx* c, mxo_fe dcr, mxo_fe dci, int mxit, int* n_out, mxo_stats* st) {
    const double ESC2 = c->esc2;
    const double LG2 = log(2.0);
    const double* zfp = c->zf;
    const int reflen = c->mx_ref_it + 1;
    int rm = 0;
    mxo_fe S = mxo_fe_sqrt(mxo_fe_add(mxo_fe_mul(dcr, dcr), mxo_fe_mul(dci, dci)));
    double wr, wi, dr, di;
    if (S.m == 0.0) {
        S = mxo_fe_make(1.0, 0);
        wr = wi = dr = di = 0.0;
    } else {
        wr = mxo_fe_to_double(mxo_fe_div(dcr, S));
        wi = mxo_fe_to_double(mxo_fe_div(dci, S));
        dr = wr;
        di = wi;
    }
    double s = mxo_fe_to_double(S);
    mxo_fe S2 = mxo_fe_mul(S, S);
    int m = 1, iter = 1;
    double zsr = 1e30, zsi = 1e30;
    int save_iter = 1, period_win = 1;
    int conf_P = 0, conf_next = 0, conf_count = 0, conf_giveup = 0;
    double conf_D2 = 0, conf_zr = 0, conf_zi = 0;
    const int ede = c->ede;
    const int deriv = ede;
    mxo_fe SJ = mxo_fe_make(1.0, 0);
    double jr = 1.0, ji = 0.0, invSJd = 1.0;
    const int detect_interior = c->use_interior;
    const int int_rep = c->int_rep && detect_interior;
    double swr = 1e300, swi = 1e300;
    mxo_fe sS = mxo_fe_make(1e300, 0);
    int sm = -1, save_iter2 = 1, twin2r = 1;

    while (iter < mxit) {
        if (int_rep) {
            if (m == sm && wr == swr && wi == swi && S.m == sS.m && S.e == sS.e) {
                *n_out = -2;
                return -2.0;
            }
            if (iter - save_iter2 >= twin2r) {
                swr = wr;
                swi = wi;
                sS = S;
                sm = m;
                save_iter2 = iter;
                twin2r += twin2r;
            }
        }
        if (c->bla && m >= 2) {
            double ab[4];
            int skip = try_bla_fe(c, rm, &S, S2, &wr, &wi, dr, di, c->per ? c->per : reflen, deriv ? ab : NULL);
            if (skip > 0) {
                if (deriv) {
                    mxo_fe Jr = mxo_fe_mul_d(SJ, jr), Ji = mxo_fe_mul_d(SJ, ji);
                    mxo_fe nJr = mxo_fe_add(mxo_fe_sub(mxo_fe_mul_d(Jr, ab[0]), mxo_fe_mul_d(Ji, ab[1])), mxo_fe_from(ab[2]));
                    mxo_fe nJi = mxo_fe_add(mxo_fe_add(mxo_fe_mul_d(Jr, ab[1]), mxo_fe_mul_d(Ji, ab[0])), mxo_fe_from(ab[3]));
                    SJ = mxo_fe_sqrt(mxo_fe_add(mxo_fe_mul(nJr, nJr), mxo_fe_mul(nJi, nJi)));
                    if (SJ.m == 0.0) {
                        SJ = mxo_fe_make(1.0, 0);
                        jr = ji = 0.0;
                        invSJd = 1.0;
                    } else {
                        jr = mxo_fe_to_double(mxo_fe_div(nJr, SJ));
                        ji = mxo_fe_to_double(mxo_fe_div(nJi, SJ));
                        invSJd = 1.0 / mxo_fe_to_double(SJ);
                    }
                }
                m += skip;
                iter += skip;
                rm += skip;
                if (c->per && rm >= c->per) rm -= c->per;
                st->bla_applies++;
                st->bla_skipped += skip;
                s = mxo_fe_to_double(S);
                S2 = mxo_fe_mul(S, S);
                dr = mxo_fe_to_double(mxo_fe_div(dcr, S));
                di = mxo_fe_to_double(mxo_fe_div(dci, S));
                zsr = 1e30;
                zsi = 1e30;
                save_iter = iter;
                period_win = 1;
                conf_P = 0;
                conf_giveup = 0;
                continue;
            }
        }
        double Xr = m ? zfp[2 * rm] : 0.0, Xi = m ? zfp[2 * rm + 1] : 0.0;
        double zmr = Xr + s * wr, zmi = Xi + s * wi;
        const int deriv_underflow = deriv && zmr == 0.0 && zmi == 0.0 && invSJd == 0.0;
        mxo_fe zmrfe = mxo_fe_make(0.0, 0), zmife = mxo_fe_make(0.0, 0);
        if (deriv_underflow) {
            mxo_fe Xrfe = m ? c->zre_fe[rm] : mxo_fe_make(0.0, 0);
            mxo_fe Xife = m ? c->zim_fe[rm] : mxo_fe_make(0.0, 0);
            zmrfe = mxo_fe_add(Xrfe, mxo_fe_mul_d(S, wr));
            zmife = mxo_fe_add(Xife, mxo_fe_mul_d(S, wi));
        }
        double nwr, nwi;
        {
            const double fR = Xr + Xr + s * wr;
            const double fI = Xi + Xi + s * wi;
            nwr = fR * wr - fI * wi + dr;
            nwi = fR * wi + fI * wr + di;
        }
        const int step_underflow = fabs(nwr) < 0x1p-1056 && fabs(nwi) < 0x1p-1056;
        if (step_underflow && (nwr != 0.0 || nwi != 0.0 || wr != 0.0 || wi != 0.0 || dcr.m != 0.0 || dci.m != 0.0)) {
            mxo_fe Xrfe = m ? c->zre_fe[rm] : mxo_fe_make(0.0, 0);
            mxo_fe Xife = m ? c->zim_fe[rm] : mxo_fe_make(0.0, 0);
            fe_step_underflow(Xrfe, Xife, dcr, dci, &S, &wr, &wi);
            s = mxo_fe_to_double(S);
            S2 = mxo_fe_mul(S, S);
            dr = mxo_fe_to_double(mxo_fe_div(dcr, S));
            di = mxo_fe_to_double(mxo_fe_div(dci, S));
            st->underflow_steps++;
        } else {
            wr = nwr;
            wi = nwi;
        }
        ++m;
        ++iter;
        ++rm;
        if (c->per && rm == c->per) rm = 0;
        st->steps++;
        if (deriv) {
            if (deriv_underflow) {
                fe_deriv_underflow(zmrfe, zmife, &SJ, &jr, &ji, &invSJd);
                st->deriv_underflows++;
            } else {
                double a = 2.0 * (zmr * jr - zmi * ji) + invSJd;
                double bb = 2.0 * (zmr * ji + zmi * jr);
                jr = a;
                ji = bb;
                double jm2 = jr * jr + ji * ji;
                if (jm2 > 1e16 || (jm2 > 0.0 && jm2 < 1e-16)) {
                    mxo_fe jmag = mxo_fe_sqrt(mxo_fe_from(jm2));
                    SJ = mxo_fe_mul(SJ, jmag);
                    double inv = 1.0 / mxo_fe_to_double(jmag);
                    jr *= inv;
                    ji *= inv;
                    invSJd = 1.0 / mxo_fe_to_double(SJ);
                }
            }
        }
        Xr = m ? zfp[2 * rm] : 0.0;
        Xi = m ? zfp[2 * rm + 1] : 0.0;
        double zr = Xr + s * wr, zi = Xi + s * wi;
        double zrad = zr * zr + zi * zi;
        if (zrad > ESC2) {
            *n_out = iter;
            if (ede) {
                double num = sqrt(zrad) * log(zrad);
                mxo_fe Jmag = mxo_fe_mul_d(SJ, sqrt(jr * jr + ji * ji));
                return mxo_fe_to_double(mxo_fe_div(mxo_fe_from(num), mxo_fe_mul(c->dxfe, Jmag)));
            }
            return (double)iter - log(log(zrad) / 2.0 / LG2) / LG2;
        }
        if (detect_interior) {
            if (conf_P > 0) {
                conf_D2 *= 4.0 * zrad;
                if (conf_D2 > 1e18) conf_D2 = 1e18;
                if (iter >= conf_next) {
                    double pr = zr - conf_zr, pi = zi - conf_zi;
                    if (pr * pr + pi * pi < c->interior_eps2 * zrad && conf_D2 < 1.0) {
                        if (++conf_count >= c->interior_confirm) {
                            *n_out = -2;
                            return -2.0;
                        }
                        conf_zr = zr;
                        conf_zi = zi;
                        conf_D2 = 1;
                        conf_next = iter + conf_P;
                    } else {
                        conf_P = 0;
                        ++conf_giveup;
                    }
                }
            } else if ((iter & 15) == 0 && conf_giveup < 3) {
                double pr = zr - zsr, pi = zi - zsi;
                if (pr * pr + pi * pi < c->interior_eps2 * zrad) {
                    conf_P = iter - save_iter;
                    if (conf_P < 1) conf_P = 1;
                    conf_zr = zr;
                    conf_zi = zi;
                    conf_D2 = 1;
                    conf_next = iter + conf_P;
                    conf_count = 0;
                }
                if (iter - save_iter >= period_win) {
                    zsr = zr;
                    zsi = zi;
                    save_iter = iter;
                    period_win += period_win;
                }
            }
        }
        double wmag2 = wr * wr + wi * wi;
        if (wmag2 > 1e16 || (wmag2 < 1e-16 && wmag2 > 0.0)) {
            mxo_fe wmag = mxo_fe_sqrt(mxo_fe_from(wmag2));
            S = mxo_fe_mul(S, wmag);
            s = mxo_fe_to_double(S);
            S2 = mxo_fe_mul(S, S);
            double inv = 1.0 / mxo_fe_to_double(wmag);
            wr *= inv;
            wi *= inv;
            dr = mxo_fe_to_double(mxo_fe_div(dcr, S));
            di = mxo_fe_to_double(mxo_fe_div(dci, S));
            st->rescales++;
        }
        double dzmag2 = s * s * (wr * wr + wi * wi);
        if (zrad < 1e-8 || zrad < dzmag2 || (!c->per && m >= reflen)) {
            mxo_fe Xmr = m ? c->zre_fe[rm] : mxo_fe_make(0.0, 0);
            mxo_fe Xmi = m ? c->zim_fe[rm] : mxo_fe_make(0.0, 0);
            mxo_fe Swr = mxo_fe_mul_d(S, wr), Swi = mxo_fe_mul_d(S, wi);
            mxo_fe zrfe = mxo_fe_add(Xmr, Swr), zife = mxo_fe_add(Xmi, Swi);
            mxo_fe zradfe = mxo_fe_add(mxo_fe_mul(zrfe, zrfe), mxo_fe_mul(zife, zife));
            mxo_fe dzfe = mxo_fe_add(mxo_fe_mul(Swr, Swr), mxo_fe_mul(Swi, Swi));
            if ((!c->per && m >= reflen) || mxo_fe_abs_less(zradfe, dzfe)) {
                mxo_fe Snew = mxo_fe_sqrt(zradfe);
                if (Snew.m == 0.0) {
                    wr = wi = 0.0;
                } else {
                    S = Snew;
                    s = mxo_fe_to_double(S);
                    S2 = mxo_fe_mul(S, S);
                    wr = mxo_fe_to_double(mxo_fe_div(zrfe, S));
                    wi = mxo_fe_to_double(mxo_fe_div(zife, S));
                    dr = mxo_fe_to_double(mxo_fe_div(dcr, S));
                    di = mxo_fe_to_double(mxo_fe_div(dci, S));
                }
                m = 0;
                rm = -1;
                st->rebases++;
            }
        }
    }
    *n_out = -1;
    return -2.0;
}

/* One lane of Mandel::solveRescaledSimd8NoBla (MP.cpp:3620-3816): the engine's
 * floatexp loop whenever BLA is off and no derivative is carried, as stepParallel
 * selects it by default (MP.cpp:2148-2149, 2172-2181: MANDEL_FESIMD=1, MANDEL_FE_DUAL=1;
 * a sub-pixel nucleus turns BLA off, MP.cpp:1629).  Against pixel_rescaled: no
 * interior detection (a sample that does not escape runs to mxit), the step-underflow
 * recovery fires when |w'|^2 is exactly 0 (MP.cpp:3741) instead of when both parts of
 * w' lie below 2^-1056, and after a recovery z is formed in FloatExp from the
 * FloatExp orbit (MP.cpp:3749-3752).  The step is the fused form (MP.cpp:3693-3701). */
static double pixel_rescaled_nobla(const mxo_ctx* c, mxo_fe dcr, mxo_fe dci, int mxit, int* n_out, mxo_stats* st) {
    const double ESC2 = c->esc2;
    const double LG2 = log(2.0);
    const double* zfp = c->zf;
    const int mx_ref_it = c->mx_ref_it, reflen = mx_ref_it + 1;
    mxo_fe S = mxo_fe_sqrt(mxo_fe_add(mxo_fe_mul(dcr, dcr), mxo_fe_mul(dci, dci)));
    double wr = 0.0, wi = 0.0, dr = 0.0, di = 0.0;
    if (S.m == 0.0) {
        S = mxo_fe_make(1.0, 0);
    } else {
        wr = mxo_fe_to_double(mxo_fe_div(dcr, S));
        wi = mxo_fe_to_double(mxo_fe_div(dci, S));
        dr = wr;
        di = wi;
    }
    double s = mxo_fe_to_double(S);
    int idx = 1, iter = 1;
    while (iter < mxit) {
        const double pr = idx > 0 ? zfp[2 * (idx - 1)] : 0.0, pi = idx > 0 ? zfp[2 * (idx - 1) + 1] : 0.0;
        int cur = idx > mx_ref_it ? mx_ref_it : idx;
        const double fR = (pr + pr) + s * wr, fI = (pi + pi) + s * wi;
        const double nwr = (fR * wr - fI * wi) + dr, nwi = (fR * wi + fI * wr) + di;
        double amag;
        {
            const double ar = zfp[2 * cur] + s * nwr, ai = zfp[2 * cur + 1] + s * nwi;
            amag = ar * ar + ai * ai;
        }
        double dmag = nwr * nwr + nwi * nwi;
        if (dmag == 0.0 && (nwr != 0.0 || nwi != 0.0 || wr != 0.0 || wi != 0.0 || dcr.m != 0.0 || dci.m != 0.0)) {
            const mxo_fe Xr = idx ? c->zre_fe[idx - 1] : mxo_fe_make(0.0, 0);
            const mxo_fe Xi = idx ? c->zim_fe[idx - 1] : mxo_fe_make(0.0, 0);
            fe_step_underflow(Xr, Xi, dcr, dci, &S, &wr, &wi);
            s = mxo_fe_to_double(S);
            dr = mxo_fe_to_double(mxo_fe_div(dcr, S));
            di = mxo_fe_to_double(mxo_fe_div(dci, S));
            const mxo_fe ar = mxo_fe_add(c->zre_fe[cur], mxo_fe_mul_d(S, wr));
            const mxo_fe ai = mxo_fe_add(c->zim_fe[cur], mxo_fe_mul_d(S, wi));
            amag = mxo_fe_to_double(mxo_fe_add(mxo_fe_mul(ar, ar), mxo_fe_mul(ai, ai)));
            dmag = wr * wr + wi * wi;
            st->underflow_steps++;
        } else {
            wr = nwr;
            wi = nwi;
        }
        ++idx;
        ++iter;
        st->steps++;
        if (amag > ESC2) {
            *n_out = iter;
            return (double)iter - log(log(amag) / 2.0 / LG2) / LG2;
        }
        if (dmag > 1e16 || (dmag < 1e-16 && dmag > 0.0)) {
            const mxo_fe nm = mxo_fe_sqrt(mxo_fe_from(dmag));
            S = mxo_fe_mul(S, nm);
            s = mxo_fe_to_double(S);
            const double inv = 1.0 / mxo_fe_to_double(nm);
            wr *= inv;
            wi *= inv;
            dr = mxo_fe_to_double(mxo_fe_div(dcr, S));
            di = mxo_fe_to_double(mxo_fe_div(dci, S));
            st->rescales++;
        }
        const double adz = s * s * (wr * wr + wi * wi);
        if (amag < 1e-8 || amag < adz || idx >= reflen) {
            const mxo_fe Xr = idx ? c->zre_fe[idx - 1] : mxo_fe_make(0.0, 0);
            const mxo_fe Xi = idx ? c->zim_fe[idx - 1] : mxo_fe_make(0.0, 0);
            const mxo_fe Swr = mxo_fe_mul_d(S, wr), Swi = mxo_fe_mul_d(S, wi);
            const mxo_fe zr = mxo_fe_add(Xr, Swr), zi = mxo_fe_add(Xi, Swi);
            const mxo_fe zsq = mxo_fe_add(mxo_fe_mul(zr, zr), mxo_fe_mul(zi, zi));
            const mxo_fe dsq = mxo_fe_add(mxo_fe_mul(Swr, Swr), mxo_fe_mul(Swi, Swi));
            if (idx >= reflen || mxo_fe_abs_less(zsq, dsq)) {
                const mxo_fe Sn = mxo_fe_sqrt(zsq);
                if (Sn.m == 0.0) {
                    wr = wi = 0.0;
                } else {
                    S = Sn;
                    s = mxo_fe_to_double(S);
                    wr = mxo_fe_to_double(mxo_fe_div(zr, S));
                    wi = mxo_fe_to_double(mxo_fe_div(zi, S));
                    dr = mxo_fe_to_double(mxo_fe_div(dcr, S));
                    di = mxo_fe_to_double(mxo_fe_div(dci, S));
                }
                idx = 0;
                st->rebases++;
            }
        }
    }
    *n_out = -1;
    return -2.0;
}

/* The engine's per-sample loop for a context: the no-BLA vector loop's semantics when
 * BLA is off and no derivative is carried on a full reference, pixelRescaled
 * otherwise (the engine builds periodic references only with BLA on). */
static double pixel_fe(const mxo_ctx* c, mxo_fe dcr, mxo_fe dci, int mxit, int* n_out, mxo_stats* st) {
    if (!c->bla && !c->ede && !c->per) return pixel_rescaled_nobla(c, dcr, dci, mxit, n_out, st);
    return pixel_rescaled(c, dcr, dci, mxit, n_out, st);
}

/* ------------------------------------------------------------------------- */
/* frame (floatexp path): Compute -> buildReferenceOrbit -> buildBLA -> probe  */
/* -> main pass (MP.cpp:1520-1761, 2128-2229)                                  */
/* ------------------------------------------------------------------------- */

typedef struct {
    /* inputs */
    const char* cre;
    const char* cim;
    const char* scale;
    int W, H, full_h, row_base;
    int mxit;
    int c_method;        /* 0 smooth, 2 EDE */
    float escape_radius; /* the engine's float radius, default 1e8f */
    int prec_bits;
    int use_bla;
    int bla_minskip;     /* MANDEL_BLA_MINSKIP, default 8 */
    double bla_eps;      /* 0 => 2^-53 */
    int interior;        /* MANDEL_INTERIOR default 1 */
    int probe;           /* run the 8x8 interior-gate probe (MANDEL_COARSE default 1) */
    int threads;
    /* outputs (W*H each) */
    double* val;
    int32_t* n_esc;
    float* out;
    /* reference info outputs */
    int ref_it;
    int ref_escaped;
    int ref_virtual;
    double ref_x, ref_y;
    int interior_used;
    int cutoff_sensitive_hint;
    double seconds_pixels;
    long long steps, bla_applies, bla_skipped, rebases;
    /* optional: the reference orbit's shadows, (re, im) interleaved, room for mxit+1
     * entries each (NULL: not wanted) */
    double* orbit_zf;
    double* orbit_m;
    int64_t* orbit_e;
    /* output cells per pixel side (0 or 1: none); with c_method's super-sampling bit
     * (1) the adaptive sub-samples.  cells/cell_val/cell_nesc: (H sub) x (W sub),
     * written where a sample is computed; ss_pixels: pixels refined */
    int sub;
    float* cells;
    double* cell_val;
    int32_t* cell_nesc;
    int ss_pixels;
    /* periodic-nucleus reference (MP.cpp:1238-1324): input no_periodic skips the
     * search (MANDEL_PERIODIC=0); outputs the nucleus period used (0: full reference)
     * and whether a nucleus was found sub-pixel (then BLA is off for the frame) */
    int no_periodic;
    int period;
    int subpixel;
    /* outputs: the frame is cutoff-sensitive (MP.cpp:4144, 4177) and how many samples
     * its brute-force fallback recomputed (MP.cpp:2173-2183) */
    int cutoff_sensitive;
    int gmp_fallback;
    /* input: reference at the exact view centre at any depth (the library's
     * MX_OPT_VIEW_CENTER_REF; the engine's createRef(view_center), MP.cpp:4021-4036) */
    int view_center;
    /* outputs (main pass): w rescales, cold step-underflow recoveries
     * (feRescaledStepUnderflow, MP.cpp:2794-2835) and derivative-underflow recoveries
     * (feDerivativeStepUnderflow, MP.cpp:2841-2876) */
    long long rescales, underflow_steps, deriv_underflows;
    /* inputs: the floatexp path at any scale above 1e6 with the centre-pixel reference
     * (MANDEL_FE=1, MP.cpp:1573-1576), and the nucleus search at any scale
     * (MANDEL_PERIODIC=1, MP.cpp:1240-1245) */
    int force_fe, force_periodic;
} mxo_frame;

/* output cell of sample (i, j, u, v) (getIndex, MP.cpp:4334-4340) */
static size_t ss_cell(int W, int sub, int i, int j, int u, int v) {
    return ((size_t)i * sub + sub / 2 + u) * (size_t)W * sub + ((size_t)j * sub + sub / 2 + v);
}

/* Adaptive supersampling's pixel test (MP.cpp:1796-1883) on the base float values
 * `out` (W x H): under EDE, a neighbour of the opposite sign; otherwise the float log
 * of the largest float difference to a neighbour above log(mxit) / 8.  flags[p] = 1
 * for the pixels refined; returns their count. */
static int ss_flag_pixels(const float* out, int W, int H, int mxit, int ede, unsigned char* flags) {
    const double thresh = log((double)mxit) / 8.0;
    int cnt = 0;
    for (int i = 0; i < H; ++i)
        for (int j = 0; j < W; ++j) {
            const float cv = out[i * W + j];
            int need = 0;
            if (ede) {
                for (int dy = -1; dy <= 1 && !need; ++dy)
                    for (int dx = -1; dx <= 1 && !need; ++dx) {
                        const int y = i + dy, x = j + dx;
                        if ((dx || dy) && x >= 0 && x < W && y >= 0 && y < H && cv * out[y * W + x] < 0.0f) need = 1;
                    }
            } else {
                float diff = 0.0f;
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dx = -1; dx <= 1; ++dx) {
                        const int y = i + dy, x = j + dx;
                        if ((dx || dy) && x >= 0 && x < W && y >= 0 && y < H) {
                            const float d = fabsf(cv - out[y * W + x]);
                            if (d > diff) diff = d;
                        }
                    }
                need = logf(diff) > thresh;
            }
            flags[i * W + j] = (unsigned char)need;
            cnt += need;
        }
    return cnt;
}

MXO_EXPORT int mxo_ss_flags(const float* out, int W, int H, int mxit, int ede, unsigned char* flags) {
    return ss_flag_pixels(out, W, H, mxit, ede, flags);
}

static int bla_minlevel_from_skip(int ms) {
    int l = 0;
    while ((1 << (l + 1)) <= ms) ++l;
    return l;
}

/* reference pixel value for a pixel (non-virtual) reference (MP.cpp:4145-4160, 4179-4185) */
static double ref_pixel_value(const mxo_orbit* o, int mxit, int c_method, double esc2, mpfr_srcptr dx, int* n_out) {
    if (!o->escaped) {
        *n_out = -1;
        return -2.0;
    }
    if (c_method) return point_mpfr(o->c_re, o->c_im, o->prec, mxit, (c_method & 2) != 0, 1, esc2, dx, n_out);
    mpfr_t t1, t2;
    mpfr_inits2(o->prec, t1, t2, (mpfr_ptr)0);
    mpfr_mul(t1, o->last_re, o->last_re, MPFR_RNDN);
    mpfr_mul(t2, o->last_im, o->last_im, MPFR_RNDN);
    mpfr_add(t1, t1, t2, MPFR_RNDN);
    double rad = mpfr_get_d(t1, MPFR_RNDZ);
    mpfr_clears(t1, t2, (mpfr_ptr)0);
    int i = o->ref_it;
    *n_out = i + 1;
    return (i + 1 - log(log(rad) / 2 / log(2)) / log(2));
}

static void run_pixels(const mxo_ctx* ctx, const mxo_frame* f, const mxo_fe* dcol, const mxo_fe* drow, const int* idx, int n, double* val, int32_t* nesc, mxo_stats* tot) {
    long long steps = 0, ap = 0, sk = 0, rb = 0, rs = 0, us = 0, du = 0;
#pragma omp parallel for schedule(dynamic, 64) num_threads(f->threads) reduction(+ : steps, ap, sk, rb, rs, us, du)
    for (int q = 0; q < n; ++q) {
        int p = idx[q];
        int i = p / f->W, j = p % f->W;
        mxo_stats st = {0, 0, 0, 0, 0, 0, 0};
        int ne;
        double v = pixel_fe(ctx, dcol[j], drow[i], f->mxit, &ne, &st);
        val[p] = v;
        nesc[p] = ne;
        steps += st.steps;
        ap += st.bla_applies;
        sk += st.bla_skipped;
        rb += st.rebases;
        rs += st.rescales;
        us += st.underflow_steps;
        du += st.deriv_underflows;
    }
    tot->steps += steps;
    tot->bla_applies += ap;
    tot->bla_skipped += sk;
    tot->rebases += rb;
    tot->rescales += rs;
    tot->underflow_steps += us;
    tot->deriv_underflows += du;
}

/* ------------------------------------------------------------------------- */
/* periodic-nucleus reference (MP.cpp:76-224 g_ballPeriod / g_newton,          */
/* 1072-1219 findNucleus / createPeriodicRef / nucleusSizeMag)                  */
/* ------------------------------------------------------------------------- */

/* z = z^2 + c, Karatsuba square in mpf (g_sqadd, MP.cpp:97-105) */
static void nuc_sqadd(mpf_ptr zr, mpf_ptr zi, mpf_srcptr cr, mpf_srcptr ci, mpf_ptr t1, mpf_ptr t2, mpf_ptr t3) {
    mpf_add(t1, zr, zi);
    mpf_sub(t2, zr, zi);
    mpf_mul(t3, zr, zi);
    mpf_mul(zr, t1, t2);
    mpf_add(zr, zr, cr);
    mpf_mul_ui(zi, t3, 2);
    mpf_add(zi, zi, ci);
}

/* g_ballPeriod (MP.cpp:116-166): the disk |c - c0| <= r through z -> z^2 + c as a
 * disk of radius R_{n+1} = (2|z_n| + R_n) R_n + r + (|z_n|^2 + |c| + 1) 2^-(prec-32);
 * the first n with |z_n|^2 <= R_n^2, or 0 once |z_n| > |c| + 2 + R_n or n > maxp */
static int nuc_ball_period(mpf_srcptr cr, mpf_srcptr ci, mpf_srcptr r, int maxp, mp_bitcnt_t prec) {
    mpf_t zr, zi, R, R2m, zabs2, nb, ru, t, t1, t2, t3;
    mpf_t* all[] = {&zr, &zi, &R, &R2m, &zabs2, &nb, &ru, &t, &t1, &t2, &t3};
    for (int k = 0; k < 11; ++k) mpf_init2(*all[k], prec);
    mpf_set_ui(zr, 0);
    mpf_set_ui(zi, 0);
    mpf_set_ui(R, 0);
    const double cab = hypot(mpf_get_d(cr), mpf_get_d(ci));
    mpf_set_ui(ru, 1);
    mpf_div_2exp(ru, ru, prec > 64 ? prec - 32 : prec / 2);
    int period = 0;
    for (int n = 1; n <= maxp; ++n) {
        const double za = hypot(mpf_get_d(zr), mpf_get_d(zi));
        mpf_set_d(t, 2.0 * za);
        mpf_add(nb, t, R);
        mpf_mul(nb, nb, R);
        mpf_add(nb, nb, r);
        mpf_set_d(t, za * za + cab + 1.0);
        mpf_mul(t, t, ru);
        mpf_add(nb, nb, t);
        mpf_set(R, nb);
        nuc_sqadd(zr, zi, cr, ci, t1, t2, t3);
        mpf_mul(t1, zr, zr);
        mpf_mul(t2, zi, zi);
        mpf_add(zabs2, t1, t2);
        mpf_mul(R2m, R, R);
        if (mpf_cmp(zabs2, R2m) <= 0) {
            period = n;
            break;
        }
        if (hypot(mpf_get_d(zr), mpf_get_d(zi)) > cab + 2.0 + mpf_get_d(R)) break;
    }
    for (int k = 0; k < 11; ++k) mpf_clear(*all[k]);
    return period;
}

/* g_newton (MP.cpp:168-222): c <- c - z_p / z_p' with (z, z') from (0, 0) by
 * z' <- 2 z z' + 1, z <- z^2 + c; converged once |step|^2 <= 2^-2(prec-48) */
static int nuc_newton(mpf_ptr cr, mpf_ptr ci, int period, mp_bitcnt_t prec, int maxit) {
    mpf_t zr, zi, dr, di, den, nr, ni, tol, step2, t1, t2, t3;
    mpf_t* all[] = {&zr, &zi, &dr, &di, &den, &nr, &ni, &tol, &step2, &t1, &t2, &t3};
    for (int k = 0; k < 12; ++k) mpf_init2(*all[k], prec);
    mpf_set_ui(tol, 1);
    mpf_div_2exp(tol, tol, prec > 96 ? prec - 48 : prec / 2);
    mpf_mul(tol, tol, tol);
    int conv = 0;
    for (int it = 0; it < maxit; ++it) {
        mpf_set_ui(zr, 0);
        mpf_set_ui(zi, 0);
        mpf_set_ui(dr, 0);
        mpf_set_ui(di, 0);
        for (int i = 0; i < period; ++i) {
            mpf_mul(t1, zr, dr);
            mpf_mul(t2, zi, di);
            mpf_sub(t3, t1, t2);
            mpf_mul_ui(t3, t3, 2);
            mpf_add_ui(t3, t3, 1);
            mpf_mul(t1, zr, di);
            mpf_mul(t2, zi, dr);
            mpf_add(di, t1, t2);
            mpf_mul_ui(di, di, 2);
            mpf_set(dr, t3);
            nuc_sqadd(zr, zi, cr, ci, t1, t2, t3);
        }
        mpf_mul(t1, dr, dr);
        mpf_mul(t2, di, di);
        mpf_add(den, t1, t2);
        if (mpf_sgn(den) == 0) break;
        mpf_mul(t1, zr, dr);
        mpf_mul(t2, zi, di);
        mpf_add(t3, t1, t2);
        mpf_div(nr, t3, den);
        mpf_mul(t1, zi, dr);
        mpf_mul(t2, zr, di);
        mpf_sub(t3, t1, t2);
        mpf_div(ni, t3, den);
        mpf_sub(cr, cr, nr);
        mpf_sub(ci, ci, ni);
        mpf_mul(t1, nr, nr);
        mpf_mul(t2, ni, ni);
        mpf_add(step2, t1, t2);
        if (mpf_cmp(step2, tol) <= 0) {
            conv = 1;
            break;
        }
    }
    for (int k = 0; k < 12; ++k) mpf_clear(*all[k]);
    return conv;
}

/* findNucleus (MP.cpp:1072-1106): search disk centred on the view centre
 * c0 + (r, r H / W), radius r = 2 / scale; the period when Newton converges to a
 * point inside the disk, else 0 */
static int nuc_find(const mxo_grid* g, const char* scale, int W, int H, int maxp, mpf_ptr nre, mpf_ptr nim) {
    const mp_bitcnt_t prec = mpf_get_prec(g->fc0_re);
    mpf_t sc, cre, cim, r, dh, dist, dr, di;
    mpf_t* all[] = {&sc, &cre, &cim, &r, &dh, &dist, &dr, &di};
    for (int k = 0; k < 8; ++k) mpf_init2(*all[k], prec);
    mpf_set_str(sc, scale, 10);
    mpf_set_ui(r, 2);
    mpf_div(r, r, sc);
    mpf_add(cre, g->fc0_re, r);
    mpf_mul_ui(dh, r, H);
    mpf_div_ui(dh, dh, W);
    mpf_add(cim, g->fc0_im, dh);
    int result = 0;
    const int period = nuc_ball_period(cre, cim, r, maxp, prec);
    if (period > 0) {
        mpf_set(nre, cre);
        mpf_set(nim, cim);
        if (nuc_newton(nre, nim, period, prec, 200)) {
            mpf_sub(dr, nre, cre);
            mpf_sub(di, nim, cim);
            mpf_mul(dr, dr, dr);
            mpf_mul(di, di, di);
            mpf_add(dist, dr, di);
            mpf_sqrt(dist, dist);
            if (mpf_cmp(dist, r) <= 0) result = period;
        }
    }
    for (int k = 0; k < 8; ++k) mpf_clear(*all[k]);
    return result;
}

static double mpf_get_ld_c(mpf_srcptr x) {
    long e;
    const double d = mpf_get_d_2exp(&e, x);
    return ldexp(d, (int)e);
}

static mxo_fe mpf_to_fe_c(mpf_srcptr x) {
    if (mpf_sgn(x) == 0) return mxo_fe_make(0.0, 0);
    long e;
    const double d = mpf_get_d_2exp(&e, x);
    return mxo_fe_make(d, (int64_t)e);
}

/* createPeriodicRef (MP.cpp:1108-1172): the nucleus's orbit for one period, X_1 = c
 * .. X_p, in mpf with the Karatsuba square (calCoefficient's mpf branch, 4230-4245);
 * shadows from mpf_get_d_2exp */
static void orbit_build_periodic(mxo_orbit* o, mpf_srcptr cre, mpf_srcptr cim, int period) {
    const mp_bitcnt_t prec = mpf_get_prec(cre);
    o->prec = (mpfr_prec_t)prec;
    o->zf = (double*)malloc(sizeof(double) * 2 * (size_t)period);
    o->zre_fe = (mxo_fe*)malloc(sizeof(mxo_fe) * (size_t)period);
    o->zim_fe = (mxo_fe*)malloc(sizeof(mxo_fe) * (size_t)period);
    o->dfr = (double*)calloc((size_t)period, sizeof(double));
    o->dfi = (double*)calloc((size_t)period, sizeof(double));
    o->z_m3 = (double*)malloc(sizeof(double) * (size_t)period);
    mpfr_inits2((mpfr_prec_t)prec, o->last_re, o->last_im, o->c_re, o->c_im, (mpfr_ptr)0);
    mpfr_set_f(o->c_re, cre, MPFR_RNDN);
    mpfr_set_f(o->c_im, cim, MPFR_RNDN);
    mpf_t zr, zi, t1, t2, nzr, nzi;
    mpf_t* all[] = {&zr, &zi, &t1, &t2, &nzr, &nzi};
    for (int k = 0; k < 6; ++k) mpf_init2(*all[k], prec);
    mpf_set(zr, cre);
    mpf_set(zi, cim);
    for (int i = 0; i < period; ++i) {
        if (i > 0) {
            mpf_add(t1, zr, zi);
            mpf_sub(t2, zr, zi);
            mpf_mul(nzi, zr, zi);
            mpf_mul(nzr, t1, t2);
            mpf_mul_ui(nzi, nzi, 2);
            mpf_add(zr, nzr, cre);
            mpf_add(zi, nzi, cim);
        }
        o->zf[2 * i] = mpf_get_ld_c(zr);
        o->zf[2 * i + 1] = mpf_get_ld_c(zi);
        o->zre_fe[i] = mpf_to_fe_c(zr);
        o->zim_fe[i] = mpf_to_fe_c(zi);
        o->z_m3[i] = (o->zf[2 * i] * o->zf[2 * i] + o->zf[2 * i + 1] * o->zf[2 * i + 1]) / 1000000;
    }
    o->dfr[0] = 1.0;
    o->escaped = 0;
    o->ref_it = period;
    o->len = period;
    mpfr_set_f(o->last_re, zr, MPFR_RNDN);
    mpfr_set_f(o->last_im, zi, MPFR_RNDN);
    for (int k = 0; k < 6; ++k) mpf_clear(*all[k]);
}

/* nucleusSizeMag (MP.cpp:1196-1219): 1 / (|b| |l|^2) with l = prod_{k<p} 2 z_k and
 * b = 1 + sum 1 / l_k, z_k = X_k, in FloatExp */
static mxo_fe nuc_size(const mxo_orbit* o, int p) {
    if (p < 2) return mxo_fe_make(1.0, 1000000);
    mxo_fe lr = mxo_fe_make(1.0, 0), li = mxo_fe_make(0.0, 0), br = mxo_fe_make(1.0, 0), bi = mxo_fe_make(0.0, 0);
    for (int k = 1; k < p; ++k) {
        const mxo_fe zr = o->zre_fe[k - 1], zi = o->zim_fe[k - 1];
        const mxo_fe nr = mxo_fe_scale2(mxo_fe_sub(mxo_fe_mul(zr, lr), mxo_fe_mul(zi, li)), 1);
        const mxo_fe ni = mxo_fe_scale2(mxo_fe_add(mxo_fe_mul(zr, li), mxo_fe_mul(zi, lr)), 1);
        lr = nr;
        li = ni;
        const mxo_fe l2 = mxo_fe_add(mxo_fe_mul(lr, lr), mxo_fe_mul(li, li));
        if (l2.m == 0.0) continue;
        br = mxo_fe_add(br, mxo_fe_div(lr, l2));
        bi = mxo_fe_sub(bi, mxo_fe_div(li, l2));
    }
    const mxo_fe lmag = mxo_fe_sqrt(mxo_fe_add(mxo_fe_mul(lr, lr), mxo_fe_mul(li, li)));
    const mxo_fe bmag = mxo_fe_sqrt(mxo_fe_add(mxo_fe_mul(br, br), mxo_fe_mul(bi, bi)));
    const mxo_fe den = mxo_fe_mul(bmag, mxo_fe_mul(lmag, lmag));
    if (den.m == 0.0) return mxo_fe_make(1.0, 1000000);
    return mxo_fe_div(mxo_fe_make(1.0, 0), den);
}

MXO_EXPORT int mxo_compute_fe(mxo_frame* f) {
    const int W = f->W, H = f->H;
    const int pb = mpf_bits(f->prec_bits);
    const double esc2 = (double)(f->escape_radius * f->escape_radius);
    const double escR = (double)f->escape_radius;
    mxo_grid g;
    grid_init(&g, f->cre, f->cim, f->scale, W, H, f->full_h, f->row_base, pb, 1);
    mxo_fe dxfe = mpfr_to_fe_trunc(g.dx), dyfe = mpfr_to_fe_trunc(g.dy);
    const int ede = (f->c_method & 2) != 0;

    /* reference at the centre pixel (H/2, W/2); a frame at scale <= 1e280 reaches this
     * path only by the double path's escalation (MP.cpp:1336-1355), whose reference
     * is the exact view centre */
    mpfr_t rre, rim;
    mpfr_inits2(pb, rre, rim, (mpfr_ptr)0);
    int escalated, shallow_fe;
    {
        mpfr_t sc;
        mpfr_init2(sc, pb);
        mpfr_set_str(sc, f->scale, 10, MPFR_RNDN);
        shallow_fe = mpfr_cmp_d(sc, 1e280) <= 0;
        escalated = (shallow_fe && !f->force_fe) || f->view_center;
        mpfr_clear(sc);
    }
    /* one integer limb: the fixed-point orbit needs |z|^2 + |c| well below 2^64
     * (bigFixedRangeSafe, MP.cpp:4093-4095); L = 64-bit limbs of the mpf centre + 2 */
    const int fxL = f->escape_radius <= 0x1p31f ? pb / 64 : 0;
    mxo_orbit o;
    int ref_virtual = 0;
    double ref_x = W / 2, ref_y = H / 2;
    f->cutoff_sensitive_hint = 0;
    f->period = 0;
    f->subpixel = 0;
    int have_ref = 0;
    /* a nucleus in view on the deep path with BLA (buildReferenceOrbit, MP.cpp:
     * 1238-1324): resolved (at least 4 pixels) it is the reference, its orbit one
     * period long; sub-pixel, the centre pixel's full reference with BLA off */
    if (!escalated && f->use_bla && !f->no_periodic && (!shallow_fe || f->force_periodic)) {
        const mp_bitcnt_t np = mpf_get_prec(g.fc0_re);
        mpf_t nre, nim, t;
        mpf_init2(nre, np);
        mpf_init2(nim, np);
        mpf_init2(t, np);
        const int p = nuc_find(&g, f->scale, W, H, f->mxit, nre, nim);
        if (p > 2 && p <= f->mxit) {
            /* the nucleus's fractional pixel position (MP.cpp:1121-1127) */
            mpf_sub(t, nre, g.fc0_re);
            mpf_div(t, t, g.fdx);
            const double rx = mpf_get_d(t);
            mpf_sub(t, nim, g.fc0_im);
            mpf_div(t, t, g.fdy);
            const double ry = mpf_get_d(t);
            orbit_build_periodic(&o, nre, nim, p);
            if (mxo_fe_abs_less(nuc_size(&o, p), mxo_fe_mul_d(dxfe, 4.0))) {
                orbit_free(&o);
                f->subpixel = 1;
            } else {
                have_ref = 1;
                f->period = p;
                ref_virtual = 1;
                ref_x = rx;
                ref_y = ry;
            }
        }
        mpf_clear(nre);
        mpf_clear(nim);
        mpf_clear(t);
    }
    if (!have_ref) {
    if (escalated)
        grid_view_center(&g, W, H, rre, rim);
    else
        grid_pixel_c(&g, H / 2, W / 2, 0, 0, 1, rre, rim);
    orbit_build(&o, rre, rim, pb, f->mxit, escR, ede, 1, fxL);
    if (escalated) {
        ref_virtual = 1;
        ref_x = (W - 1) * 0.5;
        ref_y = (H - 1) * 0.5;
        if (o.escaped && o.ref_it >= f->mxit - 16) f->cutoff_sensitive_hint = 1;
        if (!o.escaped) f->cutoff_sensitive_hint = 2;
    } else if (!o.escaped || o.ref_it >= f->mxit - 16) {
        int sensitive = o.escaped || orbit_tail_escapes(&o, 80, esc2);
        if (sensitive) {
            orbit_free(&o);
            grid_view_center(&g, W, H, rre, rim);
            orbit_build(&o, rre, rim, pb, f->mxit, escR, ede, 1, fxL);
            ref_virtual = 1;
            ref_x = (W - 1) * 0.5;
            ref_y = (H - 1) * 0.5;
            if (o.escaped && o.ref_it >= f->mxit - 16) f->cutoff_sensitive_hint = 1;
            if (!o.escaped) f->cutoff_sensitive_hint = 2; /* needs a brute-force probe at mxit+64 */
        }
    }
    }
    const int use_bla = f->use_bla && !f->subpixel;
    f->ref_it = o.ref_it;
    f->ref_escaped = o.escaped;
    f->ref_virtual = ref_virtual;
    f->ref_x = ref_x;
    f->ref_y = ref_y;

    mxo_bla bla;
    memset(&bla, 0, sizeof bla);
    if (use_bla) {
        double dxf = fabs(mpfr_to_ld_trunc(g.dx)), dyf = fabs(mpfr_to_ld_trunc(g.dy));
        double a = ref_x, bx = (double)(W - 1) - ref_x;
        double mxx = (a > bx ? a : bx) + 0.5;
        double ay = ref_y, by = (double)(H - 1) - ref_y;
        double mxy = (ay > by ? ay : by) + 0.5;
        double dcmax = sqrt(dxf * dxf * mxx * mxx + dyf * dyf * mxy * mxy);
        bla_build(&bla, o.zf, o.dfr, o.dfi, o.ref_it, dcmax, f->bla_eps > 0 ? f->bla_eps : ldexp(1.0, -53), 0);
    }

    mxo_ctx ctx;
    ctx.zf = o.zf;
    ctx.zre_fe = o.zre_fe;
    ctx.zim_fe = o.zim_fe;
    ctx.mx_ref_it = o.ref_it;
    ctx.bla = use_bla ? &bla : NULL;
    ctx.per = f->period;
    ctx.bla_minlevel = bla_minlevel_from_skip(f->bla_minskip > 0 ? f->bla_minskip : 8);
    ctx.esc2 = esc2;
    ctx.use_interior = f->interior;
    ctx.interior_eps2 = 3e-9 * 3e-9;
    ctx.interior_confirm = 4;
    ctx.int_rep = 1;
    ctx.ede = ede;
    ctx.dxfe = dxfe;

    /* per-column / per-row dc (MP.cpp:2160-2163) */
    mxo_fe* dcol = (mxo_fe*)malloc(sizeof(mxo_fe) * W);
    mxo_fe* drow = (mxo_fe*)malloc(sizeof(mxo_fe) * H);
    for (int j = 0; j < W; ++j) dcol[j] = mxo_fe_mul_d(dxfe, (double)j - ref_x);
    for (int i = 0; i < H; ++i) drow[i] = mxo_fe_mul_d(dyfe, (double)i - ref_y);

    int* idx = (int*)malloc(sizeof(int) * (size_t)W * H);
    mxo_stats tot = {0, 0, 0, 0, 0, 0, 0};
    double t0 = omp_get_wtime();
    if (f->interior && f->probe) {
        int n = 0;
        for (int i = 0; i < H; i += 8)
            for (int j = 0; j < W; j += 8) idx[n++] = i * W + j;
        mxo_stats pst = {0, 0, 0, 0, 0, 0, 0};
        run_pixels(&ctx, f, dcol, drow, idx, n, f->val, f->n_esc, &pst);
        int nint = 0;
        for (int q = 0; q < n; ++q)
            if (f->val[idx[q]] == -2.0) ++nint;
        if (nint == 0) ctx.use_interior = 0;
    }
    f->interior_used = ctx.use_interior;
    int n = 0;
    const int cidx = (H / 2) * W + (W / 2);
    for (int p = 0; p < W * H; ++p)
        if (ref_virtual || p != cidx) idx[n++] = p;
    run_pixels(&ctx, f, dcol, drow, idx, n, f->val, f->n_esc, &tot);
    f->seconds_pixels = omp_get_wtime() - t0;
    if (!ref_virtual) {
        int ne;
        f->val[cidx] = ref_pixel_value(&o, f->mxit, f->c_method, esc2, g.dx, &ne);
        f->n_esc[cidx] = ne;
    }
    for (int p = 0; p < W * H; ++p) f->out[p] = (f->n_esc[p] < 0) ? -2.f : (float)f->val[p];
    f->ss_pixels = 0;
    if (f->sub > 1) {
        const int S = f->sub, h = S / 2;
        for (int i = 0; i < H; ++i)
            for (int j = 0; j < W; ++j) {
                const size_t c = ss_cell(W, S, i, j, 0, 0);
                f->cells[c] = f->out[i * W + j];
                f->cell_val[c] = f->val[i * W + j];
                f->cell_nesc[c] = f->n_esc[i * W + j];
            }
        if (f->c_method & 1) {
            unsigned char* flags = (unsigned char*)malloc((size_t)W * H);
            f->ss_pixels = ss_flag_pixels(f->out, W, H, f->mxit, ede, flags);
            /* every other cell of a flagged pixel against the same reference and BLA
             * table; dc from the fractional pixel position (MP.cpp:2160-2163) */
            const int per = S * S - 1;
            int np = 0;
            int* pix = (int*)malloc(sizeof(int) * (size_t)(f->ss_pixels + 1));
            for (int p = 0; p < W * H; ++p)
                if (flags[p]) pix[np++] = p;
            long long steps = 0, ap = 0, sk = 0, rb = 0;
#pragma omp parallel for schedule(dynamic, 8) num_threads(f->threads) reduction(+ : steps, ap, sk, rb)
            for (int q = 0; q < np * per; ++q) {
                const int p = pix[q / per], r = q % per;
                const int k = r < (per / 2) ? r : r + 1; /* skip the centre */
                const int yi = k / S - h, xi = k % S - h;
                const int i = p / W, j = p % W;
                const double px = j + (double)xi / S, py = i + (double)yi / S;
                mxo_stats st = {0, 0, 0, 0, 0, 0, 0};
                int ne;
                const double v = pixel_fe(&ctx, mxo_fe_mul_d(dxfe, px - ref_x), mxo_fe_mul_d(dyfe, py - ref_y), f->mxit, &ne, &st);
                const size_t c = ss_cell(W, S, i, j, yi, xi);
                f->cell_val[c] = v;
                f->cell_nesc[c] = ne;
                f->cells[c] = ne < 0 ? -2.f : (float)v;
                steps += st.steps;
                ap += st.bla_applies;
                sk += st.bla_skipped;
                rb += st.rebases;
            }
            tot.steps += steps;
            tot.bla_applies += ap;
            tot.bla_skipped += sk;
            tot.rebases += rb;
            free(pix);
            free(flags);
        }
    }
    /* cutoff-sensitive frames: a view-centre reference that escapes within 16 of
     * max_iter (MP.cpp:4144), or a bounded one whose brute force escapes within
     * max_iter + 64 (MP.cpp:4177); every computed sample whose float value is negative
     * or above max_iter - 16 is then recomputed by the brute force at its own c
     * (finalize, MP.cpp:2173-2183; accuratePixelCompute, MP.cpp:896-918) */
    f->cutoff_sensitive = f->cutoff_sensitive_hint == 1;
    if (f->cutoff_sensitive_hint == 2) {
        int ne;
        f->cutoff_sensitive = point_mpfr(rre, rim, pb, f->mxit + 64, 0, 1, esc2, g.dx, &ne) >= 0;
    }
    f->gmp_fallback = 0;
    if (f->cutoff_sensitive) {
        const int S = f->sub > 1 ? f->sub : 1, h = S / 2;
        const long ncell = (long)W * H * S * S;
        int nfb = 0;
#pragma omp parallel for schedule(dynamic, 1) num_threads(f->threads) reduction(+ : nfb)
        for (long c = 0; c < ncell; ++c) {
            const int row = (int)(c / ((long)W * S)), col = (int)(c % ((long)W * S));
            const int i = row / S, j = col / S, yi = row % S - h, xi = col % S - h;
            const int base = xi == 0 && yi == 0;
            float v;
            if (base)
                v = f->out[i * W + j];
            else if (f->sub > 1 && f->cell_nesc[c] != INT32_MIN)
                v = f->cells[c];
            else
                continue;
            if (!(v < 0 || v > f->mxit - 16)) continue;
            mpfr_t cr, ci;
            mpfr_inits2(pb, cr, ci, (mpfr_ptr)0);
            grid_pixel_c(&g, i, j, yi, xi, S, cr, ci);
            int ne;
            const double r = point_mpfr(cr, ci, pb, f->mxit, ede, 1, esc2, g.dx, &ne);
            mpfr_clears(cr, ci, (mpfr_ptr)0);
            const float rf = ne < 0 ? -2.f : (float)r;
            if (base) {
                f->val[i * W + j] = r;
                f->n_esc[i * W + j] = ne;
                f->out[i * W + j] = rf;
            }
            if (f->sub > 1) {
                f->cell_val[c] = r;
                f->cell_nesc[c] = ne;
                f->cells[c] = rf;
            }
            ++nfb;
        }
        f->gmp_fallback = nfb;
    }
    f->steps = tot.steps;
    f->bla_applies = tot.bla_applies;
    f->bla_skipped = tot.bla_skipped;
    f->rebases = tot.rebases;
    f->rescales = tot.rescales;
    f->underflow_steps = tot.underflow_steps;
    f->deriv_underflows = tot.deriv_underflows;
    if (f->orbit_zf)
        for (int k = 0; k < o.len; ++k) {
            f->orbit_zf[2 * k] = o.zf[2 * k];
            f->orbit_zf[2 * k + 1] = o.zf[2 * k + 1];
            f->orbit_m[2 * k] = o.zre_fe[k].m;
            f->orbit_e[2 * k] = o.zre_fe[k].e;
            f->orbit_m[2 * k + 1] = o.zim_fe[k].m;
            f->orbit_e[2 * k + 1] = o.zim_fe[k].e;
        }

    free(idx);
    free(dcol);
    free(drow);
    if (use_bla) bla_free(&bla);
    orbit_free(&o);
    mpfr_clears(rre, rim, (mpfr_ptr)0);
    grid_clear(&g);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* pieces exported for pinning tests                                          */
/* ------------------------------------------------------------------------- */

/* Direct FloatExp entry points (tests compare them with exact rationals). op: 0 mul,
 * 1 mul_d, 2 add, 3 sub, 4 div, 5 sqrt(a), 6 from(a.m), 7 to_double(a), 8 abs_less */
MXO_EXPORT void mxo_fe_op(int op, int n, const double* am, const int64_t* ae, const double* bm, const int64_t* be, double* rm, int64_t* re) {
    for (int k = 0; k < n; ++k) {
        mxo_fe a = mxo_fe_make(am[k], ae[k]), b = mxo_fe_make(bm[k], be[k]), r = mxo_fe_make(0, 0);
        switch (op) {
        case 0: r = mxo_fe_mul(a, b); break;
        case 1: r = mxo_fe_mul_d(a, bm[k]); break;
        case 2: r = mxo_fe_add(a, b); break;
        case 3: r = mxo_fe_sub(a, b); break;
        case 4: r = mxo_fe_div(a, b); break;
        case 5: r = mxo_fe_sqrt(a); break;
        case 6: r = mxo_fe_from(am[k]); break;
        case 7: r = mxo_fe_make(mxo_fe_to_double(a), 0); break;
        case 8: r = mxo_fe_make((double)mxo_fe_abs_less(a, b), 0); break;
        }
        rm[k] = r.m;
        re[k] = r.e;
    }
}

/* The cold recoveries on explicit inputs, for tests.  step: in X (re, im), dc (re, im),
 * S, w (re, im) -> out S, w.  deriv: in z (re, im), SJ, j (re, im) -> out SJ, j, 1/SJ.
 * FloatExp values as (mantissa, exponent) pairs. */
MXO_EXPORT void mxo_step_underflow(const double* xm, const int64_t* xe, const double* cm, const int64_t* ce, double sm, int64_t se, double wr, double wi, double* out_sm, int64_t* out_se, double* out_w) {
    mxo_fe S = mxo_fe_make(sm, se);
    fe_step_underflow(mxo_fe_make(xm[0], xe[0]), mxo_fe_make(xm[1], xe[1]), mxo_fe_make(cm[0], ce[0]), mxo_fe_make(cm[1], ce[1]), &S, &wr, &wi);
    *out_sm = S.m;
    *out_se = S.e;
    out_w[0] = wr;
    out_w[1] = wi;
}

MXO_EXPORT void mxo_deriv_underflow(const double* zm, const int64_t* ze, double sjm, int64_t sje, double jr, double ji, double* out_sjm, int64_t* out_sje, double* out_j) {
    mxo_fe SJ = mxo_fe_make(sjm, sje);
    double inv = 0.0;
    fe_deriv_underflow(mxo_fe_make(zm[0], ze[0]), mxo_fe_make(zm[1], ze[1]), &SJ, &jr, &ji, &inv);
    *out_sjm = SJ.m;
    *out_sje = SJ.e;
    out_j[0] = jr;
    out_j[1] = ji;
    out_j[2] = inv;
}

/* Reference orbit shadows for a point given as decimal strings, in MPFR at prec_bits
 * (fx_limbs == 0) or in the fx_limbs-limb fixed point from the point held at
 * prec_bits. Returns ref_it. */
MXO_EXPORT int mxo_orbit_shadows(const char* cre, const char* cim, int prec_bits, int fx_limbs, int mxit, double escape_radius, int deriv, double* zf, double* zre_m, int64_t* zre_e, double* zim_m, int64_t* zim_e, double* dfr, double* dfi) {
    mpfr_t a, b;
    mpfr_inits2(prec_bits, a, b, (mpfr_ptr)0);
    mpfr_set_str(a, cre, 10, MPFR_RNDN);
    mpfr_set_str(b, cim, 10, MPFR_RNDN);
    mxo_orbit o;
    orbit_build(&o, a, b, prec_bits, mxit, escape_radius, deriv, 1, fx_limbs);
    for (int k = 0; k < o.len; ++k) {
        zf[2 * k] = o.zf[2 * k];
        zf[2 * k + 1] = o.zf[2 * k + 1];
        if (zre_m) {
            zre_m[k] = o.zre_fe[k].m;
            zre_e[k] = o.zre_fe[k].e;
            zim_m[k] = o.zim_fe[k].m;
            zim_e[k] = o.zim_fe[k].e;
        }
        if (dfr) {
            dfr[k] = o.dfr[k];
            dfi[k] = o.dfi[k];
        }
    }
    int r = o.ref_it;
    orbit_free(&o);
    mpfr_clears(a, b, (mpfr_ptr)0);
    return r;
}

/* One fixed-point product and one shadow on L-limb magnitudes (least significant limb
 * first) with separate signs, for pinning the arithmetic against exact integers. */
MXO_EXPORT void mxo_fixed_mul(int L, const uint64_t* a, int sa, const uint64_t* b, int sb, uint64_t* r, int* sr) {
    mxo_fx F;
    fx_init(&F, L);
    mpz_t x, y, z;
    mpz_inits(x, y, z, (mpz_ptr)0);
    mpz_import(x, (size_t)L, -1, sizeof(uint64_t), 0, 0, a);
    mpz_import(y, (size_t)L, -1, sizeof(uint64_t), 0, 0, b);
    if (sa < 0) mpz_neg(x, x);
    if (sa == 0) mpz_set_ui(x, 0);
    if (sb < 0) mpz_neg(y, y);
    if (sb == 0) mpz_set_ui(y, 0);
    fx_mul(&F, z, x, y);
    memset(r, 0, sizeof(uint64_t) * (size_t)L);
    *sr = mpz_sgn(z);
    /* the engine keeps L limbs; the column above them is zero for |result| < 2^64 */
    mpz_abs(z, z);
    mpz_tdiv_r_2exp(z, z, (mp_bitcnt_t)64 * L);
    mpz_export(r, NULL, -1, sizeof(uint64_t), 0, 0, z);
    mpz_clears(x, y, z, (mpz_ptr)0);
    fx_clear(&F);
}

MXO_EXPORT void mxo_fixed_to_fe(int L, const uint64_t* a, int sa, double* m, int64_t* e) {
    mpz_t x;
    mpz_init(x);
    mpz_import(x, (size_t)L, -1, sizeof(uint64_t), 0, 0, a);
    if (sa < 0) mpz_neg(x, x);
    if (sa == 0) mpz_set_ui(x, 0);
    const mxo_fe f = fx_to_fe(x, L);
    *m = f.m;
    *e = f.e;
    mpz_clear(x);
}

/* BLA table from a double orbit shadow; writes r2/l/coef (caller sizes them with
 * mxo_bla_total). Returns the number of levels. */
MXO_EXPORT int64_t mxo_bla_total(int reflen) {
    if (reflen < 3) return 0;
    int64_t n = reflen - 1, tot = 0;
    for (;;) {
        tot += n;
        if (n <= 1) break;
        n /= 2;
    }
    return tot;
}

MXO_EXPORT int mxo_bla_tables(const double* zf, const double* dfr, const double* dfi, int reflen, double dcmax, double eps, int ede, double* r2, int64_t* l, double* coef, double* der, double* rmax2, int64_t* lev_off, int64_t* lev_size) {
    mxo_bla b;
    bla_build(&b, zf, dfr, dfi, reflen, dcmax, eps, ede);
    if (b.nlev == 0) return 0;
    memcpy(r2, b.r2, sizeof(double) * b.total);
    memcpy(l, b.l, sizeof(int64_t) * b.total);
    memcpy(coef, b.coef, sizeof(double) * 4 * b.total);
    if (ede && der) memcpy(der, b.der, sizeof(double) * 4 * b.total);
    *rmax2 = b.rmax2;
    for (int p = 0; p < b.nlev; ++p) {
        lev_off[p] = b.off[p];
        lev_size[p] = b.size[p];
    }
    int nl = b.nlev;
    bla_free(&b);
    return nl;
}

/* Per-pixel rescaled loop on caller-supplied orbit/BLA (the same arrays the GPU
 * consumed), so a parity test isolates the kernel. */
typedef struct {
    const double* zf;
    const double* zre_m;
    const int64_t* zre_e;
    const double* zim_m;
    const int64_t* zim_e;
    int mx_ref_it;
    /* BLA (nlev == 0 => off) */
    int nlev;
    const int64_t* lev_off;
    const int64_t* lev_size;
    const double* r2;
    const int64_t* l;
    const double* coef;
    double rmax2;
    int bla_minlevel;
    double esc2;
    int use_interior;
    int ede;
    double dx_m;
    int64_t dx_e;
    int mxit;
    int threads;
    int period; /* > 0: a periodic reference of `period` entries (mx_ref_it == period) */
} mxo_pixels_in;

MXO_EXPORT void mxo_pixels_fe(const mxo_pixels_in* in, int n, const double* dcr_m, const int64_t* dcr_e, const double* dci_m, const int64_t* dci_e, double* val, int32_t* nesc, long long* stats4) {
    int64_t len = in->period > 0 ? in->period : in->mx_ref_it + 1;
    mxo_fe* zre = (mxo_fe*)malloc(sizeof(mxo_fe) * len);
    mxo_fe* zim = (mxo_fe*)malloc(sizeof(mxo_fe) * len);
    for (int64_t k = 0; k < len; ++k) {
        zre[k] = mxo_fe_make(in->zre_m[k], in->zre_e[k]);
        zim[k] = mxo_fe_make(in->zim_m[k], in->zim_e[k]);
    }
    mxo_bla b;
    memset(&b, 0, sizeof b);
    if (in->nlev > 0) {
        b.nlev = in->nlev;
        for (int p = 0; p < in->nlev; ++p) {
            b.off[p] = in->lev_off[p];
            b.size[p] = in->lev_size[p];
        }
        b.r2 = (double*)in->r2;
        b.l = (int64_t*)in->l;
        b.coef = (double*)in->coef;
        b.rmax2 = in->rmax2;
    }
    mxo_ctx ctx;
    ctx.zf = in->zf;
    ctx.zre_fe = zre;
    ctx.zim_fe = zim;
    ctx.mx_ref_it = in->mx_ref_it;
    ctx.bla = in->nlev > 0 ? &b : NULL;
    ctx.bla_minlevel = in->bla_minlevel;
    ctx.esc2 = in->esc2;
    ctx.use_interior = in->use_interior;
    ctx.interior_eps2 = 3e-9 * 3e-9;
    ctx.interior_confirm = 4;
    ctx.int_rep = 1;
    ctx.ede = in->ede;
    ctx.dxfe = mxo_fe_make(in->dx_m, in->dx_e);
    ctx.per = in->period > 0 ? in->period : 0;
    long long steps = 0, ap = 0, sk = 0, rb = 0, rs = 0, us = 0, du = 0;
#pragma omp parallel for schedule(dynamic, 16) num_threads(in->threads > 0 ? in->threads : 1) reduction(+ : steps, ap, sk, rb, rs, us, du)
    for (int q = 0; q < n; ++q) {
        mxo_stats st = {0, 0, 0, 0, 0, 0, 0};
        int ne;
        val[q] = pixel_fe(&ctx, mxo_fe_make(dcr_m[q], dcr_e[q]), mxo_fe_make(dci_m[q], dci_e[q]), in->mxit, &ne, &st);
        nesc[q] = ne;
        steps += st.steps;
        ap += st.bla_applies;
        sk += st.bla_skipped;
        rb += st.rebases;
        rs += st.rescales;
        us += st.underflow_steps;
        du += st.deriv_underflows;
    }
    if (stats4) { /* steps, applies, skipped, rebases, rescales, step underflows, derivative underflows */
        stats4[0] = steps;
        stats4[1] = ap;
        stats4[2] = sk;
        stats4[3] = rb;
        stats4[4] = rs;
        stats4[5] = us;
        stats4[6] = du;
    }
    free(zre);
    free(zim);
}

/* c lies inside the main cardioid or the period-2 disc by at least 2^(8 - prec):
 * q (q + x - 1/4) < y^2 / 4 with q = (x - 1/4)^2 + y^2, or (x + 1)^2 + y^2 < 1/16,
 * evaluated in MPFR with that margin for its rounding, so a point near either boundary
 * is never claimed.  Such a c has an attracting fixed point or 2-cycle: its orbit stays
 * bounded and never escapes. */
static int in_main_bulbs(mpfr_srcptr x, mpfr_srcptr y) {
    const mpfr_prec_t p = mpfr_get_prec(x);
    mpfr_t q, t, y2, m;
    mpfr_inits2(p + 16, q, t, y2, m, (mpfr_ptr)0);
    mpfr_set_ui_2exp(m, 1, 8 - (long)p, MPFR_RNDN);
    mpfr_sqr(y2, y, MPFR_RNDN);
    mpfr_sub_d(t, x, 0.25, MPFR_RNDN);
    mpfr_sqr(q, t, MPFR_RNDN);
    mpfr_add(q, q, y2, MPFR_RNDN);
    mpfr_add(t, t, q, MPFR_RNDN);
    mpfr_mul(t, t, q, MPFR_RNDN);  /* q (q + x - 1/4) */
    mpfr_add(t, t, m, MPFR_RNDN);
    mpfr_div_2ui(q, y2, 2, MPFR_RNDN);
    int in = mpfr_less_p(t, q);
    if (!in) {
        mpfr_add_ui(t, x, 1, MPFR_RNDN);
        mpfr_sqr(t, t, MPFR_RNDN);
        mpfr_add(t, t, y2, MPFR_RNDN);
        mpfr_add(t, t, m, MPFR_RNDN);
        in = mpfr_cmp_d(t, 0.0625) < 0;
    }
    mpfr_clears(q, t, y2, m, (mpfr_ptr)0);
    return in;
}

static void direct_grid(const char* cre, const char* cim, const char* scale, int W, int H, int prec_bits, int wp_bits, int mxit, int c_method, float escape_radius, int step, int threads, int track_d, int skip_bulbs, double* val, int32_t* nesc) {
    mxo_grid g;
    grid_init(&g, cre, cim, scale, W, H, 0, 0, prec_bits > wp_bits ? prec_bits : wp_bits, 0);
    const double esc2 = (double)(escape_radius * escape_radius);
    int ny = (H + step - 1) / step, nx = (W + step - 1) / step;
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads > 0 ? threads : 1)
    for (int q = 0; q < nx * ny; ++q) {
        int i = (q / nx) * step, j = (q % nx) * step;
        mpfr_t a, b;
        mpfr_inits2(g.prec, a, b, (mpfr_ptr)0);
        grid_pixel_c(&g, i, j, 0, 0, 1, a, b);
        int ne;
        if (skip_bulbs && in_main_bulbs(a, b)) {
            val[i * W + j] = -2.0;
            ne = -2;
        } else
            val[i * W + j] = point_mpfr(a, b, wp_bits, mxit, (c_method & 2) != 0, track_d, esc2, g.dx, &ne);
        nesc[i * W + j] = ne;
        mpfr_clears(a, b, (mpfr_ptr)0);
    }
    grid_clear(&g);
}

/* Brute-force pixel grid for a frame (ComputeDirect, MP.cpp:959-990): pixels on a
 * (step x step) grid; others untouched.  wp_bits = working precision. */
MXO_EXPORT void mxo_direct(const char* cre, const char* cim, const char* scale, int W, int H, int prec_bits, int wp_bits, int mxit, int c_method, float escape_radius, int step, int threads, int track_d, double* val, int32_t* nesc) {
    direct_grid(cre, cim, scale, W, H, prec_bits, wp_bits, mxit, c_method, escape_radius, step, threads, track_d, 0, val, nesc);
}

/* The same grid, with the pixels inside the main cardioid or the period-2 disc marked
 * interior (-2) without iterating: the escaping pixels and their values are those of
 * mxo_direct.  For the boundary-walk probes, which only rank escaping pixels. */
MXO_EXPORT void mxo_direct_skip_bulbs(const char* cre, const char* cim, const char* scale, int W, int H, int prec_bits, int wp_bits, int mxit, int c_method, float escape_radius, int step, int threads, int track_d, double* val, int32_t* nesc) {
    direct_grid(cre, cim, scale, W, H, prec_bits, wp_bits, mxit, c_method, escape_radius, step, threads, track_d, 1, val, nesc);
}

/* Brute-force samples idx[0..n) (row-major in the W x H grid of the frame) */
MXO_EXPORT void mxo_direct_list(const char* cre, const char* cim, const char* scale, int W, int H, int prec_bits, int wp_bits, int mxit, int c_method, float escape_radius, const int32_t* idx, int n, int threads, int track_d, double* val, int32_t* nesc) {
    mxo_grid g;
    grid_init(&g, cre, cim, scale, W, H, 0, 0, prec_bits > wp_bits ? prec_bits : wp_bits, 0);
    const double esc2 = (double)(escape_radius * escape_radius);
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads > 0 ? threads : 1)
    for (int q = 0; q < n; ++q) {
        const int i = idx[q] / W, j = idx[q] % W;
        mpfr_t a, b;
        mpfr_inits2(g.prec, a, b, (mpfr_ptr)0);
        grid_pixel_c(&g, i, j, 0, 0, 1, a, b);
        int ne;
        val[q] = point_mpfr(a, b, wp_bits, mxit, (c_method & 2) != 0, track_d, esc2, g.dx, &ne);
        nesc[q] = ne;
        mpfr_clears(a, b, (mpfr_ptr)0);
    }
    grid_clear(&g);
}

/* ------------------------------------------------------------------------- */
/* shallow direct path (scale <= 1e6)                                         */
/* ------------------------------------------------------------------------- */

/* floatPointCompute for smooth and exterior-DE colouring (MP.cpp:330-399): z = c,
 * d = 2 (dz/dz0, the interior test), dc = 1 (dz/dc, EDE); per iteration all three from
 * the old z; escape |z|^2 > esc2 first, then |d|^2 < 1e-9 (interior), and -2 once the
 * counter i reaches mxit.  Returns the double the engine casts to float: the smooth
 * count i + 1 - log2(log(|z|^2)/2/log 2), or the raw distance estimate, -2 interior
 * (n = escape count i + 1, -2 interior, -1 max_iter). */
static double shallow_scalar(double c_re, double c_im, int mxit, double esc2, int ede, int* n) {
    double z_re = c_re, z_im = c_im, d_re = 2, d_im = 0, dc_re = 1, dc_im = 0, tmp;
    int i = 1;
    while (i < mxit) {
        tmp = 2.0 * (d_re * z_re - d_im * z_im);
        d_im = 2.0 * (d_re * z_im + d_im * z_re);
        d_re = tmp;
        if (ede) {
            tmp = 2.0 * (dc_re * z_re - dc_im * z_im) + 1;
            dc_im = 2.0 * (dc_re * z_im + dc_im * z_re);
            dc_re = tmp;
        }
        tmp = z_re * z_re - z_im * z_im + c_re;
        z_im = 2.0 * z_re * z_im + c_im;
        z_re = tmp;
        tmp = z_re * z_re + z_im * z_im;
        if (tmp > esc2) {
            *n = i + 1;
            if (ede) return sqrt(tmp) * log(tmp) / sqrt(dc_re * dc_re + dc_im * dc_im);
            return i + 1 - log(log(tmp) / 2 / log(2)) / log(2);
        }
        tmp = d_re * d_re + d_im * d_im;
        if (tmp < 0.000000001) {
            *n = -2;
            return -2;
        }
        ++i;
    }
    *n = -1;
    return -2;
}

/* One lane of solveShallowSimdList (MP.cpp:416-532), the engine's default for these
 * colourings: the same recurrences with the counter jv kept in double and the three
 * exits tested after each iteration (escape, then interior, then jv + 1 >= mxit).
 * Returns the value before the lane rounds it to float for its float output.  For
 * mxit >= 2 its values equal shallow_scalar's (the engine checks that on four points,
 * verify.cpp:9980-9990). */
static double shallow_lane(double cr, double ci, int mxit, double esc2, int ede, int* n) {
    const double LG2 = log(2.0);
    double zr = cr, zi = ci, dr = 2, di = 0, dcr = 1, dci = 0, jv = 1;
    for (;;) {
        const double ndr = 2.0 * (dr * zr - di * zi), ndi = 2.0 * (dr * zi + di * zr);
        const double ndcr = 2.0 * (dcr * zr - dci * zi) + 1.0, ndci = 2.0 * (dcr * zi + dci * zr);
        const double nzr = (zr * zr - zi * zi) + cr, nzi = 2.0 * (zr * zi) + ci;
        zr = nzr;
        zi = nzi;
        dr = ndr;
        di = ndi;
        dcr = ndcr;
        dci = ndci;
        const double zrad = zr * zr + zi * zi, dmag = dr * dr + di * di, jnext = jv + 1.0;
        if (zrad > esc2) {
            *n = (int)jv + 1;
            if (ede) return sqrt(zrad) * log(zrad) / sqrt(dcr * dcr + dci * dci);
            return jv + 1 - log(log(zrad) / 2 / LG2) / LG2;
        }
        if (dmag < 1e-9) {
            *n = -2;
            return -2.0;
        }
        if (jnext >= (double)mxit) {
            *n = -1;
            return -2.0;
        }
        jv = jnext;
    }
}

/* Both loops on explicit points (scalar = 1: floatPointCompute, 0: the lane); val is
 * the unrounded value. */
MXO_EXPORT void mxo_shallow_points(const double* cr, const double* ci, int n, int mxit, double esc2, int ede, int scalar, double* val, int32_t* nesc) {
    for (int k = 0; k < n; ++k) {
        int ne = 0;
        val[k] = scalar ? shallow_scalar(cr[k], ci[k], mxit, esc2, ede, &ne) : shallow_lane(cr[k], ci[k], mxit, esc2, ede, &ne);
        nesc[k] = ne;
    }
}

/* A shallow view (Mandel::Compute's method 0, MP.cpp:1444-1519): c0, dx, dy from the
 * engine's mpf grid truncated to double (mpf_get_ld); base pixel (i, j) at
 * c0 + (dx j, dy i); with list != NULL the adaptive pass instead (MP.cpp:1933-1965): the
 * sub*sub - 1 cells of each listed pixel (row-major index), xi outer and yi inner,
 * at (c0r + dx j) + dx xi / sub, (c0i + dy i) + dy yi / sub.  Outputs per sample: val
 * is a double whose float is the engine's output (the smooth value; under EDE the
 * lane's float distance divided by dx in double, MP.cpp:1497-1499, 1952-1954), nesc as
 * shallow_scalar's n.  grid4 (optional) receives c0r, c0i, dx, dy. */
MXO_EXPORT void mxo_shallow(const char* cre, const char* cim, const char* scale, int W, int H, int full_h, int row_base, int prec_bits, int mxit, float escape_radius, int ede, int sub, const int32_t* list, int nlist, int threads, double* grid4, double* val, int32_t* nesc) {
    mxo_grid g;
    grid_init(&g, cre, cim, scale, W, H, full_h, row_base, mpf_bits(prec_bits), 1);
    const double c0r = mpf_get_ld_c(g.fc0_re), c0i = mpf_get_ld_c(g.fc0_im), dx = mpf_get_ld_c(g.fdx), dy = mpf_get_ld_c(g.fdy);
    grid_clear(&g);
    if (grid4) {
        grid4[0] = c0r;
        grid4[1] = c0i;
        grid4[2] = dx;
        grid4[3] = dy;
    }
    const double esc2 = (double)(escape_radius * escape_radius);
    const int per = sub * sub - 1, h = sub / 2;
    const long total = list ? (long)nlist * per : (long)W * H;
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 64)
    for (long q = 0; q < total; ++q) {
        double cr, ci;
        if (list) {
            const int pix = list[q / per];
            int t = (int)(q % per);
            if (t >= h * sub + h) ++t;
            const int xi = t / sub - h, yi = t % sub - h, i = pix / W, j = pix % W;
            cr = c0r + dx * j + dx * xi / sub;
            ci = c0i + dy * i + dy * yi / sub;
        } else {
            const int i = (int)(q / W), j = (int)(q % W);
            cr = c0r + dx * j;
            ci = c0i + dy * i;
        }
        int ne = 0;
        double v = shallow_lane(cr, ci, mxit, esc2, ede, &ne);
        if (ede && ne > 0) v = (double)(float)v / dx;
        val[q] = v;
        nesc[q] = ne;
    }
}

/* Brute-force single point from decimal strings. */
MXO_EXPORT double mxo_point(const char* cre, const char* cim, int wp_bits, int mxit, int ede, int track_d, double esc2, const char* dx, int* nesc) {
    mpfr_t a, b, d;
    mpfr_inits2(wp_bits, a, b, d, (mpfr_ptr)0);
    mpfr_set_str(a, cre, 10, MPFR_RNDN);
    mpfr_set_str(b, cim, 10, MPFR_RNDN);
    mpfr_set_str(d, dx ? dx : "1", 10, MPFR_RNDN);
    double v = point_mpfr(a, b, wp_bits, mxit, ede, track_d, esc2, d, nesc);
    mpfr_clears(a, b, d, (mpfr_ptr)0);
    return v;
}

/* Grid values for tests: truncated dx/dy as FloatExp and c0 as strings. */
MXO_EXPORT void mxo_grid_info(const char* cre, const char* cim, const char* scale, int W, int H, int full_h, int row_base, int prec_bits, double* dx_m, int64_t* dx_e, double* dy_m, int64_t* dy_e) {
    prec_bits = mpf_bits(prec_bits);
    mxo_grid g;
    grid_init(&g, cre, cim, scale, W, H, full_h, row_base, prec_bits, 1);
    mxo_fe a = mpfr_to_fe_trunc(g.dx), b = mpfr_to_fe_trunc(g.dy);
    *dx_m = a.m;
    *dx_e = a.e;
    *dy_m = b.m;
    *dy_e = b.e;
    grid_clear(&g);
}

/* findNucleus and nucleusSizeMag for one view: the period (0: none in the search
 * disk), the nucleus as decimal strings (digits with an implied point before them,
 * and the exponent: value = 0.<digits> * 10^exp) and the island size in FloatExp */
MXO_EXPORT int mxo_nucleus(const char* cre, const char* cim, const char* scale, int W, int H, int maxp, int prec_bits, char* re_digits, long* re_exp, char* im_digits, long* im_exp, int ndig, double* size_m, int64_t* size_e) {
    prec_bits = mpf_bits(prec_bits);
    mxo_grid g;
    grid_init(&g, cre, cim, scale, W, H, 0, 0, prec_bits, 1);
    const mp_bitcnt_t np = mpf_get_prec(g.fc0_re);
    mpf_t nre, nim;
    mpf_init2(nre, np);
    mpf_init2(nim, np);
    const int p = nuc_find(&g, scale, W, H, maxp, nre, nim);
    *size_m = 0.0;
    *size_e = 0;
    if (p > 0) {
        mpf_get_str(re_digits, re_exp, 10, (size_t)ndig - 2, nre);
        mpf_get_str(im_digits, im_exp, 10, (size_t)ndig - 2, nim);
        mxo_orbit o;
        orbit_build_periodic(&o, nre, nim, p);
        const mxo_fe sz = nuc_size(&o, p);
        *size_m = sz.m;
        *size_e = sz.e;
        orbit_free(&o);
    }
    mpf_clear(nre);
    mpf_clear(nim);
    grid_clear(&g);
    return p;
}

/* ========================================================================= */
/* plain-double perturbation path (scale 1e6..1e280)                          */
/* ========================================================================= */

#define MXO_SA_N 30

typedef struct {
    double r, i;
} mxo_c;

/* The engine uses std::complex<double> here.  Its rounding is library-specific
 * (MSVC vs libstdc++), so this restatement fixes explicit formulas, shared with the
 * GPU path: textbook multiplication, and the Smith division of libgcc's __divdc3
 * for finite operands. */
static inline mxo_c c_add(mxo_c a, mxo_c b) {
    mxo_c r = {a.r + b.r, a.i + b.i};
    return r;
}
static inline mxo_c c_mul(mxo_c a, mxo_c b) {
    mxo_c r = {a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r};
    return r;
}
static inline mxo_c c_div(mxo_c a, mxo_c b) {
    mxo_c r;
    if (fabs(b.r) < fabs(b.i)) {
        double ratio = b.r / b.i;
        double denom = (b.r * ratio) + b.i;
        r.r = ((a.r * ratio) + a.i) / denom;
        r.i = ((a.i * ratio) - a.r) / denom;
    } else {
        double ratio = b.i / b.r;
        double denom = (b.i * ratio) + b.r;
        r.r = ((a.i * ratio) + a.r) / denom;
        r.i = (a.i - (a.r * ratio)) / denom;
    }
    return r;
}

/* get_exp(long double) (float_math.cpp:24-29) */
static inline int sa_get_exp(double x) {
    if (x == 0.0) return -0x40000000;
    int e;
    (void)frexp(x, &e);
    return e - 1;
}

/* SACheckMagnitude (MP.cpp:4307-4332) */
static int sa_check(const mxo_c* A) {
    int pre_mn[MXO_SA_N], suf_mx[MXO_SA_N];
    int re = sa_get_exp(A[0].r), im = sa_get_exp(A[0].i);
    pre_mn[0] = re < im ? re : im;
    re = sa_get_exp(A[MXO_SA_N - 1].r);
    im = sa_get_exp(A[MXO_SA_N - 1].i);
    suf_mx[MXO_SA_N - 1] = re > im ? re : im;
    for (int i = 1; i < MXO_SA_N; ++i) {
        re = sa_get_exp(A[i].r);
        im = sa_get_exp(A[i].i);
        int mn = re < im ? re : im;
        pre_mn[i] = pre_mn[i - 1] > mn ? pre_mn[i - 1] : mn;
    }
    for (int i = MXO_SA_N - 2; i >= 0; --i) {
        re = sa_get_exp(A[i].r);
        im = sa_get_exp(A[i].i);
        int mx = re > im ? re : im;
        suf_mx[i] = suf_mx[i + 1] > mx ? suf_mx[i + 1] : mx;
    }
    for (int i = 0; i < MXO_SA_N - 10; ++i)
        if (pre_mn[i] - suf_mx[i + 1] >= 80) return i > 5 ? i : 5;
    return -1;
}

typedef struct {
    mxo_c A[MXO_SA_N]; /* _Adf_old */
    mxo_c B[MXO_SA_N]; /* _Bdf_old: series of the derivative delta (EDE) */
    mxo_c delta;       /* _SA_delta */
    int order, it, flag, ede;
} mxo_sa;

/* One series step (calCoefficient's SA block, MP.cpp:4273-4302): z = _zf[i-1],
 * d = _df[i-1] (EDE only). */
static void sa_step(mxo_sa* sa, mxo_c z, mxo_c d, int i) {
    if (!sa->flag) return;
    mxo_c N[MXO_SA_N], NB[MXO_SA_N];
    mxo_c two = {2.0, 0.0};
    mxo_c z2 = c_mul(two, z);
    for (int j = 0; j < MXO_SA_N; ++j) {
        N[j] = c_mul(z2, sa->A[j]);
        for (int k = 0; k < j / 2; ++k) N[j] = c_add(N[j], c_mul(c_mul(two, sa->A[k]), sa->A[j - k - 1]));
        if (j % 2) N[j] = c_add(N[j], c_mul(sa->A[j / 2], sa->A[j / 2]));
    }
    if (sa->ede) {
        /* B'_j = 2 (B_j Z + A_j D + sum_{k<j} A_k B_{j-k-1}) */
        for (int j = 0; j < MXO_SA_N; ++j) {
            NB[j] = c_add(c_mul(sa->B[j], z), c_mul(sa->A[j], d));
            for (int k = 0; k < j; ++k) NB[j] = c_add(NB[j], c_mul(sa->A[k], sa->B[j - k - 1]));
            NB[j] = c_mul(NB[j], two);
        }
    }
    N[0] = c_add(N[0], sa->delta);
    int order = sa_check(N);
    if (order < 0) {
        sa->flag = 0;
    } else {
        sa->it = i;
        sa->order = order;
        memcpy(sa->A, N, sizeof N);
        if (sa->ede) memcpy(sa->B, NB, sizeof NB);
    }
}

/* Horner seeds of the double loop (MP.cpp:2319-2330): dz from A, dd from B */
static mxo_c sa_horner(const mxo_c* C, int order, mxo_c delta, mxo_c dc) {
    mxo_c dz = {0.0, 0.0};
    for (int x = order; x >= 0; --x) {
        dz = c_add(dz, C[x]);
        dz = c_mul(dz, dc);
        dz = c_div(dz, delta);
    }
    return dz;
}
static mxo_c sa_seed(const mxo_sa* sa, mxo_c dc) { return sa_horner(sa->A, sa->order, sa->delta, dc); }

/* tryBLA (MP.cpp:2714-2755); under EDE the derivative delta is carried,
 * dd -> C dz + A dd + E dc, with the dz from before the skip (2742-2749). */
static int try_bla_d(const mxo_bla* b, int minlevel, const double* zf, int s, double* dzr, double* dzi, double* ddr, double* ddi, double dcr, double dci, int ede, double esc2, int mx_ref_it) {
    if (s < 1 || !b || b->nlev == 0) return 0;
    double zmag2 = *dzr * *dzr + *dzi * *dzi;
    int maxp = b->nlev - 1;
    int startp = (s == 1) ? maxp : tzcnt32((unsigned)(s - 1));
    if (startp > maxp) startp = maxp;
    if (startp < minlevel) return 0;
    for (int p = startp; p >= minlevel; --p) {
        int64_t i = (int64_t)((s - 1) >> p);
        if (i >= b->size[p]) continue;
        int64_t k = b->off[p] + i;
        double r2 = b->r2[k];
        if (r2 <= 0 || zmag2 >= r2) continue;
        int land = s + (int)b->l[k];
        if (land >= mx_ref_it - 1) continue;
        const double* c = b->coef + 4 * k;
        double nzr = c[0] * *dzr - c[1] * *dzi + c[2] * dcr - c[3] * dci;
        double nzi = c[0] * *dzi + c[1] * *dzr + c[2] * dci + c[3] * dcr;
        double fr = zf[2 * land] + nzr, fi = zf[2 * land + 1] + nzi;
        if (fr * fr + fi * fi > esc2) return 0;
        if (ede) {
            const double* d = b->der + 4 * k;
            double nddr = (d[0] * *dzr - d[1] * *dzi) + (c[0] * *ddr - c[1] * *ddi) + (d[2] * dcr - d[3] * dci);
            double nddi = (d[0] * *dzi + d[1] * *dzr) + (c[0] * *ddi + c[1] * *ddr) + (d[2] * dci + d[3] * dcr);
            *ddr = nddr;
            *ddi = nddi;
        }
        *dzr = nzr;
        *dzi = nzi;
        return (int)b->l[k];
    }
    return 0;
}

typedef struct {
    const double* zf;  /* interleaved double orbit, reflen entries */
    const double* z_m3; /* |Z|^2/1e6 per entry */
    const double* dfr; /* reference derivative D (EDE) */
    const double* dfi;
    int mx_ref_it;
    const mxo_bla* bla;
    int bla_minlevel;
    double esc2;
    int use_interior;
    double interior_eps2;
    int interior_confirm;
    int int_rep;
    int sa_it;
    int ede;
    double dxf;        /* mpf_get_ld(_dx), the DE normalisation */
} mxo_dctx;

/* The scalar double loop of stepParallel (MP.cpp:2331-2584): smooth output, or the
 * pixel-normalised exterior distance estimate (2444-2456, 2487-2488) under EDE.
 * Returns the value; *n_out = escape iteration, -1 reached mxit, -2 interior
 * detected, -3 unresolved (glitch or end of reference; *glitch_p set). */
static double pixel_double(const mxo_dctx* c, mxo_c dc, mxo_c dz0, mxo_c dd0, int mxit, int* n_out, double* glitch_p, mxo_stats* st) {
    const double* Z = c->zf;
    const int mx_ref_it = c->mx_ref_it;
    const int ede = c->ede;
    int j = c->sa_it + 1;
    double dzr = dz0.r, dzi = dz0.i, dcr = dc.r, dci = dc.i;
    double ddr = dd0.r, ddi = dd0.i, dr = 0.0, di = 0.0;
    int k = j;
    double zsr = 1e30, zsi = 1e30;
    int save_j = j, period_win = 1;
    int conf_P = 0, conf_next = 0, conf_count = 0, conf_giveup = 0;
    double conf_D2 = 0, conf_zr = 0, conf_zi = 0;
    const int detect_interior = c->use_interior;
    const int int_rep = c->int_rep && detect_interior;
    double sdzr = 1e300, sdzi = 1e300;
    int sk = -1, save_j2 = j, twin2 = 1;
    *glitch_p = -1;
    const double LG = log(2);
    while (j < mxit) {
        if (int_rep) {
            if (k == sk && dzr == sdzr && dzi == sdzi) {
                *n_out = -2;
                return -2.0;
            }
            if (j - save_j2 >= twin2) {
                sdzr = dzr;
                sdzi = dzi;
                sk = k;
                save_j2 = j;
                twin2 += twin2;
            }
        }
        if (c->bla && dzr * dzr + dzi * dzi < c->bla->rmax2) {
            int skip = try_bla_d(c->bla, c->bla_minlevel, Z, k - 1, &dzr, &dzi, &ddr, &ddi, dcr, dci, ede, c->esc2, mx_ref_it);
            if (skip > 0) {
                k += skip;
                j += skip;
                st->bla_applies++;
                st->bla_skipped += skip;
                zsr = 1e30;
                zsi = 1e30;
                save_j = j;
                period_win = 1;
                conf_P = 0;
                conf_giveup = 0;
                continue;
            }
        }
        double tmp;
        if (ede) {
            /* dd' = 2 (dd X + dz (D + dd)), before dz advances */
            if (k == 0) {
                tmp = (dzr * (1 + ddr) - dzi * (1 + ddi)) * 2;
                ddi = (dzr * (1 + ddi) + dzi * (1 + ddr)) * 2;
            } else {
                const double xr = Z[2 * (k - 1)], xi = Z[2 * (k - 1) + 1];
                tmp = (ddr * xr - ddi * xi + dzr * (c->dfr[k - 1] + ddr) - dzi * (c->dfi[k - 1] + ddi)) * 2;
                ddi = (ddr * xi + ddi * xr + dzr * (c->dfi[k - 1] + ddi) + dzi * (c->dfr[k - 1] + ddr)) * 2;
            }
            ddr = tmp;
            dr = ddr + c->dfr[k];
            di = ddi + c->dfi[k];
        }
        double dzr2 = dzr * dzr, dzi2 = dzi * dzi;
        if (k == 0) {
            tmp = dzr2 - dzi2 + dcr;
            dzi = dzr * dzi * 2 + dci;
        } else {
            double xr = Z[2 * (k - 1)], xi = Z[2 * (k - 1) + 1];
            tmp = (dzr * xr - dzi * xi) * 2 + dzr2 - dzi2 + dcr;
            dzi = (dzr * xi + dzi * xr) * 2 + dzr * dzi * 2 + dci;
        }
        dzr = tmp;
        double zr = dzr + Z[2 * k], zi = dzi + Z[2 * k + 1];
        double zrad = zr * zr + zi * zi;
        ++k;
        st->steps++;
        if (zrad > c->esc2) {
            *n_out = j + 1;
            if (ede) return sqrt(zrad) / c->dxf * log(zrad) / sqrt(dr * dr + di * di);
            return j + 1 - log(log(zrad) / 2 / LG) / LG;
        }
        if (detect_interior) {
            if (conf_P > 0) {
                conf_D2 *= 4 * zrad;
                if (conf_D2 > 1e18) conf_D2 = 1e18;
                if (j >= conf_next) {
                    double pr = zr - conf_zr, pi = zi - conf_zi;
                    int ret = (pr * pr + pi * pi < c->interior_eps2 * zrad);
                    if (ret && conf_D2 < 1.0) {
                        if (++conf_count >= c->interior_confirm) {
                            *n_out = -2;
                            return -2.0;
                        }
                        conf_zr = zr;
                        conf_zi = zi;
                        conf_D2 = 1;
                        conf_next = j + conf_P;
                    } else {
                        conf_P = 0;
                        ++conf_giveup;
                    }
                }
            } else if ((j & 15) == 0 && conf_giveup < 3) {
                double pr = zr - zsr, pi = zi - zsi;
                if (pr * pr + pi * pi < c->interior_eps2 * zrad) {
                    conf_P = j - save_j;
                    if (conf_P < 1) conf_P = 1;
                    conf_zr = zr;
                    conf_zi = zi;
                    conf_D2 = 1;
                    conf_next = j + conf_P;
                    conf_count = 0;
                }
                if (j - save_j >= period_win) {
                    zsr = zr;
                    zsi = zi;
                    save_j = j;
                    period_win += period_win;
                }
            }
        }
        /* The engine reads _z_m3[j] with j the pixel's iteration, which can run past
         * the entries this reference filled; the index is clamped to the last one. */
        const int jm = j < mx_ref_it ? j : mx_ref_it;
        /* The end-of-reference test is k >= mx_ref_it where the engine has k ==: a
         * series skip that ends at the reference's escape iteration starts the loop at
         * k == mx_ref_it, and one step later the engine would read past the orbit. */
        if (zrad < dzr * dzr + dzi * dzi || k >= mx_ref_it) {
            if ((dzr * dzr + dzi * dzi) / zrad > 10000000) {
                *glitch_p = zrad / c->z_m3[jm];
                *n_out = -3;
                return -10.0;
            }
            dzr = zr;
            dzi = zi;
            if (ede) {
                ddr = dr - 1;
                ddi = di;
            }
            k = 0;
            st->rebases++;
        }
        ++j;
    }
    *n_out = -1;
    return -2.0;
}

/* Reference orbit of the double path with its series coefficients (createRef +
 * calCoefficient, MP.cpp:4014-4187, 4189-4305).  sa keeps `order` from the previous
 * reference, as the engine's _SA_order does. */
static void orbit_build_sa(mxo_orbit* o, mxo_sa* sa, mpfr_srcptr cre, mpfr_srcptr cim, mpfr_prec_t prec, int mxit, double escape_radius, mxo_c delta, int ede) {
    orbit_build(o, cre, cim, prec, mxit, escape_radius, ede, 0, 0);
    for (int i = 0; i < MXO_SA_N; ++i) {
        sa->A[i].r = sa->A[i].i = 0.0;
        sa->B[i].r = sa->B[i].i = 0.0;
    }
    sa->A[0] = delta;
    sa->delta = delta;
    sa->flag = 1;
    sa->it = 0;
    sa->ede = ede;
    /* calCoefficient(i) runs the SA step after a non-escaping Z[i] */
    const int last = o->escaped ? o->ref_it - 1 : o->ref_it;
    for (int i = 1; i <= last && sa->flag; ++i) {
        mxo_c z = {o->zf[2 * (i - 1)], o->zf[2 * (i - 1) + 1]};
        mxo_c d = {ede ? o->dfr[i - 1] : 0.0, ede ? o->dfi[i - 1] : 0.0};
        sa_step(sa, z, d, i);
    }
}

typedef struct {
    const char* cre;
    const char* cim;
    const char* scale;
    int W, H, mxit;
    float escape_radius;
    int prec_bits;
    int use_bla, bla_minskip, interior, threads;
    double* val;
    int32_t* n_esc;
    float* out;
    int passes;          /* reference passes used */
    int escalate;        /* 1: the frame belongs to the floatexp path (MP.cpp:1336-1355) */
    int sa_it0, sa_order0;
    long long steps, bla_applies, bla_skipped, rebases;
    int ede;             /* exterior distance estimate output */
    /* output cells per pixel side (0 or 1: none) and adaptive supersampling (ss): the
     * (H sub) x (W sub) cells of getIndex (MP.cpp:4334-4340), written where a sample is
     * computed; ss_pixels: pixels refined; ss_passes: reference passes over their
     * sub-samples (MP.cpp:1886-2005) */
    int sub, ss;
    float* cells;
    double* cell_val;
    int32_t* cell_nesc;
    int ss_pixels, ss_passes;
} mxo_dframe;

MXO_EXPORT int mxo_compute_dbl(mxo_dframe* f) {
    const int W = f->W, H = f->H, mxit = f->mxit;
    const int pb = mpf_bits(f->prec_bits);
    const double esc2 = (double)(f->escape_radius * f->escape_radius);
    const double escR = (double)f->escape_radius;
    mxo_grid g;
    grid_init(&g, f->cre, f->cim, f->scale, W, H, 0, 0, pb, 1);
    mpfr_t t;
    mpfr_init2(t, pb);
    mxo_c delta;
    mpfr_mul_ui(t, g.dx, W, MPFR_RNDN);
    mpfr_div_ui(t, t, 2, MPFR_RNDN);
    delta.r = mpfr_to_ld_trunc(t);
    mpfr_mul_ui(t, g.dy, H, MPFR_RNDN);
    mpfr_div_ui(t, t, 2, MPFR_RNDN);
    delta.i = mpfr_to_ld_trunc(t);
    const double dxf = fabs(mpfr_to_ld_trunc(g.dx)), dyf = fabs(mpfr_to_ld_trunc(g.dy));
    /* per-pixel dc relative to a reference pixel, truncated from full precision */
    double* dcol = (double*)malloc(sizeof(double) * W);
    double* drow = (double*)malloc(sizeof(double) * H);
    int ref_i = H / 2, ref_j = W / 2;
    mpfr_t rre, rim;
    mpfr_inits2(pb, rre, rim, (mpfr_ptr)0);
    grid_pixel_c(&g, ref_i, ref_j, 0, 0, 1, rre, rim);
    mxo_orbit o;
    mxo_sa sa;
    memset(&sa, 0, sizeof sa);
    const int ede = f->ede;
    orbit_build_sa(&o, &sa, rre, rim, pb, mxit, escR, delta, ede);
    f->sa_it0 = sa.it;
    f->sa_order0 = sa.order;
    f->escalate = 0;
    {
        mpfr_t sc;
        mpfr_init2(sc, pb);
        mpfr_set_str(sc, f->scale, 10, MPFR_RNDN);
        if (mpfr_cmp_d(sc, 1e150) > 0) {
            double refmin2 = 1e300;
            int uf = 0;
            for (int q = 1; q <= o.ref_it; ++q) {
                double m2 = o.z_m3[q] * 1e6;
                if (m2 <= 0.0) {
                    uf = 1;
                    break;
                }
                if (m2 < refmin2) refmin2 = m2;
            }
            double dcmax = 2.0 / mpfr_get_d(sc, MPFR_RNDZ);
            if (uf || sqrt(refmin2) * dcmax < 1e-300) f->escalate = 1;
        }
        mpfr_clear(sc);
    }
    if (f->escalate) {
        /* the frame belongs to the floatexp path (mxo_compute_fe) */
        f->passes = 0;
        free(dcol);
        free(drow);
        orbit_free(&o);
        mpfr_clears(rre, rim, t, (mpfr_ptr)0);
        grid_clear(&g);
        return 1;
    }
    int* idx = (int*)malloc(sizeof(int) * (size_t)W * H);
    int* rest = (int*)malloc(sizeof(int) * (size_t)W * H);
    double* gp = (double*)malloc(sizeof(double) * (size_t)W * H);
    for (int p = 0; p < W * H; ++p) {
        f->val[p] = -10.0;
        f->n_esc[p] = INT32_MIN;
    }
    mxo_bla bla;
    memset(&bla, 0, sizeof bla);
    mxo_dctx ctx;
    memset(&ctx, 0, sizeof ctx);
    ctx.esc2 = esc2;
    ctx.bla_minlevel = bla_minlevel_from_skip(f->bla_minskip > 0 ? f->bla_minskip : 8);
    ctx.use_interior = f->interior;
    ctx.interior_eps2 = 3e-9 * 3e-9;
    ctx.interior_confirm = 4;
    ctx.int_rep = 1;
    ctx.ede = ede;
    ctx.dxf = mpfr_to_ld_trunc(g.dx);
    long long steps = 0, ap = 0, sk = 0, rb = 0;

    /* set-up for the current reference: BLA, dc tables */
#define MXO_SETUP_REF()                                                                                       \
    do {                                                                                                      \
        if (f->use_bla) {                                                                                     \
            bla_free(&bla);                                                                                   \
            double mxx = (ref_j > (W - 1) - ref_j ? ref_j : (W - 1) - ref_j) + 0.5;                           \
            double mxy = (ref_i > (H - 1) - ref_i ? ref_i : (H - 1) - ref_i) + 0.5;                           \
            double dcmax = sqrt(dxf * dxf * mxx * mxx + dyf * dyf * mxy * mxy);                               \
            bla_build(&bla, o.zf, o.dfr, o.dfi, o.ref_it, dcmax, ldexp(1.0, -53), ede);                   \
        }                                                                                                     \
        ctx.zf = o.zf;                                                                                        \
        ctx.z_m3 = o.z_m3;                                                                                    \
        ctx.dfr = o.dfr;                                                                                      \
        ctx.dfi = o.dfi;                                                                                      \
        ctx.mx_ref_it = o.ref_it;                                                                             \
        ctx.bla = f->use_bla ? &bla : NULL;                                                                   \
        ctx.sa_it = sa.it;                                                                                    \
        for (int jj = 0; jj < W; ++jj) {                                                                      \
            mpfr_mul_si(t, g.dx, jj - ref_j, MPFR_RNDN);                                                      \
            dcol[jj] = mpfr_to_ld_trunc(t);                                                                   \
        }                                                                                                     \
        for (int ii = 0; ii < H; ++ii) {                                                                      \
            mpfr_mul_si(t, g.dy, ii - ref_i, MPFR_RNDN);                                                      \
            drow[ii] = mpfr_to_ld_trunc(t);                                                                   \
        }                                                                                                     \
    } while (0)

    /* value of a (non-virtual) reference pixel (MP.cpp:4145-4160, 4179-4185) */
#define MXO_REF_PIXEL()                                                                                       \
    do {                                                                                                      \
        int ne_;                                                                                              \
        f->val[ref_i * W + ref_j] = ref_pixel_value(&o, mxit, ede ? 2 : 0, esc2, g.dx, &ne_);                \
        f->n_esc[ref_i * W + ref_j] = ne_;                                                                    \
    } while (0)

    MXO_SETUP_REF();
    MXO_REF_PIXEL();
    const int ref0 = ref_i * W + ref_j;

    /* one pass over idx[0..n): resolved pixels written; unresolved kept in rest */
#define MXO_PASS(nin, nout)                                                                                   \
    do {                                                                                                      \
        _Pragma("omp parallel for schedule(dynamic, 16) num_threads(f->threads) reduction(+ : steps, ap, sk, rb)") \
        for (int q = 0; q < (nin); ++q) {                                                                     \
            int p = idx[q];                                                                                   \
            int i = p / W, jx = p % W;                                                                        \
            mxo_c dc = {dcol[jx], drow[i]};                                                                   \
            mxo_c dz = sa_seed(&sa, dc);                                                                      \
            mxo_c dd0 = {0.0, 0.0};                                                                           \
            if (ede) dd0 = sa_horner(sa.B, sa.order, sa.delta, dc);                                           \
            mxo_stats st = {0, 0, 0, 0, 0, 0, 0};                                                                \
            int ne;                                                                                           \
            double gpp;                                                                                       \
            double v = pixel_double(&ctx, dc, dz, dd0, mxit, &ne, &gpp, &st);                                 \
            gp[p] = gpp;                                                                                      \
            if (ne != -3) {                                                                                   \
                f->val[p] = v;                                                                                \
                f->n_esc[p] = ne;                                                                             \
            }                                                                                                 \
            steps += st.steps;                                                                                \
            ap += st.bla_applies;                                                                             \
            sk += st.bla_skipped;                                                                             \
            rb += st.rebases;                                                                                 \
        }                                                                                                     \
        (nout) = 0;                                                                                           \
        for (int q = 0; q < (nin); ++q)                                                                       \
            if (gp[idx[q]] >= 0) rest[(nout)++] = idx[q];                                                     \
    } while (0)

    if (f->interior) {
        int n = 0, nr;
        for (int i = 0; i < H; i += 8)
            for (int jx = 0; jx < W; jx += 8) idx[n++] = i * W + jx;
        long long s0 = steps, a0 = ap, k0 = sk, r0 = rb;
        MXO_PASS(n, nr);
        steps = s0, ap = a0, sk = k0, rb = r0;
        int nint = 0;
        for (int q = 0; q < n; ++q)
            if (f->n_esc[idx[q]] == -1 || f->n_esc[idx[q]] == -2) ++nint;
        /* the probe's own centre value (a pixel reference) is restored afterwards */
        if (nint == 0) ctx.use_interior = 0;
        for (int q = 0; q < n; ++q) {
            f->val[idx[q]] = -10.0;
            f->n_esc[idx[q]] = INT32_MIN;
        }
        MXO_REF_PIXEL();
    }
    int n = 0;
    for (int p = 0; p < W * H; ++p)
        if (p != ref0) idx[n++] = p;
    int passes = 0;
    while (n > 0) {
        int nr;
        MXO_PASS(n, nr);
        ++passes;
        if (nr == 0) break;
        /* glitch-guided next reference: the unresolved pixel with the smallest
         * glitch metric, first in row-major order (MP.cpp:2587-2594, 4046-4056) */
        int best = -1;
        double mn = 1e100;
        for (int q = 0; q < nr; ++q)
            if (gp[rest[q]] < mn) {
                mn = gp[rest[q]];
                best = rest[q];
            }
        if (best < 0) best = rest[0];
        ref_i = best / W;
        ref_j = best % W;
        orbit_free(&o);
        grid_pixel_c(&g, ref_i, ref_j, 0, 0, 1, rre, rim);
        orbit_build_sa(&o, &sa, rre, rim, pb, mxit, escR, delta, ede);
        MXO_REF_PIXEL();
        n = 0;
        for (int q = 0; q < nr; ++q)
            if (rest[q] != best) idx[n++] = rest[q];
        /* with sub-samples to follow, the last reference's BLA table is built even when
         * no base pixel is left for it (MP.cpp:1750-1756) */
        if (n == 0 && !(f->sub > 1 && f->ss)) break;
        MXO_SETUP_REF();
        if (n == 0) break;
    }
    f->passes = passes;
    for (int p = 0; p < W * H; ++p) f->out[p] = (f->n_esc[p] == INT32_MIN) ? -10.f : (f->n_esc[p] < 0) ? -2.f : (float)f->val[p];
    f->ss_pixels = 0;
    f->ss_passes = 0;
    if (f->sub > 1) {
        const int S = f->sub;
        for (int i = 0; i < H; ++i)
            for (int j = 0; j < W; ++j) {
                const size_t cc = ss_cell(W, S, i, j, 0, 0);
                f->cells[cc] = f->out[i * W + j];
                f->cell_val[cc] = f->val[i * W + j];
                f->cell_nesc[cc] = f->n_esc[i * W + j];
            }
    }
    if (f->sub > 1 && f->ss) {
        /* adaptive supersampling on the double path (MP.cpp:1789-2005): the flagged
         * pixels' other cells in the engine's std::set order (i, j, yi, xi), passes
         * against the current reference -- the first with its BLA table, the later ones
         * after a glitch re-reference with BLA off -- until every sub-sample resolves */
        const int S = f->sub, h = S / 2, per = S * S - 1;
        unsigned char* flags = (unsigned char*)malloc((size_t)W * H);
        f->ss_pixels = ss_flag_pixels(f->out, W, H, mxit, ede, flags);
        const int ns = f->ss_pixels * per;
        int* sc = (int*)malloc(sizeof(int) * (size_t)(ns + 1));   /* cell column */
        int* sr = (int*)malloc(sizeof(int) * (size_t)(ns + 1));   /* cell row */
        int* pend = (int*)malloc(sizeof(int) * (size_t)(ns + 1));
        int* left = (int*)malloc(sizeof(int) * (size_t)(ns + 1));
        double* sgp = (double*)malloc(sizeof(double) * (size_t)(ns + 1));
        double* ccol = (double*)malloc(sizeof(double) * (size_t)W * S);
        double* crow = (double*)malloc(sizeof(double) * (size_t)H * S);
        int m = 0;
        for (int p = 0; p < W * H; ++p) {
            if (!flags[p]) continue;
            const int i = p / W, j = p % W;
            for (int yi = -h; yi <= h; ++yi)
                for (int xi = -h; xi <= h; ++xi) {
                    if (!xi && !yi) continue;
                    sc[m] = j * S + h + xi;
                    sr[m] = i * S + h + yi;
                    pend[m] = m;
                    ++m;
                }
        }
        /* the current reference's cell, and dc = trunc53(d * (cell - reference cell) / sub)
         * per cell column and row (stepParallel, MP.cpp:2300-2312) */
        int rc = ref_j * S + h, rr = ref_i * S + h;
#define MXO_CELL_DC()                                                                                         \
    do {                                                                                                      \
        for (int q = 0; q < W * S; ++q) {                                                                     \
            mpfr_mul_si(t, g.dx, q - rc, MPFR_RNDN);                                                          \
            mpfr_div_ui(t, t, S, MPFR_RNDN);                                                                  \
            ccol[q] = mpfr_to_ld_trunc(t);                                                                    \
        }                                                                                                     \
        for (int q = 0; q < H * S; ++q) {                                                                     \
            mpfr_mul_si(t, g.dy, q - rr, MPFR_RNDN);                                                          \
            mpfr_div_ui(t, t, S, MPFR_RNDN);                                                                  \
            crow[q] = mpfr_to_ld_trunc(t);                                                                    \
        }                                                                                                     \
    } while (0)
        MXO_CELL_DC();
        int np = m;
        while (np > 0) {
            _Pragma("omp parallel for schedule(dynamic, 16) num_threads(f->threads) reduction(+ : steps, ap, sk, rb)")
            for (int q = 0; q < np; ++q) {
                const int k = pend[q];
                mxo_c dc = {ccol[sc[k]], crow[sr[k]]};
                mxo_c dz = sa_seed(&sa, dc);
                mxo_c dd0 = {0.0, 0.0};
                if (ede) dd0 = sa_horner(sa.B, sa.order, sa.delta, dc);
                mxo_stats st = {0, 0, 0, 0, 0, 0, 0};
                int ne;
                double gpp;
                const double v = pixel_double(&ctx, dc, dz, dd0, mxit, &ne, &gpp, &st);
                sgp[k] = gpp;
                if (ne != -3) {
                    const size_t cc = (size_t)sr[k] * W * S + sc[k];
                    f->cell_val[cc] = v;
                    f->cell_nesc[cc] = ne;
                    f->cells[cc] = ne < 0 ? -2.f : (float)v;
                }
                steps += st.steps;
                ap += st.bla_applies;
                sk += st.bla_skipped;
                rb += st.rebases;
            }
            ++f->ss_passes;
            int nl = 0;
            for (int q = 0; q < np; ++q)
                if (sgp[pend[q]] >= 0) left[nl++] = pend[q];
            if (nl == 0) break;
            /* next reference: the pending sub-sample with the smallest glitch metric,
             * first in set order among equals (MP.cpp:2587-2594, 4046-4056) */
            int best = left[0];
            double mn = 1e100;
            for (int q = 0; q < nl; ++q)
                if (sgp[left[q]] < mn) {
                    mn = sgp[left[q]];
                    best = left[q];
                }
            const int bi = sr[best] / S, bj = sc[best] / S, bu = sr[best] % S - h, bv = sc[best] % S - h;
            orbit_free(&o);
            grid_ref_c(&g, bi, bj, bu, bv, S, rre, rim);
            orbit_build_sa(&o, &sa, rre, rim, pb, mxit, escR, delta, ede);
            {
                int ne_;
                const size_t cc = (size_t)sr[best] * W * S + sc[best];
                const double v = ref_pixel_value(&o, mxit, ede ? 2 : 0, esc2, g.dx, &ne_);
                f->cell_val[cc] = v;
                f->cell_nesc[cc] = ne_;
                f->cells[cc] = ne_ < 0 ? -2.f : (float)v;
            }
            ctx.zf = o.zf;
            ctx.z_m3 = o.z_m3;
            ctx.dfr = o.dfr;
            ctx.dfi = o.dfi;
            ctx.mx_ref_it = o.ref_it;
            ctx.sa_it = sa.it;
            ctx.bla = NULL; /* the base table belongs to the old reference (MP.cpp:2003) */
            rc = sc[best];
            rr = sr[best];
            np = 0;
            for (int q = 0; q < nl; ++q)
                if (left[q] != best) pend[np++] = left[q];
            if (np == 0) break;
            MXO_CELL_DC();
        }
#undef MXO_CELL_DC
        free(flags);
        free(sc);
        free(sr);
        free(pend);
        free(left);
        free(sgp);
        free(ccol);
        free(crow);
    }
    f->steps = steps;
    f->bla_applies = ap;
    f->bla_skipped = sk;
    f->rebases = rb;
#undef MXO_SETUP_REF
#undef MXO_REF_PIXEL
#undef MXO_PASS
    free(idx);
    free(rest);
    free(gp);
    free(dcol);
    free(drow);
    bla_free(&bla);
    orbit_free(&o);
    mpfr_clears(rre, rim, t, (mpfr_ptr)0);
    grid_clear(&g);
    return 0;
}

/* Series coefficients of the reference at (cre, cim) for a frame half-extent delta:
 * A[2*x], A[2*x+1] = A_x (x < 30); returns the last accepted iteration (_SA_it). */
MXO_EXPORT int mxo_sa_coeffs(const char* cre, const char* cim, int prec_bits, int mxit, double escape_radius, double delta_r, double delta_i, double* A, int* order) {
    mpfr_t re, im;
    mpfr_inits2(prec_bits, re, im, (mpfr_ptr)0);
    mpfr_set_str(re, cre, 10, MPFR_RNDN);
    mpfr_set_str(im, cim, 10, MPFR_RNDN);
    mxo_orbit o;
    mxo_sa sa;
    memset(&sa, 0, sizeof sa);
    mxo_c delta = {delta_r, delta_i};
    orbit_build_sa(&o, &sa, re, im, prec_bits, mxit, escape_radius, delta, 0);
    for (int x = 0; x < MXO_SA_N; ++x) {
        A[2 * x] = sa.A[x].r;
        A[2 * x + 1] = sa.A[x].i;
    }
    *order = sa.order;
    orbit_free(&o);
    mpfr_clears(re, im, (mpfr_ptr)0);
    return sa.it;
}

/* The series seed (MP.cpp:2319-2323) for n pixel offsets. */
MXO_EXPORT void mxo_sa_seed(const double* A, int order, double delta_r, double delta_i, int n, const double* dcr, const double* dci, double* out) {
    mxo_sa sa;
    memset(&sa, 0, sizeof sa);
    for (int x = 0; x < MXO_SA_N; ++x) {
        sa.A[x].r = A[2 * x];
        sa.A[x].i = A[2 * x + 1];
    }
    sa.order = order;
    sa.delta.r = delta_r;
    sa.delta.i = delta_i;
    for (int q = 0; q < n; ++q) {
        mxo_c dc = {dcr[q], dci[q]};
        mxo_c dz = sa_seed(&sa, dc);
        out[2 * q] = dz.r;
        out[2 * q + 1] = dz.i;
    }
}

/* Complex division as the double path's series seed does it (exposed for tests). */
MXO_EXPORT void mxo_cdiv(int n, const double* a, const double* b, double* out) {
    for (int q = 0; q < n; ++q) {
        mxo_c x = {a[2 * q], a[2 * q + 1]}, y = {b[2 * q], b[2 * q + 1]};
        mxo_c r = c_div(x, y);
        out[2 * q] = r.r;
        out[2 * q + 1] = r.i;
    }
}

/* The double loop over caller-supplied orbit/BLA/series arrays (the same arrays the
 * GPU consumed), so a parity test isolates the kernel.  gp[q] >= 0 marks an
 * unresolved sample (nesc[q] = INT32_MIN). */
MXO_EXPORT void mxo_pixels_dbl(const mxo_pixels_in* in, const double* z_m3, const double* dfr, const double* dfi, const double* der, const double* A, const double* B, int sa_it, int sa_order, double delta_r, double delta_i, double dxf, int n, const double* dcr, const double* dci, double* val, int32_t* nesc, double* gp, long long* stats4) {
    mxo_bla b;
    memset(&b, 0, sizeof b);
    if (in->nlev > 0) {
        b.nlev = in->nlev;
        for (int p = 0; p < in->nlev; ++p) {
            b.off[p] = in->lev_off[p];
            b.size[p] = in->lev_size[p];
        }
        b.r2 = (double*)in->r2;
        b.l = (int64_t*)in->l;
        b.coef = (double*)in->coef;
        b.der = (double*)der;
        b.rmax2 = in->rmax2;
    }
    mxo_sa sa;
    memset(&sa, 0, sizeof sa);
    for (int x = 0; x < MXO_SA_N; ++x) {
        sa.A[x].r = A[2 * x];
        sa.A[x].i = A[2 * x + 1];
        if (B) {
            sa.B[x].r = B[2 * x];
            sa.B[x].i = B[2 * x + 1];
        }
    }
    sa.order = sa_order;
    sa.it = sa_it;
    sa.delta.r = delta_r;
    sa.delta.i = delta_i;
    mxo_dctx ctx;
    memset(&ctx, 0, sizeof ctx);
    ctx.zf = in->zf;
    ctx.z_m3 = z_m3;
    ctx.dfr = dfr;
    ctx.dfi = dfi;
    ctx.mx_ref_it = in->mx_ref_it;
    ctx.bla = in->nlev > 0 ? &b : NULL;
    ctx.bla_minlevel = in->bla_minlevel;
    ctx.esc2 = in->esc2;
    ctx.use_interior = in->use_interior;
    ctx.interior_eps2 = 3e-9 * 3e-9;
    ctx.interior_confirm = 4;
    ctx.int_rep = 1;
    ctx.sa_it = sa_it;
    ctx.ede = in->ede;
    ctx.dxf = dxf;
    long long s0 = 0, s1 = 0, s2 = 0, s3 = 0;
#pragma omp parallel for schedule(dynamic, 16) num_threads(in->threads) reduction(+ : s0, s1, s2, s3)
    for (int q = 0; q < n; ++q) {
        mxo_c dc = {dcr[q], dci[q]};
        mxo_c dz = sa_seed(&sa, dc);
        mxo_c dd0 = {0.0, 0.0};
        if (in->ede) dd0 = sa_horner(sa.B, sa.order, sa.delta, dc);
        mxo_stats st = {0, 0, 0, 0, 0, 0, 0};
        int ne;
        double g;
        double v = pixel_double(&ctx, dc, dz, dd0, in->mxit, &ne, &g, &st);
        gp[q] = g;
        if (ne == -3) {
            nesc[q] = INT32_MIN;
            val[q] = -10.0;
        } else {
            nesc[q] = ne;
            val[q] = v;
        }
        s0 += st.steps;
        s1 += st.bla_applies;
        s2 += st.bla_skipped;
        s3 += st.rebases;
    }
    stats4[0] = s0;
    stats4[1] = s1;
    stats4[2] = s2;
    stats4[3] = s3;
}

/* First-reference set-up of the double frame (the part of mxo_compute_dbl before the
 * passes): orbit shadows and |X|^2/1e6 (capacity mxit+1), series coefficients,
 * delta and the double dc tables.  Returns ref_it. */
MXO_EXPORT int mxo_dbl_setup(const char* cre, const char* cim, const char* scale, int W, int H, int mxit, float escape_radius, int prec_bits, double* zf, double* z_m3, double* A, double* delta, int* sa_it, int* sa_order, double* dcol, double* drow, int ede, double* dfr, double* dfi, double* B) {
    prec_bits = mpf_bits(prec_bits);
    mxo_grid g;
    grid_init(&g, cre, cim, scale, W, H, 0, 0, prec_bits, 1);
    mpfr_t t, rre, rim;
    mpfr_inits2(prec_bits, t, rre, rim, (mpfr_ptr)0);
    mxo_c d;
    mpfr_mul_ui(t, g.dx, W, MPFR_RNDN);
    mpfr_div_ui(t, t, 2, MPFR_RNDN);
    d.r = mpfr_to_ld_trunc(t);
    mpfr_mul_ui(t, g.dy, H, MPFR_RNDN);
    mpfr_div_ui(t, t, 2, MPFR_RNDN);
    d.i = mpfr_to_ld_trunc(t);
    const int ref_i = H / 2, ref_j = W / 2;
    grid_pixel_c(&g, ref_i, ref_j, 0, 0, 1, rre, rim);
    mxo_orbit o;
    mxo_sa sa;
    memset(&sa, 0, sizeof sa);
    orbit_build_sa(&o, &sa, rre, rim, prec_bits, mxit, (double)escape_radius, d, ede);
    if (ede) {
        memcpy(dfr, o.dfr, sizeof(double) * (o.ref_it + 1));
        memcpy(dfi, o.dfi, sizeof(double) * (o.ref_it + 1));
    }
    memcpy(zf, o.zf, sizeof(double) * 2 * (o.ref_it + 1));
    memcpy(z_m3, o.z_m3, sizeof(double) * (o.ref_it + 1));
    for (int x = 0; x < MXO_SA_N; ++x) {
        A[2 * x] = sa.A[x].r;
        A[2 * x + 1] = sa.A[x].i;
        B[2 * x] = sa.B[x].r;
        B[2 * x + 1] = sa.B[x].i;
    }
    delta[0] = d.r;
    delta[1] = d.i;
    *sa_it = sa.it;
    *sa_order = sa.order;
    for (int j = 0; j < W; ++j) {
        mpfr_mul_si(t, g.dx, j - ref_j, MPFR_RNDN);
        dcol[j] = mpfr_to_ld_trunc(t);
    }
    for (int i = 0; i < H; ++i) {
        mpfr_mul_si(t, g.dy, i - ref_i, MPFR_RNDN);
        drow[i] = mpfr_to_ld_trunc(t);
    }
    const int ri = o.ref_it;
    orbit_free(&o);
    mpfr_clears(t, rre, rim, (mpfr_ptr)0);
    grid_clear(&g);
    return ri;
}

/* ========================================================================= */
/* colour pass (tools/render.cpp:21-35, 69-91, 252-302; engine/interpolate.cpp */
/* 7-65), float arithmetic as there                                           */
/* ========================================================================= */

#define MXO_COLP 2048

/* mono_cubic_interpolate (interpolate.cpp:7-65) */
static void mxo_mono_cubic(const float* xs, const float* ys, int n, float* out, int m) {
    float dxs[16], ms[16], c1[16], c2[16], c3[16];
    for (int i = 0; i < n - 1; ++i) {
        dxs[i] = xs[i + 1] - xs[i];
        ms[i] = (ys[i + 1] - ys[i]) / dxs[i];
    }
    c1[0] = ms[0];
    for (int i = 0; i < n - 2; ++i) {
        const float a = ms[i], b = ms[i + 1];
        if (a * b <= 0) {
            c1[i + 1] = 0;
        } else {
            const float common = dxs[i] + dxs[i + 1];
            c1[i + 1] = 3 * common / ((common + dxs[i + 1]) / a + (common + dxs[i]) / b);
        }
    }
    c1[n - 1] = ms[n - 2];
    for (int i = 0; i < n - 1; ++i) {
        const float inv = 1 / dxs[i], common = c1[i] + c1[i + 1] - ms[i] - ms[i];
        c2[i] = (ms[i] - c1[i] - common) * inv;
        c3[i] = common * inv * inv;
    }
    for (int k = 0; k < m; ++k) {
        const float x = 1.f * k / m;
        if (x == xs[n - 1]) {
            out[k] = ys[n - 1];
            continue;
        }
        int low = 0, high = n - 2, mid, found = -1;
        while (low <= high) {
            mid = (low + high) / 2;
            if (xs[mid] < x)
                low = mid + 1;
            else if (xs[mid] > x)
                high = mid - 1;
            else {
                found = mid;
                break;
            }
        }
        if (found >= 0) {
            out[k] = ys[found];
            continue;
        }
        const int i = high > 0 ? high : 0;
        const float d = x - xs[i], d2 = d * d;
        out[k] = ys[i] + c1[i] * d + c2[i] * d2 + c3[i] * d * d2;
    }
}

/* colorMapInit (render.cpp:25-30) */
MXO_EXPORT void mxo_palette(float* pal) {
    static const float pos[6] = {0.0f, 0.16f, 0.42f, 0.6425f, 0.8575f, 1.0f};
    static const float col[3][6] = {{0, 32, 237, 255, 0, 0}, {70, 107, 255, 170, 2, 70}, {100, 203, 255, 0, 0, 100}};
    for (int c = 0; c < 3; ++c) mxo_mono_cubic(pos, col[c], 6, pal + c * MXO_COLP, MXO_COLP);
}

/* getColor's palette index (render.cpp:80-81); a product out of int range converts as
 * x86's cvttss2si does (INT_MIN) */
static int mxo_palette_index(float f) {
    const float v = f * (float)MXO_COLP;
    int x = (v < 2147483648.0f && v > -2147483904.0f) ? (int)v : INT_MIN;
    x %= MXO_COLP;
    if (x < 0) x += MXO_COLP;
    return x;
}

/* the colour loop of render.cpp:252-302 without the DE overlay, normal map and relief */
MXO_EXPORT void mxo_colorize(const float* iter, int W, int H, int ss, int ede, float density, unsigned char* rgb) {
    float pal[3 * MXO_COLP];
    mxo_palette(pal);
    const size_t Ws = (size_t)W * ss;
    for (int i = 0; i < H; ++i)
        for (int j = 0; j < W; ++j) {
            float lr = 0, lg = 0, lb = 0;
            for (int a = 0; a < ss; ++a)
                for (int b = 0; b < ss; ++b) {
                    const float it = iter[((size_t)i * ss + a) * Ws + (size_t)j * ss + b];
                    float r = 0, g = 0, bb = 0;
                    if (!(it < 0)) {
                        const float f = ede ? tanhf(it * density / 3600.0f * 5.0f)
                                            : powf(logf(it + 2), logf(it + 2) * logf(it + 2) * density / 3600.0f);
                        const int x = mxo_palette_index(f);
                        r = pal[x];
                        g = pal[MXO_COLP + x];
                        bb = pal[2 * MXO_COLP + x];
                    }
                    const double clr = powf(r / 255.f, 2.2f), clg = powf(g / 255.f, 2.2f), clb = powf(bb / 255.f, 2.2f);
                    lr += clr * 1.0f;
                    lg += clg * 1.0f;
                    lb += clb * 1.0f;
                }
            const int n = ss * ss;
            unsigned char* p = rgb + ((size_t)i * W + j) * 3;
            const float gr = lr / n, gg = lg / n, gb = lb / n;
            p[0] = (unsigned char)(255.f * powf(gr < 0 ? 0 : gr, 1.f / 2.2f) + 0.5f);
            p[1] = (unsigned char)(255.f * powf(gg < 0 ? 0 : gg, 1.f / 2.2f) + 0.5f);
            p[2] = (unsigned char)(255.f * powf(gb < 0 ? 0 : gb, 1.f / 2.2f) + 0.5f);
        }
}
