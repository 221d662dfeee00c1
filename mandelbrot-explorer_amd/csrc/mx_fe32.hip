// Mixed-precision form of the per-pixel perturbation kernel for gfx950.
//
// Same algorithm as the plain form (Mandel::pixelRescaled,
// src/engine/mandel_perturbation.cpp:3287-3618: delta iteration, BLA skips,
// rescaling, Zhuoran rebasing, carried derivative for the distance estimate) with two
// representation changes chosen for the GPU:
//   * the delta is dz = w * 2^k: w is an FP32 complex, k an int32.  The scale is an
//     exact power of two, so every rescale is an exponent shift (ldexp) instead of the
//     engine's sqrt + three divisions, and d = dc / 2^k is a shift of dc's mantissa;
//   * the delta step, BLA apply and derivative run in FP32 on an FP32 copy of the
//     orbit and FP32-floatexp BLA coefficients (mantissa + int32 exponent).
// FP64 stays where it decides or produces results: the BLA validity radii and
// their test, the rebase test |Z + dz| < |dz| and the rebased value (from the FP64
// FloatExp orbit), and the final smooth-iteration / distance logarithm.
// Results are not bit-identical to the engine; their error against the MPFR oracle
// is measured (fraction of samples with identical escape count, max smooth error).
//
// Lane scheduling is the plain form's: persistent grid, per-XCD atomic work queues
// with wave-aggregated refill, 8x8 sample tiles.  Each trip runs one full step
// (every rare event handled) and then a run of lean steps that carry only the common
// case and stop, before committing, at the first step that needs the full path: a
// wave executes a branch whenever any of its 64 lanes takes it, so per-lane events
// that are individually rare (a BLA-aligned index every 8 steps, the periodicity
// check every 16) would otherwise put their code on nearly every wave step.
//
// The step arithmetic is written with explicit fused multiply-adds and no other FP
// contraction, so the full and the lean step produce identical bits.
#include <climits>

#include "mx_dev.h"

#pragma clang fp contract(off)

namespace {

// The scaled delta algorithm is written once for the delta's scalar type F: float
// is the mixed form; double gives the same power-of-two-scaled algorithm at FP64.
MX_HD float ldexp_t(float x, int e) { return ldexpf(x, e); }
MX_HD double ldexp_t(double x, int e) { return ldexp(x, e); }
MX_HD float frexp_t(float x, int* e) { return frexpf(x, e); }
MX_HD double frexp_t(double x, int* e) { return frexp(x, e); }
MX_HD float fmax_t(float a, float b) { return fmaxf(a, b); }
MX_HD double fmax_t(double a, double b) { return fmax(a, b); }
MX_HD float fmin_t(float a, float b) { return fminf(a, b); }
MX_HD double fmin_t(double a, double b) { return fmin(a, b); }
MX_HD float fabs_t(float a) { return fabsf(a); }
MX_HD double fabs_t(double a) { return fabs(a); }
MX_HD float fma_t(float a, float b, float c) { return fmaf(a, b, c); }
MX_HD double fma_t(double a, double b, double c) { return fma(a, b, c); }

template <typename F>
struct Vec2;
template <>
struct Vec2<float> {
    using T = float2;
};
template <>
struct Vec2<double> {
    using T = double2;
};

template <typename F>
struct P32 {
    using V = typename Vec2<F>::T;
    F wr, wi;    // dz = w * 2^k
    F dr, di;    // dc * 2^-k
    F s;         // 2^k (0 when it underflows)
    F dcr, dci;  // dc mantissas, exponents below
    int dcre, dcie;
    int k;
    int m, rm, iter;
    // derivative J = jw * 2^kj
    F jr, ji, jone; // jone = 2^-kj (the "+1")
    int kj;
    // interior detection
    F zsr, zsi, conf_D2, conf_zr, conf_zi;
    int save_iter, period_win, conf_P, conf_next, conf_count, conf_giveup;
    F swr, swi;
    int sk, sm, save_iter2, twin2r;
    int xv;
    V Xc, Xn;
};

struct C32 {
    unsigned long long steps, applies, skipped, rebases;
};

// minimum waves per SIMD requested from the register allocator (0: its own choice);
// the few spills this costs sit on rare paths and the added occupancy hides latency
#ifndef MX_SC64_WAVES
#define MX_SC64_WAVES 5
#endif
#ifndef MX_FE32_WAVES
#define MX_FE32_WAVES 5
#endif
#define MX_WAVES_ATTR(n) MX_WAVES_ATTR_(n)
#define MX_WAVES_ATTR_(n) __attribute__((amdgpu_waves_per_eu((n) > 0 ? (n) : 1)))

template <typename F>
MX_HD F pow2t(int e) {
    return ldexp_t((F)1, e); // 0 below the subnormal range, inf above the range
}

template <typename F>
MX_HD void set_scale(P32<F>& p) {
    // When dz is far below dc (a rebase onto a near-zero z) the delta is negligible
    // next to dc; lift k so that d = dc / 2^k stays inside the float range.
    int ec = INT_MIN;
    if (p.dcr != (F)0) ec = p.dcre;
    if (p.dci != (F)0) ec = max(ec, p.dcie);
    if (ec != INT_MIN && ec - p.k > 60) {
        const int sh = ec - p.k - 60;
        p.wr = ldexp_t(p.wr, -sh);
        p.wi = ldexp_t(p.wi, -sh);
        p.k += sh;
    }
    p.s = pow2t<F>(p.k);
    p.dr = ldexp_t(p.dcr, p.dcre - p.k);
    p.di = ldexp_t(p.dci, p.dcie - p.k);
}

// keep |w| ~ 1 by an exact power-of-two shift
template <typename F>
MX_HD void renorm_w(P32<F>& p) {
    const F a = fmax_t(fabs_t(p.wr), fabs_t(p.wi));
    if (a == (F)0) return;
    int e;
    (void)frexp_t(a, &e);
    p.wr = ldexp_t(p.wr, -e);
    p.wi = ldexp_t(p.wi, -e);
    p.k += e;
    set_scale(p);
}

template <typename F>
MX_HD void renorm_j(P32<F>& p) {
    const F a = fmax_t(fabs_t(p.jr), fabs_t(p.ji));
    if (a == (F)0) return;
    int e;
    (void)frexp_t(a, &e);
    p.jr = ldexp_t(p.jr, -e);
    p.ji = ldexp_t(p.ji, -e);
    p.kj += e;
    p.jone = pow2t<F>(-p.kj);
}

template <typename F>
MX_HD void init32(P32<F>& p, Fe dcr, Fe dci) {
    p.dcr = (F)dcr.m;
    p.dci = (F)dci.m;
    p.dcre = dcr.e;
    p.dcie = dci.e;
    // dz_1 = dc: choose k from the larger component
    const int e = max(dcr.m != 0.0 ? dcr.e : INT_MIN / 2, dci.m != 0.0 ? dci.e : INT_MIN / 2);
    if (dcr.m == 0.0 && dci.m == 0.0) {
        p.k = 0;
        p.wr = p.wi = (F)0;
    } else {
        p.k = e;
        p.wr = ldexp_t(p.dcr, p.dcre - e);
        p.wi = ldexp_t(p.dci, p.dcie - e);
    }
    set_scale(p);
    p.m = 1;
    p.iter = 1;
    p.rm = 0;
    p.jr = (F)1;
    p.ji = (F)0;
    p.kj = 0;
    p.jone = (F)1;
    p.zsr = p.zsi = (F)1e30f;
    p.save_iter = 1;
    p.period_win = 1;
    p.conf_P = p.conf_next = p.conf_count = p.conf_giveup = 0;
    p.conf_D2 = p.conf_zr = p.conf_zi = (F)0;
    p.swr = p.swi = (F)3e38f;
    p.sk = INT_MIN;
    p.sm = -1;
    p.save_iter2 = 1;
    p.twin2r = 1;
    p.xv = 0;
}

// |w|^2 * 2^(2k) < r2 (r2 an FP64 radius^2 > 0): the FloatExp comparison of
// tryBLAfe (mandel_perturbation.cpp:2891-2897) on exponents and mantissas.
template <typename F>
MX_HD bool dz2_less(F w2, int k2, double r2) {
    if (w2 == (F)0) return true;
    int ew, er;
    const double mw = frexp((double)w2, &ew);
    const double mr = frexp(r2, &er);
    const int d = ew + k2 - er;
    if (d < 0) return true;
    if (d > 0) return false;
    return mw < mr;
}

// Orbit entry and BLA coefficients in the delta's precision: FP32 copies with a
// per-entry exponent for the mixed form, the FP64 tables for the FP64 variant.
MX_HD float2 load_X(const MxRefDev& R, int i, float) { return R.zf32[i]; }
MX_HD double2 load_X(const MxRefDev& R, int i, double) { return R.zf[i]; }
MX_HD void load_coef(const MxRefDev& R, int e, float4& c, int2& ce) {
    c = R.bla_coeff[e];
    ce = R.bla_coefe[e];
}
MX_HD void load_coef(const MxRefDev& R, int e, double4& c, int2& ce) {
    c = R.bla_coef[e];
    ce = make_int2(0, 0);
}
// Orbit entries for the lean steps through a buffer resource: one VALU shift per
// address instead of a 64-bit add, and indices past the end read as 0.
template <typename F>
struct OrbitLoader;
template <>
struct OrbitLoader<float> {
    __amdgpu_buffer_rsrc_t r;
    int n;
    MX_HD explicit OrbitLoader(const MxRefDev& R) : n(R.nent) {
#ifdef __HIP_DEVICE_COMPILE__
        r = __builtin_amdgcn_make_buffer_rsrc((void*)R.zf32.p, 0, R.nent * 8, 0x00020000);
#endif
        zf = R.zf32;
    }
    GPtr<const float2> zf;
    MX_HD float2 load(int i) const {
#ifdef __HIP_DEVICE_COMPILE__
        return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, i * 8, 0, 0));
#else
        return i < n ? zf[i] : float2{0.f, 0.f};
#endif
    }
};
template <>
struct OrbitLoader<double> {
    GPtr<const double2> zf;
    int n;
    MX_HD explicit OrbitLoader(const MxRefDev& R) : zf(R.zf), n(R.nent) {}
    MX_HD double2 load(int i) const { return zf[min(i, n - 1)]; }
};

template <typename F>
struct Vec4;
template <>
struct Vec4<float> {
    using T = float4;
};
template <>
struct Vec4<double> {
    using T = double4;
};

// Results of the cold recoveries, returned by value: a reference to the lane
// state across a call would force that state out of registers into scratch.
struct WK {
    double wr, wi;
    int k;
    bool zero;
};

// Exact recovery of a step whose scaled result fell out of the delta's range
// (feRescaledStepUnderflow, mandel_perturbation.cpp:2794-2835, with S = 2^k): after a
// rebase onto a tiny z, dz^2 and dc can lie far below the current scale.  The terms
// 2 X dz, dz^2 and dc are aligned to their largest exponent in FP64 (X from the FP64
// FloatExp orbit) and the sum is renormalised into (w, k).
__host__ __device__ __noinline__ WK step_recover(const MxRefDev& R, int m, int rm, double wr, double wi, int k, double dcr, int dcre, double dci, int dcie) {
    double xrm = 0.0, xim = 0.0;
    int xre = 0, xie = 0;
    if (m) {
        const double2 mm = R.zfe_m[rm];
        const int2 ee = R.zfe_e[rm];
        xrm = mm.x;
        xim = mm.y;
        xre = ee.x;
        xie = ee.y;
    }
    const bool has_w = wr != 0.0 || wi != 0.0;
    int e = INT_MIN;
    if (has_w && xrm != 0.0) e = max(e, xre + k);
    if (has_w && xim != 0.0) e = max(e, xie + k);
    if (has_w) e = max(e, k + k);
    if (dcr != 0.0) e = max(e, dcre);
    if (dci != 0.0) e = max(e, dcie);
    if (e == INT_MIN) return WK{0.0, 0.0, k, true};
    auto al = [e](double mm, int te) { return (mm == 0.0 || te - e < -1000) ? 0.0 : ldexp(mm, te - e); };
    const double xr = al(xrm, xre + k), xi = al(xim, xie + k), s2 = al(1.0, k + k);
    const double cr = al(dcr, dcre), ci = al(dci, dcie);
    const double nr = 2.0 * (xr * wr - xi * wi) + s2 * (wr * wr - wi * wi) + cr;
    const double ni = 2.0 * (xr * wi + xi * wr) + 2.0 * s2 * wr * wi + ci;
    const double a = fmax(fabs(nr), fabs(ni));
    if (a == 0.0) return WK{0.0, 0.0, k, true};
    int ea;
    (void)frexp(a, &ea);
    return WK{ldexp(nr, -ea), ldexp(ni, -ea), e + ea, false};
}

// J' = 2 z J + 1 when z and the "+1" both vanish at their scales
// (feDerivativeStepUnderflow, mandel_perturbation.cpp:2842-2876): z from the FP64
// orbit plus dz, aligned with 2^-kj.  Returns J' as (jr, ji, kj).
__host__ __device__ __noinline__ WK deriv_recover(const MxRefDev& R, int m, int rm, double wr, double wi, int k, double jr, double ji, int kj) {
    double xrm = 0.0, xim = 0.0;
    int xre = 0, xie = 0;
    if (m) {
        const double2 mm = R.zfe_m[rm];
        const int2 ee = R.zfe_e[rm];
        xrm = mm.x;
        xim = mm.y;
        xre = ee.x;
        xie = ee.y;
    }
    // z = X + w 2^k at a common exponent ez
    int ez = INT_MIN;
    if (xrm != 0.0) ez = max(ez, xre);
    if (xim != 0.0) ez = max(ez, xie);
    if (wr != 0.0 || wi != 0.0) ez = max(ez, k);
    double zr = 0.0, zi = 0.0;
    if (ez != INT_MIN) {
        auto al = [ez](double mm, int te) { return (mm == 0.0 || te - ez < -1000) ? 0.0 : ldexp(mm, te - ez); };
        zr = al(xrm, xre) + al(wr, k);
        zi = al(xim, xie) + al(wi, k);
    }
    // 2 z J at 2^(ez + kj), 1 at 2^0
    const bool has = (zr != 0.0 || zi != 0.0) && (jr != 0.0 || ji != 0.0);
    const int e = has ? max(ez + kj, 0) : 0;
    const double t = has ? ldexp(1.0, max(ez + kj - e, -1000)) : 0.0;
    const double ar = 2.0 * (zr * jr - zi * ji) * t + ldexp(1.0, max(-e, -1000));
    const double ai = 2.0 * (zr * ji + zi * jr) * t;
    const double a = fmax(fabs(ar), fabs(ai));
    int ea;
    (void)frexp(a, &ea);
    return WK{ldexp(ar, -ea), ldexp(ai, -ea), e + ea, false};
}

// The common-case arithmetic of one step, shared by the full and the lean step.
template <typename F>
MX_HD F wmag2(F wr, F wi) { return fma_t(wr, wr, wi * wi); }

// z = X + s w (the pre-step value, which the derivative step also uses), then
// w' = (X + z) w + d = (2X + s w) w + d
template <typename F, typename V>
MX_HD void step_z(const P32<F>& p, V X, F& zmr, F& zmi) {
    zmr = fma_t(p.s, p.wr, X.x);
    zmi = fma_t(p.s, p.wi, X.y);
}
template <typename F, typename V>
MX_HD void step_w(const P32<F>& p, V X, F zmr, F zmi, F& nwr, F& nwi) {
    const F fR = zmr + X.x;
    const F fI = zmi + X.y;
    nwr = fma_t(fR, p.wr, fma_t(-fI, p.wi, p.dr));
    nwi = fma_t(fR, p.wi, fma_t(fI, p.wr, p.di));
}

// a step result far below |w| ~ 1: the terms live below the current scale
template <typename F>
MX_HD bool step_tiny(const P32<F>& p, F nwr, F nwi) {
    const F TINY = sizeof(F) == 4 ? (F)0x1p-60f : (F)0x1p-1000;
    return fabs_t(nwr) < TINY && fabs_t(nwi) < TINY && (nwr != (F)0 || nwi != (F)0 || p.wr != (F)0 || p.wi != (F)0 || p.dcr != (F)0 || p.dci != (F)0);
}

// J' = 2 z J + 1 in J's scale, z = X + s w (the pre-step delta).  Both components
// take the doubled z, so the imaginary part needs no doubling of its own: the result is
// the same as doubling fma(zr, ji, zi jr) except where that is subnormal.
template <typename F>
MX_HD void step_j(const P32<F>& p, F zmr, F zmi, F& a, F& b) {
    const F zr2 = zmr + zmr, zi2 = zmi + zmi;
    a = fma_t(zr2, p.jr, fma_t(-zi2, p.ji, p.jone));
    b = fma_t(zr2, p.ji, zi2 * p.jr);
}

template <typename F>
MX_HD bool renorm_due(F m2) { return m2 > (F)0x1p40f || (m2 < (F)0x1p-40f && m2 > (F)0); }

// BLA skip from table index s = m-1 (tryBLAfe, mandel_perturbation.cpp:2883-2970)
// with floatexp coefficients (mantissa c, exponents ce).  Updates (w, k) and J.
template <typename F>
MX_HD int try_bla32(const MxRefDev& R, GPtr<const int2> lev, int s, P32<F>& p, int reflen, F esc2f, bool ede) {
    using V = typename Vec2<F>::T;
    using V4 = typename Vec4<F>::T;
    if (s < 1) return 0;
    const int maxp = R.nlev - 1;
    int startp = (s == 1) ? maxp : __builtin_ctz((unsigned)(s - 1));
    if (startp > maxp) startp = maxp;
    if (startp < R.minlevel) return 0;
    const F w2 = wmag2(p.wr, p.wi);
    const int k2 = 2 * p.k;
    if (!(R.rmax2 > 0.0) || !dz2_less(w2, k2, R.rmax2)) return 0;
    // the first level whose entry is valid decides: it is applied, or the search ends
    // with no skip (an escaping or unrepresentable landing does not fall back to a
    // shorter level), so the level search is a loop of its own
    int q = startp, e = -1;
    for (; q >= R.minlevel; --q) {
        const int2 ls = lev[q];
        const int i = (s - 1) >> q;
        if (i >= ls.y) continue;
        const int ee = ls.x + i;
        const double r2 = R.bla_r2[ee];
        if (!(r2 > 0.0) || !dz2_less(w2, k2, r2)) continue;
        if (s + (1 << q) >= reflen - 1) continue;
        e = ee;
        break;
    }
    if (e < 0) return 0;
    const int land = s + (1 << q);
    V4 c;
    int2 ce;
    load_coef(R, e, c, ce);
    // A w and B d with their exponents, then a common exponent
    const F ar = c.x * p.wr - c.y * p.wi, ai = c.x * p.wi + c.y * p.wr;
    const F br = c.z * p.dr - c.w * p.di, bi = c.z * p.di + c.w * p.dr;
    const F aa = fmax_t(fabs_t(ar), fabs_t(ai)), bb = fmax_t(fabs_t(br), fabs_t(bi));
    int ea = INT_MIN / 2, eb = INT_MIN / 2;
    if (aa != (F)0) {
        (void)frexp_t(aa, &ea);
        ea += ce.x;
    }
    if (bb != (F)0) {
        (void)frexp_t(bb, &eb);
        eb += ce.y;
    }
    const int es = max(ea, eb);
    F nwr, nwi;
    if (es == INT_MIN / 2) {
        nwr = nwi = (F)0;
    } else {
        nwr = ldexp_t(ar, ce.x - es) + ldexp_t(br, ce.y - es);
        nwi = ldexp_t(ai, ce.x - es) + ldexp_t(bi, ce.y - es);
    }
    const int nk = p.k + es;
    // never land past an escape: z = X_land + 2^nk w'
    const V X = load_X(R, land, (F)0);
    const F sn = pow2t<F>(nk);
    const F fr = X.x + sn * nwr, fi = X.y + sn * nwi;
    if (fr * fr + fi * fi > esc2f) return 0;
    // unrepresentable landing: rejected on a full reference, kept as a zero delta at
    // unit scale on a periodic one (as the plain form)
    const bool wzero = nwr == (F)0 && nwi == (F)0;
    if (wzero && !R.period) return 0;
    if (ede) {
        // J -> A J + B (mandel_perturbation.cpp:3399-3414): A J at 2^(kj + ce.x), B at 2^ce.y
        const F tr = c.x * p.jr - c.y * p.ji, ti = c.x * p.ji + c.y * p.jr;
        const F ta = fmax_t(fabs_t(tr), fabs_t(ti));
        int ej = INT_MIN / 2, eB = INT_MIN / 2;
        if (ta != (F)0) {
            (void)frexp_t(ta, &ej);
            ej += p.kj + ce.x;
        }
        const F tb = fmax_t(fabs_t(c.z), fabs_t(c.w));
        if (tb != (F)0) {
            (void)frexp_t(tb, &eB);
            eB += ce.y;
        }
        const int em = max(ej, eB);
        if (em == INT_MIN / 2) {
            p.jr = p.ji = (F)0;
            p.kj = 0;
        } else {
            p.jr = ldexp_t(tr, p.kj + ce.x - em) + ldexp_t(c.z, ce.y - em);
            p.ji = ldexp_t(ti, p.kj + ce.x - em) + ldexp_t(c.w, ce.y - em);
            p.kj = em;
        }
        p.jone = pow2t<F>(-p.kj);
    }
    p.wr = nwr;
    p.wi = nwi;
    p.k = wzero ? 0 : nk;
    set_scale(p);
    return 1 << q;
}

template <typename F, bool EDE, bool INTERIOR>
MX_HD bool trip32(const MxRefDev& R, GPtr<const int2> lev, const MxTileDev& T, P32<F>& p, double& v, int& n, C32& c, F esc2f) {
    using V = typename Vec2<F>::T;
    const int reflen = R.mx_ref_it + 1;
    // a periodic reference wraps instead of ending (mandel_perturbation.cpp:3308-3313)
    const int per = R.period;
    const int mend = per ? INT_MAX : reflen;
    if (p.iter >= T.mxit) {
        n = -1;
        v = -2.0;
        return true;
    }
    if (INTERIOR && (T.flags & MX_F_INT_REP)) {
        if (p.m == p.sm && p.wr == p.swr && p.wi == p.swi && p.k == p.sk) {
            n = -2;
            v = -2.0;
            return true;
        }
        if (p.iter - p.save_iter2 >= p.twin2r) {
            p.swr = p.wr;
            p.swi = p.wi;
            p.sk = p.k;
            p.sm = p.m;
            p.save_iter2 = p.iter;
            p.twin2r += p.twin2r;
        }
    }
    if ((T.flags & MX_F_BLA) && p.m >= 2) {
        const int skip = try_bla32(R, lev, p.rm, p, per ? per : reflen, esc2f, EDE);
        if (skip > 0) {
            p.m += skip;
            p.iter += skip;
            p.rm += skip;
            if (per && p.rm >= per) p.rm -= per;
            c.applies++;
            c.skipped += skip;
            p.zsr = p.zsi = (F)1e30f;
            p.save_iter = p.iter;
            p.period_win = 1;
            p.conf_P = 0;
            p.conf_giveup = 0;
            p.xv = 0;
            return false;
        }
    }
    if (!p.xv) {
        p.Xc = p.m ? load_X(R, p.rm, (F)0) : V{(F)0, (F)0};
        p.Xn = load_X(R, p.rm + 1, (F)0);
    }
    const V Xpf = (p.rm + 2 < R.nent) ? load_X(R, p.rm + 2, (F)0) : V{(F)0, (F)0};
    const int rm_prev = p.rm;
    const V X = p.Xc;
    F zmr, zmi;
    step_z(p, X, zmr, zmi);
    F nwr, nwi;
    step_w(p, X, zmr, zmi, nwr, nwi);
    const F wr0 = p.wr, wi0 = p.wi;
    const int k0 = p.k;
    // |w| stays near 1; a result far below it means the step's terms live below the
    // current scale (a delta that underflowed), recomputed exactly
    if (step_tiny(p, nwr, nwi)) {
        const WK r = step_recover(R, p.m, p.rm, (double)p.wr, (double)p.wi, p.k, (double)p.dcr, p.dcre, (double)p.dci, p.dcie);
        p.wr = (F)r.wr;
        p.wi = (F)r.wi;
        if (!r.zero) {
            p.k = r.k;
            set_scale(p);
        }
    } else {
        p.wr = nwr;
        p.wi = nwi;
    }
    ++p.m;
    ++p.iter;
    ++p.rm;
    if (per && p.rm == per) p.rm = 0;
    c.steps++;
    if (EDE) {
        // J' = 2 z J + 1 in J's own scale
        if (zmr == (F)0 && zmi == (F)0 && p.jone == (F)0) {
            const WK r = deriv_recover(R, p.m - 1, rm_prev, (double)wr0, (double)wi0, k0, (double)p.jr, (double)p.ji, p.kj);
            p.jr = (F)r.wr;
            p.ji = (F)r.wi;
            p.kj = r.k;
            p.jone = pow2t<F>(-p.kj);
        } else {
            F a, b;
            step_j(p, zmr, zmi, a, b);
            p.jr = a;
            p.ji = b;
            if (renorm_due(wmag2(a, b))) renorm_j(p);
        }
    }
    const V Xn = p.Xn;
    p.Xc = Xn;
    p.Xn = Xpf;
    p.xv = 1;
    const F zr = fma_t(p.s, p.wr, Xn.x), zi = fma_t(p.s, p.wi, Xn.y);
    const F zrad = wmag2(zr, zi);
    if (zrad > esc2f) {
        n = p.iter;
        const double zd = (double)zr * (double)zr + (double)zi * (double)zi;
        if (EDE) {
            // |z| log|z|^2 / (dx |J|) with |J| = |jw| 2^kj
            const double num = sqrt(zd) * log(zd);
            const double jm = sqrt((double)p.jr * p.jr + (double)p.ji * p.ji);
            v = ldexp(num / (T.dxfe.m * jm), -p.kj - T.dxfe.e);
        } else {
            const double LG2 = 0.69314718055994530942;
            v = (double)p.iter - log(log(zd) / 2.0 / LG2) / LG2;
        }
        return true;
    }
    if (INTERIOR && (T.flags & MX_F_INTERIOR)) {
        const F eps2 = (F)T.interior_eps2;
        if (p.conf_P > 0) {
            p.conf_D2 = fmin_t(p.conf_D2 * (F)4 * zrad, (F)1e18f);
            if (p.iter >= p.conf_next) {
                const F pr = zr - p.conf_zr, pi = zi - p.conf_zi;
                if (wmag2(pr, pi) < eps2 * zrad && p.conf_D2 < (F)1) {
                    if (++p.conf_count >= T.interior_confirm) {
                        n = -2;
                        v = -2.0;
                        return true;
                    }
                    p.conf_zr = zr;
                    p.conf_zi = zi;
                    p.conf_D2 = (F)1;
                    p.conf_next = p.iter + p.conf_P;
                } else {
                    p.conf_P = 0;
                    ++p.conf_giveup;
                }
            }
        } else if ((p.iter & 15) == 0 && p.conf_giveup < 3) {
            const F pr = zr - p.zsr, pi = zi - p.zsi;
            if (wmag2(pr, pi) < eps2 * zrad) {
                p.conf_P = max(p.iter - p.save_iter, 1);
                p.conf_zr = zr;
                p.conf_zi = zi;
                p.conf_D2 = (F)1;
                p.conf_next = p.iter + p.conf_P;
                p.conf_count = 0;
            }
            if (p.iter - p.save_iter >= p.period_win) {
                p.zsr = zr;
                p.zsi = zi;
                p.save_iter = p.iter;
                p.period_win += p.period_win;
            }
        }
    }
    if (renorm_due(wmag2(p.wr, p.wi))) renorm_w(p);
    // Rebase candidates: |z|^2 < 1e-8, |z|^2 < |dz|^2 (while 2^(2k) is
    // representable), or the end of the reference; then the exact test and the new
    // delta in FP64 from the FP64 orbit (mandel_perturbation.cpp:3584-3615).
    const F dz2f = ldexp_t(wmag2(p.wr, p.wi), 2 * p.k);
    if (zrad < (F)1e-8f || zrad < dz2f || p.m >= mend) {
        // z * 2^-k = X * 2^-k + w in FP64
        double xr = 0.0, xi = 0.0;
        if (p.m) {
            const double2 mm = R.zfe_m[p.rm];
            const int2 ee = R.zfe_e[p.rm];
            xr = mm.x == 0.0 ? 0.0 : ldexp(mm.x, max(min(ee.x - p.k, 2000), -2000));
            xi = mm.y == 0.0 ? 0.0 : ldexp(mm.y, max(min(ee.y - p.k, 2000), -2000));
        }
        const double zsr = xr + (double)p.wr, zsi = xi + (double)p.wi;
        const double zs2 = zsr * zsr + zsi * zsi;
        const double ws2 = (double)p.wr * p.wr + (double)p.wi * p.wi;
        if (p.m >= mend || zs2 < ws2) {
            const double a = fmax(fabs(zsr), fabs(zsi));
            if (a == 0.0) {
                p.wr = p.wi = (F)0;
            } else {
                int e;
                (void)frexp(a, &e);
                p.wr = (F)ldexp(zsr, -e);
                p.wi = (F)ldexp(zsi, -e);
                p.k += e;
                set_scale(p);
            }
            p.m = 0;
            p.rm = -1;
            p.xv = 0;
            c.rebases++;
        }
    }
    return false;
}

// Bit pattern of a non-negative float / double: monotone in the value (NaN and +inf
// above every finite value), so a range test is one subtraction and one compare.
MX_HD unsigned fbits(float x) { return __builtin_bit_cast(unsigned, x); }
MX_HD unsigned long long fbits(double x) { return __builtin_bit_cast(unsigned long long, x); }

// Lean steps: exactly the common case of trip32, for up to `budget` steps.  A lane
// stops, with its state untouched, at the first step whose full path could do more
// than the plain step: every condition below is the full path's own test or a
// superset of it (the full path then decides), evaluated without branches and joined
// into one exit.
//   before the step: the iteration limit; an active periodicity confirmation; an
//     exact-repeat save, or the orbit index back at the saved one (a repeat needs it);
//     a BLA-aligned index whose delta is not above the largest radius by exponent;
//   after it: |w'|^2 outside [2^-40, 2^40] (renormalisation, or a step underflow, or
//     zero), the same for the derivative, a derivative step whose terms all vanish,
//     |z|^2 outside [1e-8, escape radius^2] (escape, rebase candidate), |z| < |dz|
//     (rebase candidate), the end of the reference, and a near return on a
//     periodicity check.
template <typename F, bool EDE, bool INTERIOR>
MX_HD void fast_steps(const MxRefDev& R, const MxTileDev& T, P32<F>& p, C32& c, F esc2f, int budget) {
    using V = typename Vec2<F>::T;
    using U = decltype(fbits((F)0));
    const int reflen = R.mx_ref_it + 1;
    const int amask = (1 << R.minlevel) - 1;
    const bool can_bla = (T.flags & MX_F_BLA) && R.nlev - 1 >= R.minlevel && R.rmax2 > 0.0;
    int er = 0;
    if (can_bla) (void)frexp(R.rmax2, &er);
    const F eps2 = (F)T.interior_eps2;
    const F R2LO = (F)0x1p-40f, R2HI = (F)0x1p40f;
    const U r2lo = fbits(R2LO), r2span = fbits(R2HI) - fbits(R2LO);
    const F ZLO = (F)1e-8f;
    const U zlo = fbits(ZLO), zspan = fbits(esc2f) - fbits(ZLO);
    const bool int_rep = INTERIOR && (T.flags & MX_F_INT_REP);
    const OrbitLoader<F> ld(R);
    F w2 = wmag2(p.wr, p.wi);
    // one counter for the steps taken: orbit index, reference index and iteration
    // advance together, and the tests on them alone bound the trip count instead: the
    // loop never takes the step at which one would stop it
    const int m0 = p.m, rm0 = p.rm, it0 = p.iter;
    int lim = min(budget, T.mxit - it0);              // iter >= mxit
    // m + 1 >= reflen after the step; periodic: rm + 1 == period (the wrap, taken by
    // the full step)
    lim = min(lim, R.period ? R.period - 1 - rm0 : reflen - 1 - m0);
    if (int_rep) {
        lim = min(lim, p.save_iter2 + p.twin2r - it0); // iter - save_iter2 >= twin2r
        if (p.sm - m0 >= 0) lim = min(lim, p.sm - m0);   // m == sm
    }
    if (INTERIOR && p.conf_P > 0) lim = 0;            // an active periodicity confirmation
    // a BLA-aligned index (rm - 1 a multiple of 2^minlevel, m >= 2) whose |dz|^2 is not
    // above the largest radius by exponent: 2^(2k) |w|^2 with e(|w|^2) <= er - 2k, i.e.
    // |w|^2 <= 2^(er - 2k) (equality included: a superset)
    const int emax = sizeof(F) == 4 ? 128 : 1024, emin = sizeof(F) == 4 ? -150 : -1075; // inf, 0
    const F wthr = can_bla ? ldexp_t((F)1, max(min(er - 2 * p.k, emax), emin)) : (F)0;
    const int nal = 1 - rm0, nm2 = 2 - m0;
    const bool chk_ok = INTERIOR && (T.flags & MX_F_INTERIOR) && p.conf_giveup < 3;
    int n = 0;
#ifdef __HIP_DEVICE_COMPILE__
    // the wave ends the run together once fewer than quorum/64 of the lanes that start
    // it are still stepping: a lane that has stopped idles until the run ends
    const int need = (__popcll(__ballot(lim > 0)) * T.quorum + 63) >> 6;
#endif
    for (int u = 0; u < lim; ++u) {
        const int rm = rm0 + n;
        int stop = 0;
        if (can_bla) stop = (int)(((n - nal) & amask) == 0) & (int)(n >= nm2) & (int)(w2 <= wthr);
        if (!p.xv) {
            p.Xc = m0 + n ? load_X(R, rm, (F)0) : V{(F)0, (F)0};
            p.Xn = load_X(R, rm + 1, (F)0);
            p.xv = 1;
        }
        // the next step's entry (0 past the end, where this step stops anyway)
        const V Xpf = ld.load(rm + 2);
        const V X = p.Xc;
        F zmr, zmi;
        step_z(p, X, zmr, zmi);
        F nwr, nwi;
        step_w(p, X, zmr, zmi, nwr, nwi);
        const F w2n = wmag2(nwr, nwi);
        stop |= fbits(w2n) - r2lo > r2span;
        F ja = (F)0, jb = (F)0;
        if (EDE) {
            step_j(p, zmr, zmi, ja, jb);
            stop |= ((int)(zmr == (F)0) & (int)(zmi == (F)0) & (int)(p.jone == (F)0)) | (int)(fbits(wmag2(ja, jb)) - r2lo > r2span);
        }
        const V Xn = p.Xn;
        const F zr = fma_t(p.s, nwr, Xn.x), zi = fma_t(p.s, nwi, Xn.y);
        const F zrad = wmag2(zr, zi);
        stop |= (int)(fbits(zrad) - zlo > zspan) | (int)(zrad < ldexp_t(w2n, 2 * p.k));
        const int it1 = it0 + n + 1;
        int save = 0;
        if (INTERIOR) {
            const int chk = (int)((it1 & 15) == 0) & (int)chk_ok;
            const F pr = zr - p.zsr, pi = zi - p.zsi;
            stop |= chk & (int)(wmag2(pr, pi) < eps2 * zrad);
            save = chk & (int)(it1 - p.save_iter >= p.period_win);
        }
        // commit by selects, so that the entry loaded above is consumed only by the
        // next step and its load overlaps this one
        const bool go = !stop;
        p.wr = go ? nwr : p.wr;
        p.wi = go ? nwi : p.wi;
        if (EDE) {
            p.jr = go ? ja : p.jr;
            p.ji = go ? jb : p.ji;
        }
        // the entries rotate on every step, so that the load above feeds the next step
        // straight from its register; a lane that stops reloads both (xv = 0 below)
        p.Xc = Xn;
        p.Xn = Xpf;
        if (INTERIOR && go && save) {
            p.zsr = zr;
            p.zsi = zi;
            p.save_iter = it1;
            p.period_win += p.period_win;
        }
        w2 = w2n;
        n += go;
        if (stop) {
            p.xv = 0;
            break;
        }
#ifdef __HIP_DEVICE_COMPILE__
        if (__popcll(__ballot(1)) < need) break;
#endif
    }
    p.m = m0 + n;
    p.rm = rm0 + n;
    p.iter = it0 + n;
    c.steps += (unsigned)n;
}

__device__ __forceinline__ unsigned long long wave_sum(unsigned long long x) {
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
    return x;
}

MX_HD bool item_to_sample(const MxTileDev& T, unsigned item, int& si, int& sj) {
    const unsigned tile = item >> 6, l = item & 63;
    const int ty = (int)(tile / (unsigned)T.tiles_x), tx = (int)(tile % (unsigned)T.tiles_x);
    const int gi = ty * 8 + (int)(l >> 3), gj = tx * 8 + (int)(l & 7);
    if (gi >= T.ny || gj >= T.nx) return false;
    si = gi * T.stride;
    sj = gj * T.stride;
    return si * T.W + sj != T.skip_sample;
}

// The next sample of work item `item`: a grid sample (separable dc tables, output at
// row * W + column) or an entry of the explicit-dc list (output at the entry).
MX_HD bool take_sample(const MxTileDev& T, unsigned item, int& sidx, Fe& dcr, Fe& dci) {
    if (T.sdc) {
        if (item >= (unsigned)T.list_len) return false;
        sidx = (int)item;
        dcr = T.sdc[2 * item];
        dci = T.sdc[2 * item + 1];
        return true;
    }
    int si, sj;
    if (!item_to_sample(T, item, si, sj)) return false;
    sidx = si * T.W + sj;
    dcr = T.dcol[sj];
    dci = T.drow[si];
    return true;
}

template <typename F, bool EDE, bool INTERIOR>
__device__ __forceinline__ void run_job32(const MxRefDev& R, GPtr<const int2> lev, const MxTileDev& T) {
    const int lane = __lane_id();
    const unsigned long long lanemask_lt = (1ull << lane) - 1ull;
    const F esc2f = (F)T.esc2;
    int q = blockIdx.x % MX_NQUEUE;
    int exhausted = 0;
    bool alive = true, active = false;
    P32<F> p;
    // each lane's work counters and output index live in LDS, not in registers: they
    // are touched once per trip and would otherwise hold seven VGPRs across the loop
    __shared__ C32 lds_counters[256];
    __shared__ int lds_sidx[256];
    C32& c = lds_counters[threadIdx.x];
    int& sidx = lds_sidx[threadIdx.x];
    c = C32{0, 0, 0, 0};
    sidx = 0;
    unsigned trips = 0;
    const unsigned ntiles = (unsigned)T.ntiles;
    while (true) {
        bool need = alive && !active;
        unsigned long long needmask = __ballot(need);
        while (needmask != 0ull && exhausted < MX_NQUEUE) {
            const unsigned t0 = ntiles * (unsigned)q / MX_NQUEUE, t1 = ntiles * (unsigned)(q + 1) / MX_NQUEUE;
            const unsigned nitems = (t1 - t0) * 64u;
            const int kk = __popcll(needmask);
            const int leader = __ffsll((long long)needmask) - 1;
            unsigned base = 0;
            if (lane == leader) base = atomicAdd(&T.queue[q], (unsigned)kk);
            base = __shfl(base, leader);
            if (need) {
                const unsigned it = base + (unsigned)__popcll(needmask & lanemask_lt);
                if (it < nitems) {
                    Fe dcr, dci;
                    if (take_sample(T, t0 * 64u + it, sidx, dcr, dci)) {
                        init32(p, dcr, dci);
                        need = false;
                        active = true;
                    }
                }
            }
            if (base + (unsigned)kk >= nitems) {
                q = (q + 1) % MX_NQUEUE;
                ++exhausted;
            }
            needmask = __ballot(need);
        }
        if (need) alive = false;
        if (__ballot(active) == 0ull) break;
        if (((++trips) & 4095u) == 0u && T.cancel && *T.cancel) break;
        if (active) {
            double v;
            int n;
            if (trip32<F, EDE, INTERIOR>(R, lev, T, p, v, n, c, esc2f)) {
                T.val[sidx] = v;
                T.nesc[sidx] = n;
                if (T.out) T.out[sidx] = n < 0 ? -2.0f : (float)v;
                active = false;
            } else {
                fast_steps<F, EDE, INTERIOR>(R, T, p, c, esc2f, T.chunk);
            }
        }
    }
    const unsigned long long s0 = wave_sum(c.steps), s1 = wave_sum(c.applies), s2 = wave_sum(c.skipped), s3 = wave_sum(c.rebases);
    if (lane == 0) {
        atomicAdd(&T.stats[0], s0);
        atomicAdd(&T.stats[1], s1);
        atomicAdd(&T.stats[2], s2);
        atomicAdd(&T.stats[3], s3);
    }
}

template <typename F, bool EDE, bool INTERIOR>
__global__ __launch_bounds__(256) MX_WAVES_ATTR(sizeof(F) == 8 ? MX_SC64_WAVES : MX_FE32_WAVES) void mx_fe32_kernel(const MxJob* __restrict__ jobs, int njobs) {
    for (int j = 0; j < njobs; ++j) run_job32<F, EDE, INTERIOR>(jobs[j].R, jobs[j].lev, jobs[j].T);
}

} // namespace

template <typename F>
static void launch_scaled(const MxJob* jobs, int njobs, bool ede, bool interior, int blocks, int threads, hipStream_t stream) {
    if (ede && interior)
        hipLaunchKernelGGL((mx_fe32_kernel<F, true, true>), dim3(blocks), dim3(threads), 0, stream, jobs, njobs);
    else if (ede)
        hipLaunchKernelGGL((mx_fe32_kernel<F, true, false>), dim3(blocks), dim3(threads), 0, stream, jobs, njobs);
    else if (interior)
        hipLaunchKernelGGL((mx_fe32_kernel<F, false, true>), dim3(blocks), dim3(threads), 0, stream, jobs, njobs);
    else
        hipLaunchKernelGGL((mx_fe32_kernel<F, false, false>), dim3(blocks), dim3(threads), 0, stream, jobs, njobs);
}

// FP32 delta (the mixed form); `jobs` is a device array, every job with the same EDE
// and interior-detection flags
hipError_t mx_launch_fe32(const MxJob* jobs, int njobs, bool ede, bool interior, int blocks, int threads, hipStream_t stream) {
    launch_scaled<float>(jobs, njobs, ede, interior, blocks, threads, stream);
    return hipGetLastError();
}

// The same power-of-two-scaled algorithm with an FP64 delta
hipError_t mx_launch_sc64(const MxJob* jobs, int njobs, bool ede, bool interior, int blocks, int threads, hipStream_t stream) {
    launch_scaled<double>(jobs, njobs, ede, interior, blocks, threads, stream);
    return hipGetLastError();
}
