// Plain form of the per-pixel rescaled perturbation kernel for gfx950: IEEE FP64
// with the engine's operation order and no FP contraction (this file is built with
// -ffp-contract=off), so every escape count equals the CPU engine's.
//
// Per-pixel algorithm: Mandel::pixelRescaled (src/engine/mandel_perturbation.cpp:
// 3287-3618) with tryBLAfe (2883-2970) and the cold underflow recoveries (2794-2876):
// dz = S*w with w an O(1) double and S a FloatExp, fused step
// w' = (2X + s w) w + d, BLA skips tested in FloatExp, rescale of |w| into S,
// Zhuoran rebase when |z| < |dz|, and the carried derivative J for the exterior
// distance estimate.
//
// Execution model: a persistent grid; each wavefront lane owns one sample.  A lane
// whose sample finishes takes the next one from a per-XCD atomic work queue (one
// queue per XCD, wave-aggregated fetch-add; an XCD that drains its queue moves on
// to the others), the GPU analogue of the engine's AVX2 lane refill
// (mandel_perturbation.cpp:2978-3282).  Samples are handed out in 8x8 tiles so the
// 64 lanes of a wave start on neighbouring pixels, which follow similar orbit
// indices and BLA decisions.
#include <climits>

#include "mx_dev.h"

namespace {

struct Px {
    Fe dcr, dci;
    Fe S, S2, SJ, sS;
    double wr, wi, dr, di, s;
    double zsr, zsi, conf_D2, conf_zr, conf_zi;
    double swr, swi;
    double jr, ji, invSJd;
    int m, rm, iter;
    int save_iter, period_win, conf_P, conf_next, conf_count, conf_giveup;
    int sm, save_iter2, twin2r;
    int xv;          // 1: Xc = X_m and Xn = X_{m+1} are loaded
    double2 Xc, Xn;  // orbit entries carried from the previous trip
};

struct Counters {
    unsigned long long steps, applies, skipped, rebases;
};

// minimum waves per SIMD requested from the register allocator (0: its own choice):
// 4 (128 VGPRs) for the variants whose spills then sit on the full path's rare
// branches (the added occupancy hides more of the trips' latency: the 5760x3240 EDE
// frame runs 5 % faster than at 3 waves), 3 for EDE with interior detection, which
// would spill on its common path
#ifndef MX_FE64_WAVES
#define MX_FE64_WAVES 4
#endif
#ifndef MX_FE64_WAVES_EDE_INTERIOR
#define MX_FE64_WAVES_EDE_INTERIOR 3
#endif
#define MX_WAVES_ATTR(n) MX_WAVES_ATTR_(n)
#define MX_WAVES_ATTR_(n) __attribute__((amdgpu_waves_per_eu((n) > 0 ? (n) : 1)))

MX_HD void init_pixel(Px& p, Fe dcr, Fe dci) {
    p.dcr = dcr;
    p.dci = dci;
    p.S = fe_sqrt(fe_add(fe_mul(dcr, dcr), fe_mul(dci, dci)));
    if (p.S.m == 0.0) {
        p.S = fe_make(1.0, 0);
        p.wr = p.wi = p.dr = p.di = 0.0;
    } else {
        p.wr = fe_to_double(fe_div(dcr, p.S));
        p.wi = fe_to_double(fe_div(dci, p.S));
        p.dr = p.wr;
        p.di = p.wi;
    }
    p.s = fe_to_double(p.S);
    p.S2 = fe_mul(p.S, p.S);
    p.m = 1;
    p.iter = 1;
    p.rm = 0;
    p.zsr = 1e30;
    p.zsi = 1e30;
    p.save_iter = 1;
    p.period_win = 1;
    p.conf_P = p.conf_next = p.conf_count = p.conf_giveup = 0;
    p.conf_D2 = p.conf_zr = p.conf_zi = 0.0;
    p.SJ = fe_make(1.0, 0);
    p.jr = 1.0;
    p.ji = 0.0;
    p.invSJd = 1.0;
    p.swr = 1e300;
    p.swi = 1e300;
    p.sS = fe_make(1e300, 0);
    p.sm = -1;
    p.save_iter2 = 1;
    p.twin2r = 1;
    p.xv = 0;
}

// Scaled delta (S, w) returned by the cold helpers by value: references into the
// per-lane state would force that state out of registers into scratch memory.
struct SW {
    Fe S;
    double wr, wi;
    bool ok;
};

struct SJW {
    Fe SJ;
    double jr, ji, invSJd;
};

// feRescaledStepUnderflow (mandel_perturbation.cpp:2794-2835)
__host__ __device__ __noinline__ SW step_underflow(Fe Xr, Fe Xi, Fe dcr, Fe dci, Fe S, double wr, double wi) {
    const double sm = S.m, sm2 = sm * sm;
    const int se = S.e;
    int e = INT_MIN;
    const bool has_w = wr != 0.0 || wi != 0.0;
    if (has_w && Xr.m != 0.0) e = max(e, Xr.e + se);
    if (has_w && Xi.m != 0.0) e = max(e, Xi.e + se);
    if (has_w) e = max(e, se + se);
    if (dcr.m != 0.0) e = max(e, dcr.e);
    if (dci.m != 0.0) e = max(e, dci.e);
    if (e == INT_MIN) return SW{fe_make(1.0, 0), 0.0, 0.0, true};
    auto align = [e](double mm, int te) {
        int d = te - e;
        if (mm == 0.0 || d < -54) return 0.0;
        return mm * fe_pow2(d);
    };
    const double xs_r = align(Xr.m * sm, Xr.e + se);
    const double xs_i = align(Xi.m * sm, Xi.e + se);
    const double s2 = align(sm2, se + se);
    const double cr = align(dcr.m, dcr.e);
    const double ci = align(dci.m, dci.e);
    const double nr = 2.0 * (xs_r * wr - xs_i * wi) + s2 * (wr * wr - wi * wi) + cr;
    const double ni = 2.0 * (xs_r * wi + xs_i * wr) + 2.0 * s2 * wr * wi + ci;
    const double mag2 = nr * nr + ni * ni;
    const double mag = mag2 > 0.0 ? sqrt(mag2) : hypot(nr, ni);
    if (mag == 0.0) return SW{fe_make(1.0, 0), 0.0, 0.0, true};
    return SW{fe_renorm(mag, e), nr / mag, ni / mag, true};
}

// feDerivativeStepUnderflow (mandel_perturbation.cpp:2842-2876)
__host__ __device__ __noinline__ SJW deriv_underflow(Fe zr, Fe zi, Fe SJ, double jr, double ji) {
    double im = 0.5 / SJ.m;
    int ie = 1 - SJ.e;
    if (im >= 1.0) {
        im *= 0.5;
        ++ie;
    }
    Fe invSJ = fe_make(im, ie);
    int e = invSJ.e;
    if (zr.m != 0.0) e = max(e, zr.e);
    if (zi.m != 0.0) e = max(e, zi.e);
    auto align = [e](Fe x) {
        if (x.m == 0.0) return 0.0;
        int d = x.e - e;
        if (d < -54) return 0.0;
        return x.m * fe_pow2(d);
    };
    double azr = align(zr), azi = align(zi);
    double ar = 2.0 * (azr * jr - azi * ji) + align(invSJ);
    double ai = 2.0 * (azr * ji + azi * jr);
    double mag = sqrt(ar * ar + ai * ai);
    if (mag == 0.0) return SJW{fe_make(1.0, 0), 0.0, 0.0, 1.0};
    const Fe nSJ = fe_mul(SJ, fe_renorm(mag, e));
    return SJW{nSJ, ar / mag, ai / mag, 1.0 / fe_to_double(nSJ)};
}

// Overflow-safe FloatExp landing of a BLA skip (mandel_perturbation.cpp:2943-2957).
// Returns false when the landing escapes.
__host__ __device__ __noinline__ SW bla_land_fe(GPtr<const double2> zf, int land, double esc2, double4 bc, Fe S, double wr, double wi, double dr, double di) {
    Fe dzr = fe_mul_d(S, wr), dzi = fe_mul_d(S, wi);
    Fe dcrfe = fe_mul_d(S, dr), dcife = fe_mul_d(S, di);
    Fe nzr = fe_add(fe_sub(fe_mul_d(dzr, bc.x), fe_mul_d(dzi, bc.y)), fe_sub(fe_mul_d(dcrfe, bc.z), fe_mul_d(dcife, bc.w)));
    Fe nzi = fe_add(fe_add(fe_mul_d(dzr, bc.y), fe_mul_d(dzi, bc.x)), fe_add(fe_mul_d(dcrfe, bc.w), fe_mul_d(dcife, bc.z)));
    double2 X = zf[land];
    double fr = X.x + fe_to_double(nzr), fi = X.y + fe_to_double(nzi);
    if (fr * fr + fi * fi > esc2) return SW{S, wr, wi, false};
    Fe Snew = fe_sqrt(fe_add(fe_mul(nzr, nzr), fe_mul(nzi, nzi)));
    if (Snew.m == 0.0) return SW{fe_make(1.0, 0), 0.0, 0.0, true};
    return SW{Snew, fe_to_double(fe_div(nzr, Snew)), fe_to_double(fe_div(nzi, Snew)), true};
}

// tryBLAfe (mandel_perturbation.cpp:2883-2970).  s = m-1, reflen = mx_ref_it+1, or
// the period for a periodic reference (landings stay inside one period).
MX_HD int try_bla(const MxRefDev& R, GPtr<const int2> lev, int s, Fe& S, Fe S2, double& wr, double& wi, double dr, double di, int reflen, double esc2, int& kbla) {
    if (s < 1) return 0;
    const int maxp = R.nlev - 1;
    int startp = (s == 1) ? maxp : __builtin_ctz((unsigned)(s - 1));
    if (startp > maxp) startp = maxp;
    if (startp < R.minlevel) return 0;
    const Fe dz2 = fe_mul(S2, fe_from(wr * wr + wi * wi));
    if (R.rmax2 <= 0.0 || !fe_abs_less(dz2, fe_from(R.rmax2))) return 0;
    // the first level whose entry is valid decides: it is applied, or the search ends
    // with no skip (an escaping or unrepresentable landing does not fall back to a
    // shorter level), so the level search is a loop of its own
    int p = startp, k = -1;
    for (; p >= R.minlevel; --p) {
        const int2 ls = lev[p];
        const int i = (s - 1) >> p;
        if (i >= ls.y) continue;
        const int kk = ls.x + i;
        const double r2 = R.bla_r2[kk];
        if (r2 <= 0 || !fe_abs_less(dz2, fe_from(r2))) continue;
        if (s + (1 << p) >= reflen - 1) continue;
        k = kk;
        break;
    }
    if (k < 0) return 0;
    const int land = s + (1 << p);
    const double4 bc = R.bla_coef[k];
    const double nwr = (bc.x * wr - bc.y * wi) + (bc.z * dr - bc.w * di);
    const double nwi = (bc.x * wi + bc.y * wr) + (bc.z * di + bc.w * dr);
    const double BIG = 1e150;
    if (__builtin_isfinite(nwr) && __builtin_isfinite(nwi) && fabs(nwr) < BIG && fabs(nwi) < BIG) {
        const double2 X = R.zf[land];
        const double fr = X.x + fe_to_double(fe_mul_d(S, nwr));
        const double fi = X.y + fe_to_double(fe_mul_d(S, nwi));
        if (fr * fr + fi * fi > esc2) return 0;
        const double wm2 = nwr * nwr + nwi * nwi;
        if (wm2 == 0.0) {
            // a full reference rejects the unrepresentable landing; a periodic one
            // keeps the nonzero unit delta (blaReRescaleUnderflow, 2770-2781)
            if (!R.period) return 0;
            if (nwr == 0.0 && nwi == 0.0) {
                S = fe_make(1.0, 0);
                wr = wi = 0.0;
            } else {
                const Fe wmag = fe_sqrt(fe_add(fe_mul(fe_from(nwr), fe_from(nwr)), fe_mul(fe_from(nwi), fe_from(nwi))));
                S = fe_mul(S, wmag);
                const double inv = 1.0 / fe_to_double(wmag);
                wr = nwr * inv;
                wi = nwi * inv;
            }
        } else {
            const double wmag = sqrt(wm2);
            S = fe_mul(S, fe_from(wmag));
            const double inv = 1.0 / wmag;
            wr = nwr * inv;
            wi = nwi * inv;
        }
    } else {
        const SW r = bla_land_fe(R.zf, land, esc2, bc, S, wr, wi, dr, di);
        if (!r.ok) return 0;
        S = r.S;
        wr = r.wr;
        wi = r.wi;
    }
    kbla = k;
    return 1 << p;
}

// The common-case arithmetic of one step, shared by the full and the lean step
// (this file is built without FP contraction: identical expressions, identical bits).
// w' = (2X + s w) w + d, in the engine's order (mandel_perturbation.cpp:3446-3450)
MX_HD void step_w(const Px& p, double2 X, double& nwr, double& nwi) {
    const double fR = X.x + X.x + p.s * p.wr;
    const double fI = X.y + X.y + p.s * p.wi;
    nwr = fR * p.wr - fI * p.wi + p.dr;
    nwi = fR * p.wi + fI * p.wr + p.di;
}

// J' = 2 z J + 1/SJ (mandel_perturbation.cpp:3483-3488)
MX_HD void step_j(const Px& p, double zmr, double zmi, double& a, double& b) {
    a = 2.0 * (zmr * p.jr - zmi * p.ji) + p.invSJd;
    b = 2.0 * (zmr * p.ji + zmi * p.jr);
}

// One trip of the pixel loop.  Returns true when the sample is finished (value and
// n_esc written to v/n).  n: escape iteration, -1 = reached mxit, -2 = detected interior.
template <bool EDE, bool INTERIOR>
MX_HD bool trip_step(const MxRefDev& R, const MxTileDev& T, Px& p, double& v, int& n, Counters& c, bool& s_changed,
                     bool& j_changed);

template <bool EDE, bool INTERIOR>
MX_HD bool trip(const MxRefDev& R, GPtr<const int2> lev, const MxTileDev& T, Px& p, double& v, int& n, Counters& c) {
    const double ESC2 = T.esc2;
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
        if (p.m == p.sm && p.wr == p.swr && p.wi == p.swi && p.S.m == p.sS.m && p.S.e == p.sS.e) {
            n = -2;
            v = -2.0;
            return true;
        }
        if (p.iter - p.save_iter2 >= p.twin2r) {
            p.swr = p.wr;
            p.swi = p.wi;
            p.sS = p.S;
            p.sm = p.m;
            p.save_iter2 = p.iter;
            p.twin2r += p.twin2r;
        }
    }
    bool applied = false, j_changed = false;
    if ((T.flags & MX_F_BLA) && p.m >= 2) {
        int kbla = 0;
        const int skip = try_bla(R, lev, p.rm, p.S, p.S2, p.wr, p.wi, p.dr, p.di, per ? per : reflen, ESC2, kbla);
        if (skip > 0) {
            if constexpr (EDE) {
                // the applied entry's coefficients are read again rather than carried
                // out of the level search (fewer registers live across the landing)
                const double4 ab = R.bla_coef[kbla];
                Fe Jr = fe_mul_d(p.SJ, p.jr), Ji = fe_mul_d(p.SJ, p.ji);
                Fe nJr = fe_add(fe_sub(fe_mul_d(Jr, ab.x), fe_mul_d(Ji, ab.y)), fe_from(ab.z));
                Fe nJi = fe_add(fe_add(fe_mul_d(Jr, ab.y), fe_mul_d(Ji, ab.x)), fe_from(ab.w));
                p.SJ = fe_sqrt(fe_add(fe_mul(nJr, nJr), fe_mul(nJi, nJi)));
                if (p.SJ.m == 0.0) {
                    p.SJ = fe_make(1.0, 0);
                    p.jr = p.ji = 0.0;
                    p.invSJd = 1.0;
                } else {
                    p.jr = fe_to_double(fe_div(nJr, p.SJ));
                    p.ji = fe_to_double(fe_div(nJi, p.SJ));
                    j_changed = true;
                }
            }
            p.m += skip;
            p.iter += skip;
            p.rm += skip;
            if (per && p.rm >= per) p.rm -= per;
            c.applies++;
            c.skipped += skip;
            p.zsr = 1e30;
            p.zsi = 1e30;
            p.save_iter = p.iter;
            p.period_win = 1;
            p.conf_P = 0;
            p.conf_giveup = 0;
            p.xv = 0;
            applied = true;
        }
    }
    // a trip that changes S (or SJ) leaves what derives from it to be recomputed once, in
    // one place for every lane of the wave, after whichever branch changed it: s, S^2
    // and d = dc / S (1 / SJ in double)
    bool s_changed = applied;
    if (!applied && trip_step<EDE, INTERIOR>(R, T, p, v, n, c, s_changed, j_changed)) return true;
    if (s_changed) {
        p.s = fe_to_double(p.S);
        p.S2 = fe_mul(p.S, p.S);
        p.dr = fe_to_double(fe_div(p.dcr, p.S));
        p.di = fe_to_double(fe_div(p.dci, p.S));
    }
    if (EDE && j_changed) p.invSJd = 1.0 / fe_to_double(p.SJ);
    return false;
}

// The step part of a trip (no BLA skip taken at this index).  Returns true when the
// sample is finished; s_changed / j_changed are set when S / SJ changed (what derives
// from them is then left to the caller; s itself is kept current here where the step
// reads it).
template <bool EDE, bool INTERIOR>
MX_HD bool trip_step(const MxRefDev& R, const MxTileDev& T, Px& p, double& v, int& n, Counters& c, bool& s_changed,
                     bool& j_changed) {
    const double ESC2 = T.esc2;
    const int reflen = R.mx_ref_it + 1;
    const int per = R.period;
    const int mend = per ? INT_MAX : reflen;
    (void)reflen;
    // X_m and X_{m+1} come from the previous trip; X_{m+2} is fetched one trip ahead
    // (m <= mx_ref_it here, so rm+1 is in range; rm+2 only when m+1 <= mx_ref_it; a
    // periodic orbit carries copies of its first two entries after the last, so
    // rm+1 and rm+2 read the wrapped entries without a modulo).
    if (!p.xv) {
        p.Xc = p.m ? double2(R.zf[p.rm]) : make_double2(0.0, 0.0);
        p.Xn = R.zf[p.rm + 1];
    }
    const double2 Xpf = (p.rm + 2 < R.nent) ? double2(R.zf[p.rm + 2]) : make_double2(0.0, 0.0);
    double2 X = p.Xc;
    double zmr = 0.0, zmi = 0.0;
    bool d_underflow = false;
    Fe zmrfe = fe_make(0.0, 0), zmife = fe_make(0.0, 0);
    if constexpr (EDE) {
        zmr = X.x + p.s * p.wr;
        zmi = X.y + p.s * p.wi;
        d_underflow = zmr == 0.0 && zmi == 0.0 && p.invSJd == 0.0;
        if (d_underflow) {
            Fe Xrfe = fe_make(0.0, 0), Xife = fe_make(0.0, 0);
            if (p.m) {
                const double2 mm = R.zfe_m[p.rm];
                const int2 ee = R.zfe_e[p.rm];
                Xrfe = fe_make(mm.x, ee.x);
                Xife = fe_make(mm.y, ee.y);
            }
            zmrfe = fe_add(Xrfe, fe_mul_d(p.S, p.wr));
            zmife = fe_add(Xife, fe_mul_d(p.S, p.wi));
        }
    }
    double nwr, nwi;
    step_w(p, X, nwr, nwi);
    // step underflow: pixelRescaled recovers when both parts of w' are below 2^-1056
    // (mandel_perturbation.cpp:3459-3460), the no-BLA vector loop when |w'|^2 is 0
    // (3741), and then forms the escape test's z from the FloatExp orbit (3749-3752)
    const bool nobla_lane = (T.flags & MX_F_NOBLA_LANE) != 0;
    const bool su = nobla_lane ? nwr * nwr + nwi * nwi == 0.0 : fabs(nwr) < 0x1p-1056 && fabs(nwi) < 0x1p-1056;
    bool z_fe = false;
    if (su && (nwr != 0.0 || nwi != 0.0 || p.wr != 0.0 || p.wi != 0.0 || p.dcr.m != 0.0 || p.dci.m != 0.0)) {
        z_fe = nobla_lane;
        Fe Xrfe = fe_make(0.0, 0), Xife = fe_make(0.0, 0);
        if (p.m) {
            const double2 mm = R.zfe_m[p.rm];
            const int2 ee = R.zfe_e[p.rm];
            Xrfe = fe_make(mm.x, ee.x);
            Xife = fe_make(mm.y, ee.y);
        }
        const SW r = step_underflow(Xrfe, Xife, p.dcr, p.dci, p.S, p.wr, p.wi);
        p.S = r.S;
        p.wr = r.wr;
        p.wi = r.wi;
        p.s = fe_to_double(p.S);
        s_changed = true;
    } else {
        p.wr = nwr;
        p.wi = nwi;
    }
    ++p.m;
    ++p.iter;
    ++p.rm;
    if (per && p.rm == per) p.rm = 0;
    c.steps++;
    if constexpr (EDE) {
        if (d_underflow) {
            const SJW r = deriv_underflow(zmrfe, zmife, p.SJ, p.jr, p.ji);
            p.SJ = r.SJ;
            p.jr = r.jr;
            p.ji = r.ji;
            p.invSJd = r.invSJd;
        } else {
            double a, b;
            step_j(p, zmr, zmi, a, b);
            p.jr = a;
            p.ji = b;
            const double jm2 = p.jr * p.jr + p.ji * p.ji;
            if (jm2 > 1e16 || (jm2 > 0.0 && jm2 < 1e-16)) {
                const Fe jmag = fe_sqrt(fe_from(jm2));
                p.SJ = fe_mul(p.SJ, jmag);
                const double inv = 1.0 / fe_to_double(jmag);
                p.jr *= inv;
                p.ji *= inv;
                j_changed = true;
            }
        }
    }
    X = p.Xn;
    p.Xc = p.Xn;
    p.Xn = Xpf;
    p.xv = 1;
    const double zr = X.x + p.s * p.wr, zi = X.y + p.s * p.wi;
    double zrad = zr * zr + zi * zi;
    if (z_fe) {
        // X_{m} of the new m from the FloatExp orbit (the no-BLA loop's `current` index)
        const double2 mm = R.zfe_m[p.rm];
        const int2 ee = R.zfe_e[p.rm];
        const Fe ar = fe_add(fe_make(mm.x, ee.x), fe_mul_d(p.S, p.wr)), ai = fe_add(fe_make(mm.y, ee.y), fe_mul_d(p.S, p.wi));
        zrad = fe_to_double(fe_add(fe_mul(ar, ar), fe_mul(ai, ai)));
    }
    if (zrad > ESC2) {
        n = p.iter;
        if constexpr (EDE) {
            const double num = sqrt(zrad) * log(zrad);
            const Fe Jmag = fe_mul_d(p.SJ, sqrt(p.jr * p.jr + p.ji * p.ji));
            v = fe_to_double(fe_div(fe_from(num), fe_mul(T.dxfe, Jmag)));
        } else {
            const double LG2 = 0.69314718055994530942; // log(2.0), exactly the double libm returns
            v = (double)p.iter - log(log(zrad) / 2.0 / LG2) / LG2;
        }
        return true;
    }
    if (INTERIOR && (T.flags & MX_F_INTERIOR)) {
        if (p.conf_P > 0) {
            p.conf_D2 *= 4.0 * zrad;
            if (p.conf_D2 > 1e18) p.conf_D2 = 1e18;
            if (p.iter >= p.conf_next) {
                const double pr = zr - p.conf_zr, pi = zi - p.conf_zi;
                if (pr * pr + pi * pi < T.interior_eps2 * zrad && p.conf_D2 < 1.0) {
                    if (++p.conf_count >= T.interior_confirm) {
                        n = -2;
                        v = -2.0;
                        return true;
                    }
                    p.conf_zr = zr;
                    p.conf_zi = zi;
                    p.conf_D2 = 1;
                    p.conf_next = p.iter + p.conf_P;
                } else {
                    p.conf_P = 0;
                    ++p.conf_giveup;
                }
            }
        } else if ((p.iter & 15) == 0 && p.conf_giveup < 3) {
            const double pr = zr - p.zsr, pi = zi - p.zsi;
            if (pr * pr + pi * pi < T.interior_eps2 * zrad) {
                p.conf_P = p.iter - p.save_iter;
                if (p.conf_P < 1) p.conf_P = 1;
                p.conf_zr = zr;
                p.conf_zi = zi;
                p.conf_D2 = 1;
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
    const double wmag2 = p.wr * p.wr + p.wi * p.wi;
    if (wmag2 > 1e16 || (wmag2 < 1e-16 && wmag2 > 0.0)) {
        const Fe wmag = fe_sqrt(fe_from(wmag2));
        p.S = fe_mul(p.S, wmag);
        p.s = fe_to_double(p.S);
        const double inv = 1.0 / fe_to_double(wmag);
        p.wr *= inv;
        p.wi *= inv;
        s_changed = true;
    }
    const double dzmag2 = p.s * p.s * (p.wr * p.wr + p.wi * p.wi);
    if (zrad < 1e-8 || zrad < dzmag2 || p.m >= mend) {
        Fe Xmr = fe_make(0.0, 0), Xmi = fe_make(0.0, 0);
        if (p.m) {
            const double2 mm = R.zfe_m[p.rm];
            const int2 ee = R.zfe_e[p.rm];
            Xmr = fe_make(mm.x, ee.x);
            Xmi = fe_make(mm.y, ee.y);
        }
        const Fe Swr = fe_mul_d(p.S, p.wr), Swi = fe_mul_d(p.S, p.wi);
        const Fe zrfe = fe_add(Xmr, Swr), zife = fe_add(Xmi, Swi);
        const Fe zradfe = fe_add(fe_mul(zrfe, zrfe), fe_mul(zife, zife));
        const Fe dzfe = fe_add(fe_mul(Swr, Swr), fe_mul(Swi, Swi));
        if (p.m >= mend || fe_abs_less(zradfe, dzfe)) {
            const Fe Snew = fe_sqrt(zradfe);
            if (Snew.m == 0.0) {
                p.wr = p.wi = 0.0;
            } else {
                p.S = Snew;
                p.wr = fe_to_double(fe_div(zrfe, p.S));
                p.wi = fe_to_double(fe_div(zife, p.S));
                s_changed = true;
            }
            p.m = 0;
            p.rm = -1;
            p.xv = 0;
            c.rebases++;
        }
    }
    return false;
}

// Bit pattern of a non-negative double: monotone in the value (NaN and +inf above every
// finite value), so a range test is one subtraction and one compare.
MX_HD unsigned long long dbits(double x) { return __builtin_bit_cast(unsigned long long, x); }

// Lean steps: exactly the common case of trip(), for up to `budget` steps.  A lane
// stops, with its state untouched, at the first step whose full path could do more
// than the plain step; every condition is the full path's own test or a superset of
// it (the full path then decides), evaluated without branches and joined into one
// exit.  Before the step: the iteration limit, an active periodicity confirmation, an
// exact-repeat save or the orbit index back at the saved one, a BLA-aligned index
// whose |dz|^2 is not above the largest radius by exponent.  After it: |w'|^2
// outside [1e-16, 1e16] (renormalisation; 0 included), a step underflow, the same
// renormalisation test for the derivative, a derivative step at z = 0,
// |z|^2 outside [1e-8, escape
// radius^2], |z| < |dz|, the end of the reference, a near return on a periodicity
// check.  A wave executes a branch whenever any of its 64 lanes takes it, so these
// per-lane rare events would otherwise put the full path on nearly every wave step.
template <bool EDE, bool INTERIOR>
MX_HD void fast_steps(const MxRefDev& R, const MxTileDev& T, Px& p, Counters& c, int budget) {
    const double ESC2 = T.esc2;
    const int reflen = R.mx_ref_it + 1;
    const int amask = (1 << R.minlevel) - 1;
    const bool can_bla = (T.flags & MX_F_BLA) && R.nlev - 1 >= R.minlevel && R.rmax2 > 0.0;
    const int er = can_bla ? fe_from(R.rmax2).e : 0;
    const unsigned long long mlo = dbits(1e-16), mspan = dbits(1e16) - dbits(1e-16);
    const unsigned long long zlo = dbits(1e-8), zspan = dbits(ESC2) - dbits(1e-8);
    const bool int_rep = INTERIOR && (T.flags & MX_F_INT_REP);
#ifdef __HIP_DEVICE_COMPILE__
    // orbit entries through a buffer resource: a 32-bit offset per load, and indices
    // past the end read as 0
    const __amdgpu_buffer_rsrc_t zrs = __builtin_amdgcn_make_buffer_rsrc((void*)R.zf.p, 0, R.nent * 16, 0x00020000);
#endif
    double w2 = p.wr * p.wr + p.wi * p.wi;
    // one counter for the steps taken: orbit index, reference index and iteration
    // advance together, and the tests on them alone bound the trip count instead: the
    // loop never takes the step at which one would stop it
    const int m0 = p.m, rm0 = p.rm, it0 = p.iter;
    int lim = min(budget, T.mxit - it0); // iter >= mxit
    // m + 1 >= reflen after the step; periodic: rm + 1 == period (the wrap, taken by
    // the full step)
    lim = min(lim, R.period ? R.period - 1 - rm0 : reflen - 1 - m0);
    if (int_rep) {
        lim = min(lim, p.save_iter2 + p.twin2r - it0);  // iter - save_iter2 >= twin2r
        if (p.sm - m0 >= 0) lim = min(lim, p.sm - m0); // m == sm
    }
    if (INTERIOR && p.conf_P > 0) lim = 0; // an active periodicity confirmation
    // a BLA-aligned index (rm - 1 a multiple of 2^minlevel, m >= 2) whose |dz|^2 is not
    // above the largest radius by exponent: dz^2 = S^2 |w|^2 has exponent S2.e + e(|w|^2)
    // or one less, tested as e(|w|^2) <= er + 1 - S2.e, i.e. |w|^2 <= 2^(er + 1 - S2.e)
    // (equality included: a superset)
    const double wthr = can_bla ? ldexp(1.0, max(min(er + 1 - p.S2.e, 1024), -1075)) : 0.0;
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
            p.Xc = m0 + n ? double2(R.zf[rm]) : make_double2(0.0, 0.0);
            p.Xn = R.zf[rm + 1];
            p.xv = 1;
        }
        // the next step's entry (0 past the end, where this step stops anyway)
#ifdef __HIP_DEVICE_COMPILE__
        const double2 Xpf = __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(zrs, (rm + 2) * 16, 0, 0));
#else
        const double2 Xpf = rm + 2 < R.nent ? R.zf[rm + 2] : make_double2(0.0, 0.0);
#endif
        const double2 X = p.Xc;
        double nwr, nwi;
        step_w(p, X, nwr, nwi);
        // supersets of the full path's renormalisation test and of both step-underflow
        // tests: |w'|^2 outside [1e-16, 1e16] (0 included), or both parts of w' below
        // 2^-1056
        const double w2n = nwr * nwr + nwi * nwi;
        stop |= (int)(dbits(w2n) - mlo > mspan) | ((int)(fabs(nwr) < 0x1p-1056) & (int)(fabs(nwi) < 0x1p-1056));
        double ja = 0.0, jb = 0.0;
        if (EDE) {
            const double zmr = X.x + p.s * p.wr, zmi = X.y + p.s * p.wi;
            step_j(p, zmr, zmi, ja, jb);
            const double jm2 = ja * ja + jb * jb;
            stop |= ((int)(zmr == 0.0) & (int)(zmi == 0.0)) | (int)(dbits(jm2) - mlo > mspan);
        }
        const double2 Xn = p.Xn;
        const double zr = Xn.x + p.s * nwr, zi = Xn.y + p.s * nwi;
        const double zrad = zr * zr + zi * zi;
        const double dzmag2 = p.s * p.s * w2n;
        stop |= (int)(dbits(zrad) - zlo > zspan) | (int)(zrad < dzmag2);
        const int it1 = it0 + n + 1;
        int save = 0;
        if (INTERIOR) {
            const int chk = (int)((it1 & 15) == 0) & (int)chk_ok;
            const double pr = zr - p.zsr, pi = zi - p.zsi;
            stop |= chk & (int)(pr * pr + pi * pi < T.interior_eps2 * zrad);
            save = chk & (int)(it1 - p.save_iter >= p.period_win);
        }
        // commit by selects, so that the entry loaded above is consumed only by the
        // next step and its load overlaps this one; the cached entries rotate on every
        // step, and a lane that stops has them reloaded by the full step (xv = 0 below)
        const bool go = !stop;
        p.wr = go ? nwr : p.wr;
        p.wi = go ? nwi : p.wi;
        if (EDE) {
            p.jr = go ? ja : p.jr;
            p.ji = go ? jb : p.ji;
        }
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

// Map a queue item to a sample.  Items are 8x8 lane tiles laid row-major over the
// launch grid; returns false for the padding lanes of edge tiles.
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

template <bool EDE, bool INTERIOR>
__device__ __forceinline__ void run_job64(const MxRefDev& R, GPtr<const int2> lev, const MxTileDev& T) {
    const int lane = __lane_id();
    const unsigned long long lanemask_lt = (1ull << lane) - 1ull;
    int q = blockIdx.x % MX_NQUEUE;
    int exhausted = 0;
    bool alive = true, active = false;
    Px p;
    // each lane's work counters and output index live in LDS, not in registers: they
    // are touched once per trip and would otherwise hold seven VGPRs across the loop
    __shared__ Counters lds_counters[256];
    __shared__ int lds_sidx[256];
    Counters& c = lds_counters[threadIdx.x];
    int& sidx = lds_sidx[threadIdx.x];
    c = Counters{0, 0, 0, 0};
    sidx = 0;
    unsigned trips = 0;
    const unsigned ntiles = (unsigned)T.ntiles;
    while (true) {
        bool need = alive && !active;
        unsigned long long needmask = __ballot(need);
        while (needmask != 0ull && exhausted < MX_NQUEUE) {
            const unsigned t0 = ntiles * (unsigned)q / MX_NQUEUE, t1 = ntiles * (unsigned)(q + 1) / MX_NQUEUE;
            const unsigned nitems = (t1 - t0) * 64u;
            const int k = __popcll(needmask);
            const int leader = __ffsll((long long)needmask) - 1;
            unsigned base = 0;
            if (lane == leader) base = atomicAdd(&T.queue[q], (unsigned)k);
            base = __shfl(base, leader);
            if (need) {
                const unsigned it = base + (unsigned)__popcll(needmask & lanemask_lt);
                if (it < nitems) {
                    Fe dcr, dci;
                    if (take_sample(T, t0 * 64u + it, sidx, dcr, dci)) {
                        init_pixel(p, dcr, dci);
                        need = false;
                        active = true;
                    }
                }
            }
            if (base + (unsigned)k >= nitems) { // this queue is drained
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
            if (trip<EDE, INTERIOR>(R, lev, T, p, v, n, c)) {
                T.val[sidx] = v;
                T.nesc[sidx] = n;
                if (T.out) T.out[sidx] = n < 0 ? -2.0f : (float)v;
                active = false;
            } else {
                fast_steps<EDE, INTERIOR>(R, T, p, c, T.chunk);
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

template <bool EDE, bool INTERIOR>
__global__ __launch_bounds__(256) MX_WAVES_ATTR((EDE && INTERIOR) ? MX_FE64_WAVES_EDE_INTERIOR : MX_FE64_WAVES) void mx_fe64_kernel(
    const MxJob* __restrict__ jobs, int njobs) {
    for (int j = 0; j < njobs; ++j) run_job64<EDE, INTERIOR>(jobs[j].R, jobs[j].lev, jobs[j].T);
}

} // namespace

// Host-side launcher (called from the engine's C++); `jobs` is a device array, every
// job with the same EDE and interior-detection flags.
hipError_t mx_launch_fe64(const MxJob* jobs, int njobs, bool ede, bool interior, int blocks, int threads, hipStream_t stream) {
    if (ede && interior)
        hipLaunchKernelGGL((mx_fe64_kernel<true, true>), dim3(blocks), dim3(threads), 0, stream, jobs, njobs);
    else if (ede)
        hipLaunchKernelGGL((mx_fe64_kernel<true, false>), dim3(blocks), dim3(threads), 0, stream, jobs, njobs);
    else if (interior)
        hipLaunchKernelGGL((mx_fe64_kernel<false, true>), dim3(blocks), dim3(threads), 0, stream, jobs, njobs);
    else
        hipLaunchKernelGGL((mx_fe64_kernel<false, false>), dim3(blocks), dim3(threads), 0, stream, jobs, njobs);
    return hipGetLastError();
}
