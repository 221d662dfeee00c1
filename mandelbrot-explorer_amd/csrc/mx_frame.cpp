// Host-side frame preparation for the perturbation path: the serial part of
// Mandel::Compute (src/engine/mandel_perturbation.cpp:1388-1634) that stays on the
// CPU.  Grid set-up and the reference orbit run in GMP mpf with the engine's
// operations, so the double and FloatExp shadows handed to the GPU are the truncated
// values the engine would hold; the BLA table is built in FP64 like buildBLA.
#include <gmp.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "mx_fe.h"
#include "mx_fixed.h"
#include "mx_internal.h"

namespace mx {

static double now_s() { return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count(); }

// mpf_to_fe / mpf_get_ld (mandel_perturbation.cpp:69-74, float_math.cpp:15-19):
// truncation toward zero to a 53-bit mantissa.
static void mpf_shadow(mpf_srcptr x, double& m, long& e) {
    if (mpf_sgn(x) == 0) {
        m = 0.0;
        e = 0;
        return;
    }
    m = mpf_get_d_2exp(&e, x);
}

static double mpf_get_ld(mpf_srcptr x) {
    long e;
    double d = mpf_get_d_2exp(&e, x);
    return ldexp(d, (int)e);
}

struct Gmp {
    mpf_t v;
    explicit Gmp(mp_bitcnt_t prec) { mpf_init2(v, prec); }
    ~Gmp() { mpf_clear(v); }
    Gmp(const Gmp&) = delete;
    Gmp& operator=(const Gmp&) = delete;
    operator mpf_ptr() { return v; }
    operator mpf_srcptr() const { return v; }
};

// Reference orbit: createRef + calCoefficient (mandel_perturbation.cpp:4014-4187,
// 4189-4249) with the Karatsuba square re = (a+b)(a-b), im = 2ab: in mpf on the
// plain-double path, in fixed point on the floatexp path (the engine's choice there,
// 4084-4111, 4197-4227), where the shadows come from the fixed-point value.
// With period_len > 0 the orbit of a nucleus: entries 0 .. period_len-1 in mpf, no
// escape test (createPeriodicRef, mandel_perturbation.cpp:1108-1172, which builds it
// with calCoefficient's mpf branch).
void Frame::build_orbit(mpf_srcptr cre, mpf_srcptr cim, int period_len) {
    // one integer limb: the fixed-point form needs |z|^2 + |c| well below 2^64
    // (bigFixedRangeSafe, 4093-4095)
    if (!period_len && !dbl && escape_radius <= 0x1p31) {
        build_orbit_fixed(cre, cim);
        return;
    }
    const int mxit = period_len ? period_len - 1 : max_iter;
    zf.assign(2 * (size_t)(mxit + 1), 0.0);
    zfe_m.assign(2 * (size_t)(mxit + 1), 0.0);
    zfe_e.assign(2 * (size_t)(mxit + 1), 0);
    z_m3.assign((size_t)(mxit + 1), 0.0);
    const bool der = dbl && ede;
    dfr.assign(der ? (size_t)(mxit + 1) : 0, 0.0);
    dfi.assign(der ? (size_t)(mxit + 1) : 0, 0.0);
    if (der) dfr[0] = 1.0; // D_0 = 1
    Gmp zr(prec), zi(prec), t1(prec), t2(prec), nzr(prec), nzi(prec);
    mpf_set(ref_re.v, cre);
    mpf_set(ref_im.v, cim);
    mpf_set(zr, cre);
    mpf_set(zi, cim);
    auto store = [&](int i) {
        double m;
        long e;
        mpf_shadow(zr, m, e);
        zfe_m[2 * i] = m;
        zfe_e[2 * i] = (int32_t)e;
        zf[2 * i] = m == 0.0 ? 0.0 : ldexp(m, (int)e);
        mpf_shadow(zi, m, e);
        zfe_m[2 * i + 1] = m;
        zfe_e[2 * i + 1] = (int32_t)e;
        zf[2 * i + 1] = m == 0.0 ? 0.0 : ldexp(m, (int)e);
        // the glitch metric's |X|^2 (calCoefficient, mandel_perturbation.cpp:4249)
        z_m3[i] = i == 0 ? 0.0 : (zf[2 * i] * zf[2 * i] + zf[2 * i + 1] * zf[2 * i + 1]) / 1000000;
    };
    store(0);
    if (dbl) {
        // series approximation restarts with every reference; the order persists
        // (createRef, mandel_perturbation.cpp:4113-4124)
        for (int x = 0; x < MX_SA_N; ++x) sa_A[x] = sa_B[x] = MxC{0.0, 0.0};
        sa_A[0] = sa_delta;
        sa_flag = true;
        sa_it = 0;
    }
    ref_escaped = false;
    ref_it = mxit;
    const double R = (double)escape_radius;
    for (int i = 1; i <= mxit; ++i) {
        if ((i & 1023) == 0 && halted()) {
            ref_it = i - 1;
            break;
        }
        mpf_add(t1, zr, zi);
        mpf_sub(t2, zr, zi);
        mpf_mul(nzi, zr, zi);
        mpf_mul(nzr, t1, t2);
        mpf_mul_ui(nzi, nzi, 2);
        mpf_add(zr, nzr, cre);
        mpf_add(zi, nzi, cim);
        store(i);
        if (der) {
            // D_i = 2 D_{i-1} X_{i-1} + 1 in double (calCoefficient, 4256-4260)
            const double a = dfr[i - 1], b = dfi[i - 1], xr = zf[2 * (i - 1)], xi = zf[2 * (i - 1) + 1];
            dfr[i] = 2.0 * (a * xr - b * xi) + 1.0;
            dfi[i] = 2.0 * (a * xi + b * xr);
        }
        if (!period_len && std::hypot(zf[2 * i], zf[2 * i + 1]) > R) { // std::abs(Comp) > R
            ref_escaped = true;
            ref_it = i;
            break;
        }
        if (dbl && sa_flag) sa_step(i);
    }
    mpf_set(tail_re.v, zr);
    mpf_set(tail_im.v, zi);
    if (period_len) {
        // one period: buildBLA(period) reads indices up to period-1
        ref_escaped = false;
        ref_it = period_len;
        zf.resize(2 * (size_t)period_len);
        zfe_m.resize(2 * (size_t)period_len);
        zfe_e.resize(2 * (size_t)period_len);
        z_m3.resize((size_t)period_len);
        if (der) {
            dfr.resize((size_t)period_len);
            dfi.resize((size_t)period_len);
        }
        return;
    }
    zf.resize(2 * (size_t)(ref_it + 1));
    zfe_m.resize(2 * (size_t)(ref_it + 1));
    zfe_e.resize(2 * (size_t)(ref_it + 1));
    z_m3.resize((size_t)(ref_it + 1));
    if (der) {
        dfr.resize((size_t)(ref_it + 1));
        dfi.resize((size_t)(ref_it + 1));
    }
}

void Frame::build_orbit_fixed(mpf_srcptr cre, mpf_srcptr cim) {
    const int mxit = max_iter;
    zf.assign(2 * (size_t)(mxit + 1), 0.0);
    zfe_m.assign(2 * (size_t)(mxit + 1), 0.0);
    zfe_e.assign(2 * (size_t)(mxit + 1), 0);
    z_m3.assign((size_t)(mxit + 1), 0.0);
    mpf_set(ref_re.v, cre);
    mpf_set(ref_im.v, cim);
    // L = precision/64 + 2 limbs: fraction limbs for the precision, the integer limb
    // and a guard limb (createRef, mandel_perturbation.cpp:4099-4101)
    const int L = ((int)mpf_get_prec(ref_re.v) + 63) / 64 + 2;
    FixedArith fx(L, 9);
    Fixed cr, ci, zr, zi, t1, t2, ab, re, im;
    for (Fixed* v : {&cr, &ci, &zr, &zi, &t1, &t2, &ab, &re, &im}) fx.make(*v);
    fx.from_mpf(cr, cre);
    fx.from_mpf(ci, cim);
    fx.copy(zr, cr);
    fx.copy(zi, ci);
    // entry 0 from the mpf centre itself (4113-4119)
    {
        double m;
        long e;
        mpf_shadow(cre, m, e);
        zfe_m[0] = m;
        zfe_e[0] = (int32_t)e;
        zf[0] = m == 0.0 ? 0.0 : ldexp(m, (int)e);
        mpf_shadow(cim, m, e);
        zfe_m[1] = m;
        zfe_e[1] = (int32_t)e;
        zf[1] = m == 0.0 ? 0.0 : ldexp(m, (int)e);
    }
    ref_escaped = false;
    ref_it = mxit;
    const double R = (double)escape_radius;
    for (int i = 1; i <= mxit; ++i) {
        if ((i & 1023) == 0 && halted()) {
            ref_it = i - 1;
            break;
        }
        fx.add(t1, zr, zi);
        fx.sub(t2, zr, zi);
        fx.mul(ab, zr, zi);
        fx.mul(re, t1, t2);
        fx.add(im, ab, ab);
        fx.add(zr, re, cr);
        fx.add(zi, im, ci);
        const Fe fr = fx.to_fe(zr), fi = fx.to_fe(zi);
        zfe_m[2 * (size_t)i] = fr.m;
        zfe_e[2 * (size_t)i] = fr.e;
        zfe_m[2 * (size_t)i + 1] = fi.m;
        zfe_e[2 * (size_t)i + 1] = fi.e;
        const double dr = fe_to_double(fr), di = fe_to_double(fi);
        zf[2 * (size_t)i] = dr;
        zf[2 * (size_t)i + 1] = di;
        z_m3[i] = (dr * dr + di * di) / 1000000;
        if (std::hypot(dr, di) > R) {
            ref_escaped = true;
            ref_it = i;
            break;
        }
    }
    fx.to_mpf(tail_re.v, zr);
    fx.to_mpf(tail_im.v, zi);
    zf.resize(2 * (size_t)(ref_it + 1));
    zfe_m.resize(2 * (size_t)(ref_it + 1));
    zfe_e.resize(2 * (size_t)(ref_it + 1));
    z_m3.resize((size_t)(ref_it + 1));
}

// ---- periodic-nucleus reference (mandel_perturbation.cpp:76-224, 1072-1219) ----
//
// Three stages on mpf numbers at the frame's precision:
//   1. DiskOrbit: the disk of c-values around the view centre, iterated as a disk of
//      z-values (centre z_n, radius R_n), until it contains 0: the period p;
//   2. Newton on z_p(c) = 0 starting at the view centre, with z_p and dz_p/dc from the
//      orbit-derivative recurrence (split over three threads for long periods);
//   3. the island's size from the one-period orbit, in FloatExp.
// The sequence of mpf operations in each recurrence is the engine's, operand for
// operand: the nucleus's last bits, and with them the FloatExp entries of the
// periodic orbit next to its return to 0, depend on it.

namespace {

struct Scratch3 {
    Gmp a, b, c;
    explicit Scratch3(mp_bitcnt_t p) : a(p), b(p), c(p) {}
};

struct MpfPair {
    Gmp re, im;
    explicit MpfPair(mp_bitcnt_t p) : re(p), im(p) {}
    void zero() {
        mpf_set_ui(re, 0);
        mpf_set_ui(im, 0);
    }
};

// z <- z^2 + c as (zr + zi)(zr - zi) + cr, 2 zr zi + ci: two products
void sq_add(MpfPair& z, mpf_srcptr cr, mpf_srcptr ci, Scratch3& s) {
    mpf_add(s.a, z.re, z.im);
    mpf_sub(s.b, z.re, z.im);
    mpf_mul(s.c, z.re, z.im);
    mpf_mul(z.re, s.a, s.b);
    mpf_add(z.re, z.re, cr);
    mpf_mul_ui(z.im, s.c, 2);
    mpf_add(z.im, z.im, ci);
}

// out <- |z|^2
void norm2(mpf_ptr out, const MpfPair& z, Scratch3& s) {
    mpf_mul(s.a, z.re, z.re);
    mpf_mul(s.b, z.im, z.im);
    mpf_add(out, s.a, s.b);
}

// The two halves of d <- 2 z d + 1 (d = dz/dc), each reading z = (zr, zi) and the
// previous d = (dr, di): the real part, and the imaginary part (which may overwrite
// di itself once the real half has read it).
void deriv_real(mpf_ptr out, mpf_srcptr zr, mpf_srcptr zi, mpf_srcptr dr, mpf_srcptr di, Scratch3& s) {
    mpf_mul(s.a, zr, dr);
    mpf_mul(s.b, zi, di);
    mpf_sub(out, s.a, s.b);
    mpf_mul_ui(out, out, 2);
    mpf_add_ui(out, out, 1);
}
void deriv_imag(mpf_ptr out, mpf_srcptr zr, mpf_srcptr zi, mpf_srcptr dr, mpf_srcptr di, Scratch3& s) {
    mpf_mul(s.a, zr, di);
    mpf_mul(s.b, zi, dr);
    mpf_add(out, s.a, s.b);
    mpf_mul_ui(out, out, 2);
}

// The disk |c - centre| <= r under z -> z^2 + c, as a disk of z: the radius recurrence
// R' = (2|z| + R) R + r plus a roundoff allowance (|z|^2 + |c| + 1) 2^-(prec-32),
// with |z| taken in double (orbit values are O(1); R stays in mpf, far below double's
// range at depth) (g_ballPeriod, mandel_perturbation.cpp:115-161).
class DiskOrbit {
  public:
    DiskOrbit(mpf_srcptr cr, mpf_srcptr ci, mpf_srcptr r, mp_bitcnt_t prec)
        : _cr(cr), _ci(ci), _r(r), _z(prec), _R(prec), _R2(prec), _z2(prec), _next(prec), _ulp(prec), _t(prec), _s(prec) {
        _z.zero();
        mpf_set_ui(_R, 0);
        _cabs = std::hypot(mpf_get_d(cr), mpf_get_d(ci));
        mpf_set_ui(_ulp, 1);
        mpf_div_2exp(_ulp, _ulp, prec > 64 ? prec - 32 : prec / 2);
    }
    // one iteration: +1 once the disk contains 0, -1 once it has left the bailout
    // region, 0 otherwise
    int step() {
        const double za = std::hypot(mpf_get_d(_z.re), mpf_get_d(_z.im));
        mpf_set_d(_t, 2.0 * za);
        mpf_add(_next, _t, _R);
        mpf_mul(_next, _next, _R);
        mpf_add(_next, _next, _r);
        mpf_set_d(_t, za * za + _cabs + 1.0);
        mpf_mul(_t, _t, _ulp);
        mpf_add(_next, _next, _t);
        mpf_set(_R, _next);
        sq_add(_z, _cr, _ci, _s);
        norm2(_z2, _z, _s);
        mpf_mul(_R2, _R, _R);
        if (mpf_cmp(_z2, _R2) <= 0) return 1;
        return std::hypot(mpf_get_d(_z.re), mpf_get_d(_z.im)) > _cabs + 2.0 + mpf_get_d(_R) ? -1 : 0;
    }

  private:
    mpf_srcptr _cr, _ci, _r;
    MpfPair _z;
    Gmp _R, _R2, _z2, _next, _ulp, _t;
    Scratch3 _s;
    double _cabs;
};

// the first n <= maxp at which the disk contains 0; 0 when it escapes first, when
// maxp is reached, or when `halt` is raised
int disk_period(mpf_srcptr cr, mpf_srcptr ci, mpf_srcptr r, int maxp, mp_bitcnt_t prec, const volatile int* halt) {
    DiskOrbit disk(cr, ci, r, prec);
    for (int n = 1; n <= maxp; ++n) {
        if ((n & 1023) == 0 && halt && *halt) return 0;
        const int st = disk.step();
        if (st > 0) return n;
        if (st < 0) return 0;
    }
    return 0;
}

// z_p and d_p = dz_p/dc from z_0 = d_0 = 0, one thread per recurrence: this thread
// iterates z and publishes z_i in a ring; one worker computes the real part of d,
// another the imaginary part.  Step i of either half reads both halves of d_{i-1}
// (double-buffered by parity), so it starts once the other half has finished step
// i-1; a ring slot is reused once both halves have read it.  On `halt` the ring stops
// and the workers return early (the result is then not used).
class OrbitDerivative {
  public:
    OrbitDerivative(mpf_srcptr cr, mpf_srcptr ci, long period, mp_bitcnt_t prec, const volatile int* halt)
        : _cr(cr), _ci(ci), _period(period), _prec(prec), _halt(halt), _ring(RING), _d{MpfPair(prec), MpfPair(prec)} {
        for (auto& slot : _ring) {
            mpf_init2(slot.re, prec);
            mpf_init2(slot.im, prec);
        }
    }
    ~OrbitDerivative() {
        for (auto& slot : _ring) {
            mpf_clear(slot.re);
            mpf_clear(slot.im);
        }
    }
    // false when halted
    bool run(MpfPair& z, MpfPair& d) {
        _d[0].zero();
        std::thread wr([this] { half(0); }), wi([this] { half(1); });
        Scratch3 s(_prec);
        z.zero();
        bool ok = true;
        for (long i = 0; i < _period; ++i) {
            if ((i & 1023) == 0 && _halt && *_halt) {
                ok = false;
                _abort.store(true, std::memory_order_release);
                break;
            }
            while (std::min(_read[0].load(std::memory_order_acquire), _read[1].load(std::memory_order_acquire)) <= i - RING)
                std::this_thread::yield();
            mpf_set(_ring[i % RING].re, z.re);
            mpf_set(_ring[i % RING].im, z.im);
            _published.store(i + 1, std::memory_order_release);
            sq_add(z, _cr, _ci, s);
        }
        wr.join();
        wi.join();
        if (!ok) return false;
        mpf_set(d.re, _d[_period & 1].re);
        mpf_set(d.im, _d[_period & 1].im);
        return true;
    }

  private:
    static constexpr int RING = 1024;
    struct Slot {
        mpf_t re, im;
    };
    bool wait_until(const std::atomic<long>& a, long v) const {
        for (int spin = 0; a.load(std::memory_order_acquire) < v; ++spin) {
            if (_abort.load(std::memory_order_acquire)) return false;
            if (spin < 4096)
                __builtin_ia32_pause();
            else
                std::this_thread::yield();
        }
        return true;
    }
    void half(int which) {
        Scratch3 s(_prec);
        for (long i = 0; i < _period; ++i) {
            if (!wait_until(_published, i + 1) || !wait_until(_finished[1 - which], i)) return;
            const Slot& slot = _ring[i % RING];
            const MpfPair& prev = _d[i & 1];
            if (which == 0)
                deriv_real(_d[(i + 1) & 1].re, slot.re, slot.im, prev.re, prev.im, s);
            else
                deriv_imag(_d[(i + 1) & 1].im, slot.re, slot.im, prev.re, prev.im, s);
            _read[which].store(i + 1, std::memory_order_release);
            _finished[which].store(i + 1, std::memory_order_release);
        }
    }
    mpf_srcptr _cr, _ci;
    long _period;
    mp_bitcnt_t _prec;
    const volatile int* _halt;
    std::vector<Slot> _ring;
    MpfPair _d[2];
    std::atomic<long> _published{0}, _read[2] = {{0}, {0}}, _finished[2] = {{0}, {0}};
    std::atomic<bool> _abort{false};
};

// Newton's method for the period-p nucleus from the start value in c: the step
// c <- c - z_p / d_p; converged once |step|^2 <= (2^-(prec-48))^2 (g_newton,
// mandel_perturbation.cpp:163-218).  False on no convergence in `maxit` steps, a
// vanishing derivative, or `halt`.
bool newton_nucleus(MpfPair& c, long period, mp_bitcnt_t prec, int maxit, const volatile int* halt) {
    MpfPair z(prec), d(prec), step(prec);
    Gmp den(prec), tol(prec), step2(prec), dr_next(prec);
    Scratch3 s(prec);
    mpf_set_ui(tol, 1);
    mpf_div_2exp(tol, tol, prec > 96 ? prec - 48 : prec / 2);
    mpf_mul(tol, tol, tol);
    for (int it = 0; it < maxit; ++it) {
        if (halt && *halt) return false;
        z.zero();
        d.zero();
        if (period >= 256) {
            OrbitDerivative od(c.re, c.im, period, prec, halt);
            if (!od.run(z, d)) return false;
        } else {
            for (long i = 0; i < period; ++i) {
                deriv_real(dr_next, z.re, z.im, d.re, d.im, s);
                deriv_imag(d.im, z.re, z.im, d.re, d.im, s);
                mpf_set(d.re, dr_next);
                sq_add(z, c.re, c.im, s);
            }
        }
        norm2(den, d, s);
        if (mpf_sgn(den.v) == 0) return false;
        // z / d = z conj(d) / |d|^2
        mpf_mul(s.a, z.re, d.re);
        mpf_mul(s.b, z.im, d.im);
        mpf_add(s.c, s.a, s.b);
        mpf_div(step.re, s.c, den);
        mpf_mul(s.a, z.im, d.re);
        mpf_mul(s.b, z.re, d.im);
        mpf_sub(s.c, s.a, s.b);
        mpf_div(step.im, s.c, den);
        mpf_sub(c.re, c.re, step.re);
        mpf_sub(c.im, c.im, step.im);
        norm2(step2, step, s);
        if (mpf_cmp(step2, tol) <= 0) return true;
    }
    return false;
}

// complex FloatExp
struct FeZ {
    Fe re, im;
};
FeZ fz_mul(FeZ a, FeZ b) { return {fe_sub(fe_mul(a.re, b.re), fe_mul(a.im, b.im)), fe_add(fe_mul(a.re, b.im), fe_mul(a.im, b.re))}; }
Fe fz_norm2(FeZ a) { return fe_add(fe_mul(a.re, a.re), fe_mul(a.im, a.im)); }

} // namespace

// findNucleus (mandel_perturbation.cpp:1072-1106): the disk around the view centre of
// radius 2/scale; its period, refined by Newton to the nucleus, which must lie in the
// disk.  Returns the period, or 0 (no period, no convergence, outside the disk,
// cancelled).
int Frame::find_nucleus(mpf_ptr nre, mpf_ptr nim, mpf_srcptr scale, int maxp) {
    const mp_bitcnt_t p = mpf_get_prec(c0_re.v);
    MpfPair centre(p), c(p);
    Gmp r(p), dh(p), dist(p);
    mpf_set_ui(r, 2);
    mpf_div(r, r, scale);
    mpf_add(centre.re, c0_re.v, r);
    mpf_mul_ui(dh, r, H);
    mpf_div_ui(dh, dh, W);
    mpf_add(centre.im, c0_im.v, dh);
    const int period = disk_period(centre.re, centre.im, r, maxp, p, halt);
    if (period <= 0) return 0;
    mpf_set(c.re, centre.re);
    mpf_set(c.im, centre.im);
    if (!newton_nucleus(c, period, p, 200, halt)) return 0;
    mpf_set(nre, c.re);
    mpf_set(nim, c.im);
    // distance to the disk centre
    MpfPair off(p);
    mpf_sub(off.re, nre, centre.re);
    mpf_sub(off.im, nim, centre.im);
    mpf_mul(off.re, off.re, off.re);
    mpf_mul(off.im, off.im, off.im);
    mpf_add(dist, off.re, off.im);
    mpf_sqrt(dist, dist);
    return mpf_cmp(dist, r) <= 0 ? period : 0;
}

// nucleusSizeMag (mandel_perturbation.cpp:1196-1219): the island's linear size
// 1 / (|b| |l|^2) over the one-period orbit z_1 .. z_{p-1}, with the multiplier
// l = prod 2 z_k and b = 1 + sum conj(l_k) / |l_k|^2; "huge" without a period.
Fe Frame::nucleus_size() const {
    const Fe huge = fe_make(1.0, 1000000);
    if (period < 2) return huge;
    FeZ l{fe_make(1.0, 0), fe_make(0.0, 0)}, b{fe_make(1.0, 0), fe_make(0.0, 0)};
    for (int k = 1; k < period; ++k) {
        const size_t q = 2 * (size_t)(k - 1);
        const FeZ z{fe_make(zfe_m[q], zfe_e[q]), fe_make(zfe_m[q + 1], zfe_e[q + 1])};
        const FeZ zl = fz_mul(z, l);
        l = {fe_mul_d(zl.re, 2.0), fe_mul_d(zl.im, 2.0)};
        const Fe l2 = fz_norm2(l);
        if (l2.m == 0.0) continue;
        b = {fe_add(b.re, fe_div(l.re, l2)), fe_sub(b.im, fe_div(l.im, l2))};
    }
    const Fe lmag = fe_sqrt(fz_norm2(l)), bmag = fe_sqrt(fz_norm2(b));
    const Fe denom = fe_mul(bmag, fe_mul(lmag, lmag));
    return denom.m == 0.0 ? huge : fe_div(fe_make(1.0, 0), denom);
}

// The periodic branch of buildReferenceOrbit (mandel_perturbation.cpp:1238-1324): a
// nucleus in view whose island spans at least 4 pixels becomes the reference, its
// orbit one period long and read modulo the period; a sub-pixel island or none
// leaves the full reference to the caller (false).
bool Frame::try_periodic(mpf_srcptr scale) {
    const mp_bitcnt_t np = mpf_get_prec(c0_re.v);
    Gmp nre(np), nim(np), t(np);
    const int p = find_nucleus(nre, nim, scale, max_iter);
    if (!(p > 2 && p <= max_iter)) return false;
    // the nucleus's fractional pixel position (createPeriodicRef, 1121-1127)
    mpf_sub(t, nre, c0_re.v);
    mpf_div(t, t, dx.v);
    const double rx = mpf_get_d(t);
    mpf_sub(t, nim, c0_im.v);
    mpf_div(t, t, dy.v);
    const double ry = mpf_get_d(t);
    build_orbit(nre, nim, p);
    period = p;
    if (fe_abs_less(nucleus_size(), fe_mul_d(dxfe, 4.0))) {
        period = 0;
        subpixel = true;
        return false;
    }
    ref_virtual = true;
    ref_x = rx;
    ref_y = ry;
    cutoff_sensitive = false;
    return true;
}

// get_exp (float_math.cpp:23-29): floor(log2|x|), zero -> -2^30
static int sa_get_exp(double x) {
    if (x == 0.0) return -0x40000000;
    int e;
    (void)std::frexp(x, &e);
    return e - 1;
}

// SACheckMagnitude (mandel_perturbation.cpp:4307-4332): the first order after which
// every later coefficient is 2^80 below every earlier one, at least 5; -1 if none.
static int sa_check(const MxC* A) {
    int pre_mn[MX_SA_N], suf_mx[MX_SA_N];
    pre_mn[0] = std::min(sa_get_exp(A[0].r), sa_get_exp(A[0].i));
    suf_mx[MX_SA_N - 1] = std::max(sa_get_exp(A[MX_SA_N - 1].r), sa_get_exp(A[MX_SA_N - 1].i));
    for (int i = 1; i < MX_SA_N; ++i) pre_mn[i] = std::max(pre_mn[i - 1], std::min(sa_get_exp(A[i].r), sa_get_exp(A[i].i)));
    for (int i = MX_SA_N - 2; i >= 0; --i) suf_mx[i] = std::max(suf_mx[i + 1], std::max(sa_get_exp(A[i].r), sa_get_exp(A[i].i)));
    for (int i = 0; i < MX_SA_N - 10; ++i)
        if (pre_mn[i] - suf_mx[i + 1] >= 80) return std::max(i, 5);
    return -1;
}

// One step of the series coefficients after a non-escaping X_i (calCoefficient's
// series block, mandel_perturbation.cpp:4273-4302), with X = _zf[i-1]:
// N_j = 2X A_j + sum_{k<j/2} 2 A_k A_{j-k-1} (+ A_{j/2}^2 for odd j), N_0 += delta.
void Frame::sa_step(int i) {
    MxC N[MX_SA_N];
    const MxC two = {2.0, 0.0};
    const MxC z2 = mxc_mul(two, MxC{zf[2 * (i - 1)], zf[2 * (i - 1) + 1]});
    for (int j = 0; j < MX_SA_N; ++j) {
        N[j] = mxc_mul(z2, sa_A[j]);
        for (int k = 0; k < j / 2; ++k) N[j] = mxc_add(N[j], mxc_mul(mxc_mul(two, sa_A[k]), sa_A[j - k - 1]));
        if (j % 2) N[j] = mxc_add(N[j], mxc_mul(sa_A[j / 2], sa_A[j / 2]));
    }
    MxC NB[MX_SA_N];
    if (ede) {
        // B'_j = 2 (B_j X + A_j D + sum_{k<j} A_k B_{j-k-1}), D = _df[i-1] (4279-4285)
        const MxC X = {zf[2 * (i - 1)], zf[2 * (i - 1) + 1]}, D = {dfr[i - 1], dfi[i - 1]};
        for (int j = 0; j < MX_SA_N; ++j) {
            NB[j] = mxc_add(mxc_mul(sa_B[j], X), mxc_mul(sa_A[j], D));
            for (int k = 0; k < j; ++k) NB[j] = mxc_add(NB[j], mxc_mul(sa_A[k], sa_B[j - k - 1]));
            NB[j] = mxc_mul(NB[j], two);
        }
    }
    N[0] = mxc_add(N[0], sa_delta);
    const int order = sa_check(N);
    if (order < 0) {
        sa_flag = false;
        return;
    }
    sa_it = i;
    sa_order = order;
    for (int j = 0; j < MX_SA_N; ++j) sa_A[j] = N[j];
    if (ede)
        for (int j = 0; j < MX_SA_N; ++j) sa_B[j] = NB[j];
}

// Value of a (non-virtual) reference pixel from its own orbit
// (mandel_perturbation.cpp:4145-4160, 4179-4185): getEscapeTime, or the GMP point
// under EDE, -2 when the reference does not escape.  With `async` the GMP point runs
// on a host thread; ref_pixel_wait() joins it.
void Frame::ref_pixel(bool async) {
    ref_pixel_wait();
    if (!ref_escaped) {
        ref_pixel_n = -1;
        ref_pixel_value = -2.0;
    } else if (ede) {
        if (async) {
            ref_pixel_n = -1;
            ref_pixel_value = -2.0;
            ref_job = std::thread([this] { job_value = point_value(ref_re.v, ref_im.v, max_iter, true, job_n); });
            job_pending = true;
        } else {
            ref_pixel_value = point_value(ref_re.v, ref_im.v, max_iter, true, ref_pixel_n);
        }
    } else {
        Gmp t1(prec), t2(prec);
        mpf_mul(t1, tail_re.v, tail_re.v);
        mpf_mul(t2, tail_im.v, tail_im.v);
        mpf_add(t1, t1, t2);
        const double rad = mpf_get_d(t1);
        ref_pixel_n = ref_it + 1;
        ref_pixel_value = (ref_it + 1 - std::log(std::log(rad) / 2 / std::log(2)) / std::log(2));
    }
}

void Frame::ref_pixel_wait() {
    if (ref_job.joinable()) ref_job.join();
    if (job_pending) {
        ref_pixel_value = job_value;
        ref_pixel_n = job_n;
        job_pending = false;
    }
}

// refTailEscapes (mandel_perturbation.cpp:927-957)
bool Frame::tail_escapes(int extra) {
    Gmp zr(prec), zi(prec), t1(prec), t2(prec);
    mpf_set(zr, tail_re.v);
    mpf_set(zi, tail_im.v);
    bool esc = false;
    for (int k = 0; k < extra && !esc; ++k) {
        mpf_mul(t1, zr, zi);
        mpf_mul(zr, zr, zr);
        mpf_mul(t2, zi, zi);
        mpf_sub(zr, zr, t2);
        mpf_add(zr, zr, ref_re.v);
        mpf_mul_ui(zi, t1, 2);
        mpf_add(zi, zi, ref_im.v);
        mpf_mul(t1, zr, zr);
        mpf_mul(t2, zi, zi);
        mpf_add(t1, t1, t2);
        if (mpf_get_d(t1) > esc2) esc = true;
    }
    return esc;
}

// accuratePointCompute (mandel_perturbation.cpp:719-894) for the reference pixel's
// own value under a coloring method; smooth or pixel-normalised EDE, -2 interior.
// This is the loop as the engine runs it; point_value below gives the same result
// faster and falls back to it where it cannot decide.
double Frame::point_value_serial(mpf_srcptr c_re, mpf_srcptr c_im, int mxit, bool ede, int& n_out) {
    Gmp t1(prec), t2(prec), zr(prec), zi(prec), dr(prec), di(prec), er(prec), ei(prec), e1(prec), e2(prec);
    mpf_set(zr, c_re);
    mpf_set(zi, c_im);
    mpf_set_ui(dr, 2);
    mpf_set_ui(di, 0);
    mpf_set_ui(er, 1);
    mpf_set_ui(ei, 0);
    double res = -2.0;
    n_out = -1;
    for (int i = 1; i < mxit; ++i) {
        if ((i & 255) == 0 && halted()) { // the engine's cancellation point (mandel_perturbation.cpp:778-782)
            n_out = INT_MIN;
            return -10.0;
        }
        mpf_mul(t1, dr, zi);
        mpf_mul(dr, dr, zr);
        mpf_mul(t2, di, zi);
        mpf_mul(di, di, zr);
        mpf_sub(dr, dr, t2);
        mpf_mul_ui(dr, dr, 2);
        mpf_add(di, di, t1);
        mpf_mul_ui(di, di, 2);
        if (ede) {
            mpf_mul(e1, er, zi);
            mpf_mul(e2, ei, zr);
            mpf_add(e2, e1, e2);
            mpf_mul(e1, er, zr);
            mpf_mul(t2, ei, zi);
            mpf_sub(e1, e1, t2);
            mpf_mul_ui(er, e1, 2);
            mpf_add_ui(er, er, 1);
            mpf_mul_ui(ei, e2, 2);
        }
        mpf_mul(t1, zr, zi);
        mpf_mul(zr, zr, zr);
        mpf_mul(zi, zi, zi);
        mpf_sub(zr, zr, zi);
        mpf_add(zr, zr, c_re);
        mpf_mul_ui(zi, t1, 2);
        mpf_add(zi, zi, c_im);
        mpf_mul(t1, zr, zr);
        mpf_mul(t2, zi, zi);
        mpf_add(t1, t1, t2);
        double rad = mpf_get_d(t1);
        if (rad > esc2) {
            n_out = i + 1;
            if (ede) {
                mpf_mul(e1, er, er);
                mpf_mul(e2, ei, ei);
                mpf_add(e1, e1, e2);
                mpf_sqrt(e1, e1);
                mpf_abs(t1, dx.v);
                mpf_mul(e1, e1, t1);
                double num = std::isfinite(rad) && rad > 0.0 ? std::sqrt(rad) * std::log(rad) : 0.0;
                mpf_set_d(e2, num);
                if (mpf_sgn(e1.v) != 0) {
                    mpf_div(e1, e2, e1);
                    res = mpf_get_d(e1);
                } else
                    res = 0.0;
            } else {
                res = (i + 1 - std::log(std::log(rad) / 2 / std::log(2)) / std::log(2));
            }
            break;
        }
        mpf_mul(t1, dr, dr);
        mpf_mul(t2, di, di);
        mpf_add(t1, t1, t2);
        if (mpf_get_d(t1) < 0.000000001) {
            n_out = -2;
            break;
        }
    }
    return res;
}

namespace {

// e with 2^(e-1) <= |x| < 2^e for x != 0
long mpf_exp2(mpf_srcptr x) {
    long e;
    (void)mpf_get_d_2exp(&e, x);
    return e;
}

// z values from the orbit thread to the EDE derivative thread
struct ZRing {
    const int n;                      // slots: 2048, at most ~64 MB
    std::vector<mpf_t> zr, zi;        // slot k % n holds z_k
    std::atomic<long> count{0};       // z_0 .. z_{count-1} are in the ring
    std::atomic<long> used[2];        // z values each consumer is done with
    std::atomic<long> esc_at{LONG_MAX}, d_at{LONG_MAX};
    explicit ZRing(mp_bitcnt_t prec)
        : n((int)std::max<mp_bitcnt_t>(16, std::min<mp_bitcnt_t>(2048, ((mp_bitcnt_t)1 << 29) / (2 * prec + 256)))),
          zr(n), zi(n) {
        used[0].store(0);
        used[1].store(0);
        for (int k = 0; k < n; ++k) {
            mpf_init2(zr[k], prec);
            mpf_init2(zi[k], prec);
        }
    }
    ~ZRing() {
        for (int k = 0; k < n; ++k) {
            mpf_clear(zr[k]);
            mpf_clear(zi[k]);
        }
    }
};

} // namespace

// The serial loop's result from fewer multiplies, the orbit on this thread and the
// EDE derivative dc on two more (one per component) reading the z values through a
// ring:
//  - |z|^2 is formed only when a component can reach the escape radius
//    (|z|^2 < 2^(2 emax + 1) otherwise);
//  - d = 2 d z (the interior early-out |d|^2 < 1e-9) is not iterated in mpf: |d_i|^2 =
//    4^i prod |z_k|^2 exactly, and the mpf d differs from that by a relative
//    i * 2^-(precision - 4) at most, so a FloatExp product of the |z_k|^2 shadows
//    (relative error below 2e-15 per step) decides the test whenever it lies outside
//    1e-9 * (1 +- (1e-6 + 2e-15 i)); inside that window the serial loop runs instead.
double Frame::point_value(mpf_srcptr c_re, mpf_srcptr c_im, int mxit, bool ede, int& n_out) {
    ZRing Q(prec);
    const int RB = Q.n;
    int eth;
    (void)std::frexp(esc2, &eth); // esc2 >= 2^(eth-1)
    const long last = mxit - 1;   // iterations 1 .. mxit-1
    double rad_esc = 0.0;
    bool undecided = false, stopped_by_halt = false;

    // the EDE derivative dc, one thread per component: iteration i of either reads
    // both components of iteration i-1 (buffers by parity) and starts once the other
    // component has finished i-1, so neither overwrites a value still being read
    Gmp er0(prec), er1(prec), ei0(prec), ei1(prec);
    mpf_ptr ER[2] = {er0.v, er1.v}, EI[2] = {ei0.v, ei1.v};
    std::atomic<long> done_c[2];
    done_c[0].store(0);
    done_c[1].store(0);
    mpf_set_ui(ER[0], 1);
    mpf_set_ui(EI[0], 0);
    auto stopped = [&](long i) {
        return Q.esc_at.load(std::memory_order_acquire) < i || Q.d_at.load(std::memory_order_acquire) < i;
    };
    auto wait_for = [&](const std::atomic<long>& a, long v, long i) {
        for (int spin = 0; a.load(std::memory_order_acquire) < v; ++spin) {
            if (stopped(i)) return false;
            if (spin < 4096)
                __builtin_ia32_pause();
            else
                std::this_thread::yield();
        }
        return true;
    };
    auto eloop = [&](int which) {
        Gmp a(prec), b(prec);
        std::atomic<long>& mine = done_c[which];
        const std::atomic<long>& other = done_c[1 - which];
        for (long i = 1; i <= last; ++i) {
            // needs z_{i-1} and the other component of iteration i-1
            if (!wait_for(Q.count, i, i) || !wait_for(other, i - 1, i)) break;
            mpf_srcptr zr = Q.zr[(i - 1) % RB], zi = Q.zi[(i - 1) % RB];
            mpf_srcptr pr = ER[(i - 1) & 1], pi = EI[(i - 1) & 1];
            if (which == 0) {
                // re: 2 (er zr - ei zi) + 1
                mpf_mul(a, pr, zr);
                mpf_mul(b, pi, zi);
                Q.used[0].store(i, std::memory_order_release);
                mpf_sub(a, a, b);
                mpf_mul_ui(ER[i & 1], a, 2);
                mpf_add_ui(ER[i & 1], ER[i & 1], 1);
            } else {
                // im: 2 (er zi + ei zr)
                mpf_mul(a, pr, zi);
                mpf_mul(b, pi, zr);
                Q.used[1].store(i, std::memory_order_release);
                mpf_add(a, a, b);
                mpf_mul_ui(EI[i & 1], a, 2);
            }
            mine.store(i, std::memory_order_release);
            if (Q.esc_at.load(std::memory_order_acquire) == i) break;
        }
        Q.used[which].store(LONG_MAX, std::memory_order_release);
    };
    std::thread te0, te1;
    if (ede) {
        te0 = std::thread(eloop, 0);
        te1 = std::thread(eloop, 1);
    } else {
        Q.used[0].store(LONG_MAX);
        Q.used[1].store(LONG_MAX);
    }
    {
        Gmp t1(prec), t2(prec), zr(prec), zi(prec);
        mpf_set(zr, c_re);
        mpf_set(zi, c_im);
        if (ede) {
            mpf_set(Q.zr[0], zr);
            mpf_set(Q.zi[0], zi);
            Q.count.store(1, std::memory_order_release);
        }
        double dm = 4.0; // |d|^2 = dm * 2^de, d_0 = 2
        long de = 0;
        for (long i = 1; i <= last; ++i) {
            if ((i & 1023) == 0 && halted()) {
                stopped_by_halt = true;
                Q.d_at.store(i, std::memory_order_release); // stops the derivative threads
                break;
            }
            // |d_i|^2 = 4 |d_{i-1}|^2 |z_{i-1}|^2
            {
                const int sr = mpf_sgn(zr.v), si = mpf_sgn(zi.v);
                if (sr == 0 && si == 0) {
                    dm = 0.0;
                } else if (dm != 0.0) {
                    long xr = 0, xi = 0;
                    const double mr = sr ? mpf_get_d_2exp(&xr, zr) : 0.0, mi = si ? mpf_get_d_2exp(&xi, zi) : 0.0;
                    const long em = sr == 0 ? xi : si == 0 ? xr : std::max(xr, xi);
                    const double a = sr ? std::ldexp(mr * mr, (int)std::max(-2000L, 2 * (xr - em))) : 0.0;
                    const double b = si ? std::ldexp(mi * mi, (int)std::max(-2000L, 2 * (xi - em))) : 0.0;
                    int k;
                    dm = std::frexp(4.0 * dm * (a + b), &k);
                    de += 2 * em + k;
                }
            }
            mpf_mul(t1, zr, zi);
            mpf_mul(zr, zr, zr);
            mpf_mul(zi, zi, zi);
            mpf_sub(zr, zr, zi);
            mpf_add(zr, zr, c_re);
            mpf_mul_ui(zi, t1, 2);
            mpf_add(zi, zi, c_im);
            // |z|^2 < 2^(2 emax + 1) <= 2^(eth - 1) <= esc2: cannot escape
            const long ez = std::max(mpf_sgn(zr.v) ? mpf_exp2(zr) : LONG_MIN / 4, mpf_sgn(zi.v) ? mpf_exp2(zi) : LONG_MIN / 4);
            if (2 * ez + 1 > eth - 1) {
                mpf_mul(t1, zr, zr);
                mpf_mul(t2, zi, zi);
                mpf_add(t1, t1, t2);
                const double rad = mpf_get_d(t1);
                if (rad > esc2) {
                    rad_esc = rad;
                    Q.esc_at.store(i, std::memory_order_release);
                    break;
                }
            }
            // the d test of iteration i
            {
                const double w = 1e-6 + 2e-15 * (double)i;
                const double v = dm == 0.0 ? 0.0 : de < -1000 ? 0.0 : de > 1000 ? 1e300 : std::ldexp(dm, (int)de);
                if (v < 1e-9 * (1.0 - w)) {
                    Q.d_at.store(i, std::memory_order_release);
                    break;
                }
                if (v <= 1e-9 * (1.0 + w)) {
                    undecided = true;
                    Q.d_at.store(i, std::memory_order_release); // stops the derivative thread
                    break;
                }
            }
            if (ede && i < last) {
                while (std::min(Q.used[0].load(std::memory_order_acquire), Q.used[1].load(std::memory_order_acquire)) <= i - RB)
                    std::this_thread::yield();
                mpf_set(Q.zr[i % RB], zr);
                mpf_set(Q.zi[i % RB], zi);
                Q.count.store(i + 1, std::memory_order_release);
            }
        }
    }
    if (ede) {
        te0.join();
        te1.join();
    }
    if (stopped_by_halt) {
        n_out = INT_MIN;
        return -10.0;
    }
    if (undecided) return point_value_serial(c_re, c_im, mxit, ede, n_out);

    const long ie = Q.esc_at.load(), id = Q.d_at.load();
    double res = -2.0;
    n_out = -1;
    if (ie != LONG_MAX) {
        const double rad = rad_esc;
        n_out = (int)ie + 1;
        if (ede) {
            Gmp t1(prec), e1(prec), e2(prec);
            mpf_srcptr er = ER[ie & 1], ei = EI[ie & 1];
            mpf_mul(e1, er, er);
            mpf_mul(e2, ei, ei);
            mpf_add(e1, e1, e2);
            mpf_sqrt(e1, e1);
            mpf_abs(t1, dx.v);
            mpf_mul(e1, e1, t1);
            double num = std::isfinite(rad) && rad > 0.0 ? std::sqrt(rad) * std::log(rad) : 0.0;
            mpf_set_d(e2, num);
            if (mpf_sgn(e1.v) != 0) {
                mpf_div(e1, e2, e1);
                res = mpf_get_d(e1);
            } else
                res = 0.0;
        } else {
            res = ((double)ie + 1 - std::log(std::log(rad) / 2 / std::log(2)) / std::log(2));
        }
    } else if (id != LONG_MAX) {
        n_out = -2;
    }
    return res;
}

// A view-centre reference makes the frame cutoff-sensitive when it escapes within 16
// of max_iter (mandel_perturbation.cpp:4144) or, bounded, when the brute force sees
// it escape within max_iter + 64 (4171-4178); the second test runs the whole orbit
// again, so only when the caller applies the fallback (check_cutoff).
bool Frame::virtual_ref_sensitive() {
    if (ref_escaped) return ref_it >= max_iter - 16;
    if (!check_cutoff || halted()) return false;
    int n = 0;
    return point_value(ref_re.v, ref_im.v, max_iter + 64, false, n) >= 0.0 && n > 0;
}

// accuratePixelCompute (mandel_perturbation.cpp:896-918): the brute-force value of
// cell (u, v) of pixel (i, j) at the frame's precision, c formed in mpf as the engine
// does (c0 + d k, then the signed sub-cell offset d |u| / sub).
double Frame::pixel_value(int i, int j, int u, int v, int& n_out) {
    Gmp cre(prec), cim(prec), off(prec);
    mpf_set(cre, c0_re.v);
    mpf_set(cim, c0_im.v);
    mpf_mul_ui(off, dx.v, (unsigned long)j);
    mpf_add(cre, cre, off);
    mpf_mul_ui(off, dx.v, (unsigned long)std::abs(v));
    if (v < 0) mpf_neg(off, off);
    mpf_div_ui(off, off, (unsigned long)sub);
    mpf_add(cre, cre, off);
    mpf_mul_ui(off, dy.v, (unsigned long)i);
    mpf_add(cim, cim, off);
    mpf_mul_ui(off, dy.v, (unsigned long)std::abs(u));
    if (u < 0) mpf_neg(off, off);
    mpf_div_ui(off, off, (unsigned long)sub);
    mpf_add(cim, cim, off);
    return point_value_serial(cre, cim, max_iter, ede, n_out);
}

// buildBLA (mandel_perturbation.cpp:2613-2709) without the EDE derivative couplings,
// which the rescaled path does not read (it carries J through A and B).
void Frame::build_bla(double dcmax_in) {
    lev_off.clear();
    lev_size.clear();
    bla_r2.clear();
    bla_coef.clear();
    rmax2 = 0.0;
    dcmax = dcmax_in;
    /* buildBLA(ref_it): level 0 covers reference indices 1..ref_it-1 (MP.cpp:1632, 2618) */
    const int reflen = ref_it;
    if (reflen < 3) return;
    const double eps = std::ldexp(1.0, -53);
    int64_t n = reflen - 1, tot = 0;
    for (;;) {
        lev_off.push_back((int32_t)tot);
        lev_size.push_back((int32_t)n);
        tot += n;
        if (n <= 1) break;
        n /= 2;
    }
    bla_r2.assign((size_t)tot, 0.0);
    bla_coef.assign(4 * (size_t)tot, 0.0);
    // the double path under EDE also carries the derivative delta through a skip:
    // level 0 C = 2 D_s, E = 0 (mandel_perturbation.cpp:2649-2651)
    const bool der = dbl && ede;
    bla_der.assign(der ? 4 * (size_t)tot : 0, 0.0);
    for (int s = 1; s < reflen; ++s) {
        const double zr = zf[2 * s], zi = zf[2 * s + 1];
        const double Zmag = std::sqrt(zr * zr + zi * zi);
        const double Amag = 2.0 * Zmag;
        const double R = eps * (Zmag - dcmax) / (Amag + 1.0);
        const double r2 = R > 0 ? R * R : 0.0;
        if (r2 > rmax2) rmax2 = r2;
        const size_t k = (size_t)(s - 1);
        bla_r2[k] = r2;
        bla_coef[4 * k + 0] = 2.0 * zr;
        bla_coef[4 * k + 1] = 2.0 * zi;
        bla_coef[4 * k + 2] = 1.0;
        bla_coef[4 * k + 3] = 0.0;
        if (der) {
            bla_der[4 * k + 0] = 2.0 * dfr[s];
            bla_der[4 * k + 1] = 2.0 * dfi[s];
        }
    }
    for (size_t p = 1; p < lev_off.size(); ++p) {
        const size_t po = (size_t)lev_off[p - 1], no = (size_t)lev_off[p];
        for (size_t i = 0; i < (size_t)lev_size[p]; ++i) {
            const size_t x = po + 2 * i, y = x + 1, z = no + i;
            const double* xc = &bla_coef[4 * x];
            const double* yc = &bla_coef[4 * y];
            double* zc = &bla_coef[4 * z];
            zc[0] = yc[0] * xc[0] - yc[1] * xc[1];
            zc[1] = yc[0] * xc[1] + yc[1] * xc[0];
            zc[2] = yc[0] * xc[2] - yc[1] * xc[3] + yc[2];
            zc[3] = yc[0] * xc[3] + yc[1] * xc[2] + yc[3];
            if (der) {
                // C_z = C_y A_x + A_y C_x, E_z = C_y B_x + A_y E_x + E_y (2679-2685)
                const double* xd = &bla_der[4 * x];
                const double* yd = &bla_der[4 * y];
                double* zd = &bla_der[4 * z];
                zd[0] = (yd[0] * xc[0] - yd[1] * xc[1]) + (yc[0] * xd[0] - yc[1] * xd[1]);
                zd[1] = (yd[0] * xc[1] + yd[1] * xc[0]) + (yc[0] * xd[1] + yc[1] * xd[0]);
                zd[2] = (yd[0] * xc[2] - yd[1] * xc[3]) + (yc[0] * xd[2] - yc[1] * xd[3]) + yd[2];
                zd[3] = (yd[0] * xc[3] + yd[1] * xc[2]) + (yc[0] * xd[3] + yc[1] * xd[2]) + yd[3];
            }
            const double Rx = std::sqrt(bla_r2[x]), Ry = std::sqrt(bla_r2[y]);
            const double Axmag = std::sqrt(xc[0] * xc[0] + xc[1] * xc[1]);
            const double Bxmag = std::sqrt(xc[2] * xc[2] + xc[3] * xc[3]);
            const double cand = Axmag > 0 ? (Ry - Bxmag * dcmax) / Axmag : 0.0;
            const double R = std::min(Rx, cand);
            bla_r2[z] = R > 0 ? R * R : 0.0;
        }
    }
}

static bool parse_decimal(mpf_ptr x, const char* s) {
    if (!s || !*s) return false;
    return mpf_set_str(x, s, 10) == 0;
}

int Frame::prepare(const mx_request& r, std::string& err) {
    const double t_start = now_s();
    const int full_h = r.full_height > 0 ? r.full_height : r.height;
    if (!r.center_re || !r.center_im || !r.scale || r.width < 2 || r.height < 1 || full_h < 2 || r.row_base < 0 ||
        r.row_base > full_h - r.height || r.sub < 1 || r.max_iterations < 1) {
        err = "invalid request";
        return MX_E_INVALID;
    }
    // sub samples per pixel side: the output is the (H sub) x (W sub) buffer of
    // getIndex (mandel_perturbation.cpp:4334-4340), pixel (i, j) at its centre cell; an
    // even sub would place the adaptive sub-samples of the last row and column past
    // the buffer
    if (r.sub > 1 && (r.sub % 2 == 0 || r.sub > 15)) {
        err = "sub must be odd and at most 15";
        return MX_E_INVALID;
    }
    sub = r.sub;
    ss = sub > 1 && (r.coloring_method & MX_SUPER_SAMPLING) != 0;
    if ((r.coloring_method & ~(MX_EXTERIOR_DIST_EST | MX_SUPER_SAMPLING)) != 0) {
        err = "coloring method not on this path (smooth and exterior DE only)";
        return MX_E_UNSUPPORTED;
    }
    W = r.width;
    H = r.height;
    max_iter = r.max_iterations;
    ede = (r.coloring_method & MX_EXTERIOR_DIST_EST) != 0;
    escape_radius = r.escape_radius > 0 ? r.escape_radius : 1e8f;
    const float R2f = escape_radius * escape_radius; // Mandel::escapeRadiusSquared (float product)
    esc2 = (double)R2f;
    const std::string sc = r.scale;
    prec = r.precision_bits > 0 ? (mp_bitcnt_t)r.precision_bits : (mp_bitcnt_t)((int)(sc.size() * std::log(10) / std::log(2)) + 64);
    if (prec > (1u << 24)) {
        err = "precision too large";
        return MX_E_INVALID;
    }
    init_mpf();
    Gmp cre(prec), cim(prec), scale(prec);
    if (!parse_decimal(cre, r.center_re) || !parse_decimal(cim, r.center_im) || !parse_decimal(scale, r.scale) || mpf_sgn(scale.v) <= 0) {
        err = "could not parse centre/scale";
        return MX_E_INVALID;
    }
    // path selection (mandel_perturbation.cpp:1434-1435, 1569-1576): the shallow direct
    // path at scale <= 1e6, else perturbation in double up to 1e280, else floatexp
    shallow = mpf_cmp_d(scale, 1e6) <= 0;
    dbl = !shallow && mpf_cmp_d(scale, 1e280) <= 0 && !(r.options & MX_OPT_FLOATEXP);
    if (dbl && (r.options & MX_OPT_PERIODIC) && !(r.options & MX_OPT_NO_PERIODIC) && r.use_bla != 0) {
        err = "the periodic (nucleus) reference on the plain-double path (scale <= 1e280) is not on this path "
              "(MX_OPT_FLOATEXP runs such frames on the floatexp path)";
        return MX_E_UNSUPPORTED;
    }
    const bool view_ref = (r.options & MX_OPT_VIEW_CENTER_REF) != 0;
    // depth guard for the int32 exponent of the device FloatExp: |log2 scale| < 2^26
    {
        long se;
        (void)mpf_get_d_2exp(&se, scale);
        if (se > (1L << 26)) {
            err = "scale out of range";
            return MX_E_INVALID;
        }
    }
    // grid (mandel_perturbation.cpp:1403-1424)
    {
        Gmp dw(prec), dh(prec);
        mpf_set_ui(dw, 2);
        mpf_div(dw, dw, scale);
        mpf_set(dh, dw);
        mpf_div_ui(dh, dh, W);
        mpf_mul_ui(dh, dh, full_h);
        mpf_sub(c0_re.v, cre, dw);
        mpf_sub(c0_im.v, cim, dh);
        mpf_div_ui(dx.v, dw, W - 1);
        mpf_mul_ui(dx.v, dx.v, 2);
        mpf_div_ui(dy.v, dh, full_h - 1);
        mpf_mul_ui(dy.v, dy.v, 2);
        if (r.row_base > 0) {
            Gmp off(prec);
            mpf_mul_ui(off, dy.v, r.row_base);
            mpf_add(c0_im.v, c0_im.v, off);
        }
    }
    if (shallow) {
        // the method-0 loops take c0, dx, dy as doubles (mandel_perturbation.cpp:1445-1448)
        sh_c0r = mpf_get_ld(c0_re.v);
        sh_c0i = mpf_get_ld(c0_im.v);
        sh_dx = mpf_get_ld(dx.v);
        sh_dy = mpf_get_ld(dy.v);
        interior = false;
        prepare_seconds = now_s() - t_start;
        return MX_OK;
    }
    double m;
    long e;
    mpf_shadow(dx.v, m, e);
    dxfe = fe_make(m, (int)e);
    mpf_shadow(dy.v, m, e);
    dyfe = fe_make(m, (int)e);

    use_bla = r.use_bla != 0;
    {
        const int ms = r.bla_minskip > 0 ? r.bla_minskip : 8;
        bla_minlevel = 0;
        while ((1 << (bla_minlevel + 1)) <= ms) ++bla_minlevel;
    }
    interior = r.interior != 0;
    escalated = false;
    if (dbl) {
        int rc = prepare_dbl(scale, err);
        if (rc == MX_OK && halted()) {
            err = "cancelled";
            rc = MX_E_CANCELLED;
        }
        if (rc != MX_OK || !escalated) {
            prepare_seconds = now_s() - t_start;
            return rc;
        }
        // the double path's reference passes too close to zero: the frame runs on the
        // floatexp path from the exact view centre (mandel_perturbation.cpp:1344-1354)
        dbl = false;
    }

    // reference: the centre pixel, then the exact view centre when the pixel
    // reference is bounded-but-sensitive (mandel_perturbation.cpp:1326-1373)
    const double t_ref = now_s();
    period = 0;
    subpixel = false;
    // a nucleus in view on the deep floatexp path with BLA (buildReferenceOrbit,
    // mandel_perturbation.cpp:1238-1250): a resolved one becomes the reference
    const bool periodic_want = (r.options & MX_OPT_PERIODIC) || mpf_cmp_d(scale, 1e280) > 0;
    if (!escalated && !view_ref && use_bla && !(r.options & MX_OPT_NO_PERIODIC) && periodic_want && try_periodic(scale)) {
        ref_seconds = now_s() - t_ref;
    } else {
        Gmp pre(prec), pim(prec), t(prec);
        const int pi = H / 2, pj = W / 2;
        mpf_mul_ui(t, dx.v, pj);
        mpf_add(pre, c0_re.v, t);
        mpf_mul_ui(t, dy.v, pi);
        mpf_add(pim, c0_im.v, t);
        ref_virtual = false;
        ref_x = pj;
        ref_y = pi;
        if (escalated || view_ref) {
            mpf_mul_ui(t, dx.v, W - 1);
            mpf_div_ui(t, t, 2);
            mpf_add(pre, c0_re.v, t);
            mpf_mul_ui(t, dy.v, H - 1);
            mpf_div_ui(t, t, 2);
            mpf_add(pim, c0_im.v, t);
            ref_virtual = true;
            ref_x = (W - 1) * 0.5;
            ref_y = (H - 1) * 0.5;
            build_orbit(pre, pim);
            cutoff_sensitive = virtual_ref_sensitive();
        } else {
            build_orbit(pre, pim);
        }
        if (!escalated && !view_ref && (!ref_escaped || ref_it >= max_iter - 16)) {
            const bool sensitive = ref_escaped || tail_escapes(80);
            if (sensitive) {
                mpf_mul_ui(t, dx.v, W - 1);
                mpf_div_ui(t, t, 2);
                mpf_add(pre, c0_re.v, t);
                mpf_mul_ui(t, dy.v, H - 1);
                mpf_div_ui(t, t, 2);
                mpf_add(pim, c0_im.v, t);
                ref_virtual = true;
                ref_x = (W - 1) * 0.5;
                ref_y = (H - 1) * 0.5;
                build_orbit(pre, pim);
                cutoff_sensitive = virtual_ref_sensitive();
            }
        }
        // value of the reference pixel itself (mandel_perturbation.cpp:4145-4160, 4179-4185)
        if (!ref_virtual) ref_pixel(true);
        ref_seconds = now_s() - t_ref;
    }
    if (halted()) {
        err = "cancelled";
        return MX_E_CANCELLED;
    }
    // a sub-pixel nucleus: BLA off for the frame (Compute, mandel_perturbation.cpp:1624-1629)
    if (subpixel) use_bla = false;

    const double t_bla = now_s();
    if (use_bla) {
        // dcmax from the farthest image corner (mandel_perturbation.cpp:2626-2629)
        const double dxf = std::fabs(mpf_get_ld(dx.v)), dyf = std::fabs(mpf_get_ld(dy.v));
        const double mxx = std::max(ref_x, (double)(W - 1) - ref_x) + 0.5;
        const double mxy = std::max(ref_y, (double)(H - 1) - ref_y) + 0.5;
        build_bla(std::sqrt(dxf * dxf * mxx * mxx + dyf * dyf * mxy * mxy));
    } else {
        build_bla(0.0);
        lev_off.clear();
        lev_size.clear();
        bla_r2.clear();
        bla_coef.clear();
        rmax2 = 0.0;
    }
    bla_seconds = now_s() - t_bla;

    // per-column / per-row dc (mandel_perturbation.cpp:2160-2163)
    dcol_m.resize(W);
    dcol_e.resize(W);
    drow_m.resize(H);
    drow_e.resize(H);
    for (int j = 0; j < W; ++j) {
        Fe f = fe_mul_d(dxfe, (double)j - ref_x);
        dcol_m[j] = f.m;
        dcol_e[j] = f.e;
    }
    for (int i = 0; i < H; ++i) {
        Fe f = fe_mul_d(dyfe, (double)i - ref_y);
        drow_m[i] = f.m;
        drow_e[i] = f.e;
    }
    prepare_seconds = now_s() - t_start;
    return MX_OK;
}

// The plain-double path's first reference: the centre pixel (h/2, w/2)
// (mandel_perturbation.cpp:1326-1328) with its series approximation over the frame's
// half extent _SA_delta = (dx*w/2, dy*h/2) (4114-4119), and the content-aware
// escalation test of scales above 1e150 (1336-1355).
int Frame::prepare_dbl(mpf_srcptr scale, std::string& err) {
    {
        Gmp t(prec);
        mpf_mul_ui(t, dx.v, W);
        mpf_div_ui(t, t, 2);
        sa_delta.r = mpf_get_ld(t);
        mpf_mul_ui(t, dy.v, H);
        mpf_div_ui(t, t, 2);
        sa_delta.i = mpf_get_ld(t);
    }
    sa_order = 0;
    n_refs = 0;
    set_pixel_reference(H / 2, W / 2);
    if (mpf_cmp_d(scale, 1e150) > 0) {
        double refmin2 = 1e300;
        bool underflowed = false;
        for (int q = 1; q <= ref_it; ++q) {
            const double m2 = z_m3[q] * 1e6;
            if (m2 <= 0.0) {
                underflowed = true;
                break;
            }
            if (m2 < refmin2) refmin2 = m2;
        }
        const double dcmax_c = 2.0 / mpf_get_d(scale);
        if (underflowed || std::sqrt(refmin2) * dcmax_c < 1e-300) escalated = true;
    }
    return MX_OK;
}

void Frame::set_pixel_reference(int pi, int pj, int u, int v, bool with_bla) {
    const double t_ref = now_s();
    Gmp pre(prec), pim(prec), t(prec);
    // c0 + dx pj + (+-dx |v|) / sub, c0 + dy pi + (+-dy |u|) / sub (mandel_perturbation.cpp:4057-4074)
    mpf_mul_ui(t, dx.v, pj);
    mpf_add(pre, c0_re.v, t);
    mpf_mul_ui(t, dx.v, (unsigned long)std::abs(v));
    if (v < 0) mpf_neg(t, t);
    mpf_div_ui(t, t, (unsigned long)sub);
    mpf_add(pre, pre, t);
    mpf_mul_ui(t, dy.v, pi);
    mpf_add(pim, c0_im.v, t);
    mpf_mul_ui(t, dy.v, (unsigned long)std::abs(u));
    if (u < 0) mpf_neg(t, t);
    mpf_div_ui(t, t, (unsigned long)sub);
    mpf_add(pim, pim, t);
    ref_virtual = false;
    ref_i = pi;
    ref_j = pj;
    ref_u = u;
    ref_v = v;
    ref_x = pj + (double)v / sub;
    ref_y = pi + (double)u / sub;
    build_orbit(pre, pim);
    ref_pixel(false);
    ++n_refs;
    ref_seconds += now_s() - t_ref;
    const double t_bla = now_s();
    if (use_bla && with_bla) {
        const double dxf = std::fabs(mpf_get_ld(dx.v)), dyf = std::fabs(mpf_get_ld(dy.v));
        const double mxx = std::max(ref_x, (double)(W - 1) - ref_x) + 0.5;
        const double mxy = std::max(ref_y, (double)(H - 1) - ref_y) + 0.5;
        build_bla(std::sqrt(dxf * dxf * mxx * mxx + dyf * dyf * mxy * mxy));
    } else {
        lev_off.clear();
        lev_size.clear();
        bla_r2.clear();
        bla_coef.clear();
        rmax2 = 0.0;
    }
    bla_seconds += now_s() - t_bla;
    // dc = trunc53(dx * (column - ref column)) from mpf (mandel_perturbation.cpp:2300-2312)
    dcol_d.resize(W);
    drow_d.resize(H);
    for (int j = 0; j < W; ++j) {
        const int xp = j - pj;
        mpf_mul_ui(t, dx.v, (unsigned long)std::abs(xp));
        if (xp < 0) mpf_neg(t, t);
        dcol_d[j] = mpf_get_ld(t);
    }
    for (int i = 0; i < H; ++i) {
        const int yp = i - pi;
        mpf_mul_ui(t, dy.v, (unsigned long)std::abs(yp));
        if (yp < 0) mpf_neg(t, t);
        drow_d[i] = mpf_get_ld(t);
    }
}

void Frame::cell_dc(std::vector<double>& col, std::vector<double>& row) const {
    const int h = sub / 2;
    const long rc = (long)ref_j * sub + ref_v + h, rr = (long)ref_i * sub + ref_u + h;
    Gmp t(prec);
    col.resize((size_t)W * sub);
    row.resize((size_t)H * sub);
    for (long c = 0; c < (long)W * sub; ++c) {
        const long xp = c - rc;
        mpf_mul_ui(t, dx.v, (unsigned long)std::labs(xp));
        if (xp < 0) mpf_neg(t, t);
        mpf_div_ui(t, t, (unsigned long)sub);
        col[(size_t)c] = mpf_get_ld(t);
    }
    for (long r = 0; r < (long)H * sub; ++r) {
        const long yp = r - rr;
        mpf_mul_ui(t, dy.v, (unsigned long)std::labs(yp));
        if (yp < 0) mpf_neg(t, t);
        mpf_div_ui(t, t, (unsigned long)sub);
        row[(size_t)r] = mpf_get_ld(t);
    }
}

void Frame::init_mpf() {
    for (mpf_ptr x : {c0_re.v, c0_im.v, dx.v, dy.v, ref_re.v, ref_im.v, tail_re.v, tail_im.v}) mpf_set_prec(x, prec);
}

Frame::Frame() {
    for (mpf_ptr x : {c0_re.v, c0_im.v, dx.v, dy.v, ref_re.v, ref_im.v, tail_re.v, tail_im.v}) mpf_init2(x, 64);
}

Frame::~Frame() {
    ref_pixel_wait();
    for (mpf_ptr x : {c0_re.v, c0_im.v, dx.v, dy.v, ref_re.v, ref_im.v, tail_re.v, tail_im.v}) mpf_clear(x);
}

void Frame::fill_view(mx_frame_view& v) const {
    std::memset(&v, 0, sizeof v);
    v.len = period ? period : ref_it + 1;
    v.period = period;
    v.subpixel = subpixel;
    v.zf = zf.data();
    v.zfe_m = zfe_m.data();
    v.zfe_e = zfe_e.data();
    v.nlev = (int32_t)lev_off.size();
    v.lev_off = lev_off.data();
    v.lev_size = lev_size.data();
    v.bla_total = (int64_t)bla_r2.size();
    v.bla_r2 = bla_r2.data();
    v.bla_coef = bla_coef.data();
    v.rmax2 = rmax2;
    v.dcmax = dcmax;
    v.dcol_m = dcol_m.data();
    v.dcol_e = dcol_e.data();
    v.drow_m = drow_m.data();
    v.drow_e = drow_e.data();
    v.dx_m = dxfe.m;
    v.dx_e = dxfe.e;
    v.dy_m = dyfe.m;
    v.dy_e = dyfe.e;
    v.esc2 = esc2;
    v.bla_minlevel = bla_minlevel;
    v.ref_it = ref_it;
    v.ref_escaped = ref_escaped;
    v.ref_virtual = ref_virtual;
    v.ref_x = ref_x;
    v.ref_y = ref_y;
    v.ref_pixel_n = ref_pixel_n;
    v.ref_pixel_value = ref_pixel_value;
    v.cutoff_sensitive = cutoff_sensitive;
    v.precision_bits = (int32_t)prec;
    v.ref_seconds = ref_seconds;
    v.bla_seconds = bla_seconds;
    v.path = dbl ? 1 : 2;
    v.escalated = escalated;
    if (dbl) {
        v.z_m3 = z_m3.data();
        v.sa_it = sa_it;
        v.sa_order = sa_order;
        v.sa_coef = &sa_A[0].r;
        v.sa_delta_re = sa_delta.r;
        v.sa_delta_im = sa_delta.i;
        v.dcol_d = dcol_d.data();
        v.drow_d = drow_d.data();
        if (ede) {
            v.dfr = dfr.data();
            v.dfi = dfi.data();
            v.bla_der = bla_der.data();
            v.sa_coef_b = &sa_B[0].r;
        }
    }
}

} // namespace mx

extern "C" int mx_frame_prepare(const mx_request* req, mx_frame** out, mx_frame_view* view) {
    if (!req || !out) return MX_E_INVALID;
    *out = nullptr;
    auto* f = new (std::nothrow) mx_frame();
    if (!f) return MX_E_NOMEM;
    std::string err;
    int rc = f->impl.prepare(*req, err);
    if (rc == MX_OK && f->impl.shallow) rc = MX_E_UNSUPPORTED; // no reference: mx_compute runs such views
    if (rc != MX_OK) {
        delete f;
        return rc;
    }
    f->impl.ref_pixel_wait();
    if (view) f->impl.fill_view(*view);
    *out = f;
    return MX_OK;
}

extern "C" void mx_frame_free(mx_frame* f) { delete f; }
