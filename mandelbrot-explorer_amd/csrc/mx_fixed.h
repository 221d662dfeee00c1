// Fixed-point reference orbit arithmetic of the engine's floatexp path
// (src/engine/bigfixed.h, used by createRef/calCoefficient when the frame is on the
// floatexp path: mandel_perturbation.cpp:4084-4111, 4197-4227).
//
// A value is sign * M * 2^(-64(L-1)) with M an L-limb magnitude: the top limb is the
// integer part, L-1 limbs of fraction.  Reference-orbit values are O(1), so adds and
// subtracts are limbwise; a product keeps limbs L-1 .. 2L-2 of the 2L-limb product,
// rounded on the most significant dropped bit.  For 16 <= L < 72 only the columns
// from L-4 up are accumulated (the dropped partial products would cost half the
// multiply), which is the engine's rule, so the rounding of such products may differ
// from the exact product's by the carry of the skipped columns.
//
// The double shadow of a value is built from its top two nonzero limbs,
// (double)hi * 2^64 + (double)lo, normalised into a FloatExp.
//
// Storage: one FixedArith owns every limb array its values use (a fixed number of
// values, allocated once), so the orbit loop allocates nothing.
#pragma once

#include <gmp.h>

#include <cstdlib>
#include <cstring>
#include <new>

#include "mx_fe.h"

namespace mx {

struct Fixed {
    int sign = 0;              // +1, -1, 0 (exact zero)
    mp_limb_t* m = nullptr;    // L limbs, least significant first, owned by a FixedArith
};

class FixedArith {
  public:
    // room for `nvals` values of `limbs` limbs plus the product scratch
    FixedArith(int limbs, int nvals) : L(limbs), cap(nvals) {
        const size_t n = (size_t)limbs * (size_t)nvals + 2 * (size_t)limbs + 8;
        store = static_cast<mp_limb_t*>(std::calloc(n, sizeof(mp_limb_t)));
        if (!store) throw std::bad_alloc();
        tmp = store + (size_t)limbs * (size_t)nvals;
    }
    ~FixedArith() { std::free(store); }
    FixedArith(const FixedArith&) = delete;
    FixedArith& operator=(const FixedArith&) = delete;

    int limbs() const { return L; }

    // a zero value in the next free slot; false once every slot is handed out
    bool make(Fixed& a) {
        if (used >= cap) return false;
        a.m = store + (size_t)L * (size_t)used++;
        a.sign = 0;
        std::memset(a.m, 0, sizeof(mp_limb_t) * (size_t)L);
        return true;
    }

    void copy(Fixed& r, const Fixed& a) const {
        if (r.m != a.m) std::memcpy(r.m, a.m, sizeof(mp_limb_t) * (size_t)L);
        r.sign = a.sign;
    }

    // truncation toward zero of |x| * 2^(64(L-1))
    void from_mpf(Fixed& a, mpf_srcptr x) const {
        std::memset(a.m, 0, sizeof(mp_limb_t) * (size_t)L);
        a.sign = mpf_sgn(x);
        if (a.sign == 0) return;
        mpf_t t;
        mpf_init2(t, (mp_bitcnt_t)64 * L + 8);
        mpf_abs(t, x);
        mpf_mul_2exp(t, t, (mp_bitcnt_t)64 * (L - 1));
        mpz_t z;
        mpz_init(z);
        mpz_set_f(z, t);
        if (mpz_size(z) <= (size_t)L) mpz_export(a.m, nullptr, -1, sizeof(mp_limb_t), 0, 0, z);
        mpz_clear(z);
        mpf_clear(t);
    }

    // the value into an mpf (rounded to the mpf's own precision by mpf_set_z)
    void to_mpf(mpf_ptr out, const Fixed& a) const {
        mpz_t z;
        mpz_init(z);
        mpz_import(z, (size_t)L, -1, sizeof(mp_limb_t), 0, 0, a.m);
        mpf_set_z(out, z);
        mpf_div_2exp(out, out, (mp_bitcnt_t)64 * (L - 1));
        if (a.sign < 0) mpf_neg(out, out);
        mpz_clear(z);
    }

    // r = sa*|a| + sb*|b| (signs passed separately: a subtraction flips sb).  r may be
    // a or b.
    void addsub(Fixed& r, const Fixed& a, int sa, const Fixed& b, int sb) const {
        if (sa == 0) {
            if (r.m != b.m) std::memcpy(r.m, b.m, sizeof(mp_limb_t) * (size_t)L);
            r.sign = sb;
            return;
        }
        if (sb == 0) {
            if (r.m != a.m) std::memcpy(r.m, a.m, sizeof(mp_limb_t) * (size_t)L);
            r.sign = sa;
            return;
        }
        if (sa == sb) {
            mpn_add_n(r.m, a.m, b.m, L); // bounded values: no carry out
            r.sign = sa;
            return;
        }
        const int c = mpn_cmp(a.m, b.m, L);
        if (c == 0) {
            std::memset(r.m, 0, sizeof(mp_limb_t) * (size_t)L);
            r.sign = 0;
        } else if (c > 0) {
            mpn_sub_n(r.m, a.m, b.m, L);
            r.sign = sa;
        } else {
            mpn_sub_n(r.m, b.m, a.m, L);
            r.sign = sb;
        }
    }
    void add(Fixed& r, const Fixed& a, const Fixed& b) const { addsub(r, a, a.sign, b, b.sign); }
    void sub(Fixed& r, const Fixed& a, const Fixed& b) const { addsub(r, a, a.sign, b, -b.sign); }

    // r = a * b: limbs L-1 .. 2L-2 of the product, rounded on the top dropped bit.  A
    // magnitude that rounds to zero keeps its sign (every consumer reads it as 0).  r
    // must not be a or b.
    void mul(Fixed& r, const Fixed& a, const Fixed& b) {
        if (a.sign == 0 || b.sign == 0) {
            std::memset(r.m, 0, sizeof(mp_limb_t) * (size_t)L);
            r.sign = 0;
            return;
        }
        mp_limb_t* h = tmp;
        mp_limb_t round_limb;
        if (L >= 16 && L < 72) {
            // columns lo .. 2L-1 of the product, every partial product a_i b_j with
            // i + j >= lo accumulated row by row
            const int G = 3, lo = L - 1 - G, nout = 2 * L - lo;
            std::memset(h, 0, sizeof(mp_limb_t) * (size_t)(nout + 1));
            for (int i = 0; i < L; ++i) {
                const int j0 = lo - i > 0 ? lo - i : 0;
                const int n = L - j0;
                const int t = i + j0 - lo;
                mp_limb_t cy = mpn_addmul_1(h + t, b.m + j0, n, a.m[i]);
                for (int u = t + n; cy && u <= nout; ++u) {
                    h[u] += cy;
                    cy = h[u] < cy;
                }
            }
            std::memcpy(r.m, h + G, sizeof(mp_limb_t) * (size_t)L);
            round_limb = h[G - 1];
        } else {
            mpn_mul_n(h, a.m, b.m, L);
            std::memcpy(r.m, h + (L - 1), sizeof(mp_limb_t) * (size_t)L);
            round_limb = L >= 2 ? h[L - 2] : 0;
        }
        if (round_limb >> 63) mpn_add_1(r.m, r.m, L, 1);
        r.sign = a.sign == b.sign ? 1 : -1;
    }

    // FloatExp shadow from the top two nonzero limbs
    Fe to_fe(const Fixed& a) const {
        if (a.sign == 0) return fe_make(0.0, 0);
        int top = L - 1;
        while (top >= 0 && a.m[top] == 0) --top;
        if (top < 0) return fe_make(0.0, 0);
        const double hi = (double)a.m[top];
        const double lo = top >= 1 ? (double)a.m[top - 1] : 0.0;
        const double mant = hi * 18446744073709551616.0 + lo;
        Fe r = fe_from(mant);
        r.e += 64 * (top - L);
        if (a.sign < 0) r.m = -r.m;
        return r;
    }

  private:
    int L, cap, used = 0;
    mp_limb_t* store = nullptr;
    mp_limb_t* tmp = nullptr;
};

} // namespace mx
