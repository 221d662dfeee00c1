// Plain-double perturbation kernel for gfx950 (frames at scale 1e6..1e280): IEEE FP64
// in the engine's operation order, no FP contraction (-ffp-contract=off), so escape
// counts equal the CPU engine's.
//
// Per-pixel algorithm: the scalar loop of Mandel::stepParallel
// (src/engine/mandel_perturbation.cpp:2291-2584) with tryBLA (2714-2755): the
// series-approximation seed dz0 = sum A_x (dc/delta)^(x+1) (2319-2323) starting at
// iteration sa_it+1, the step dz' = 2 X dz + dz^2 + dc against the double orbit,
// Zhuoran rebase when |z| < |dz| or at the end of the reference, and glitch
// detection (|dz|^2/|z|^2 > 1e7 at a rebase) that leaves the sample unresolved with
// its glitch metric |z|^2/_z_m3[j] for the host's next reference choice.  Under the
// exterior distance estimate the derivative delta dd = dz/dc - D rides along
// (2444-2456; through a BLA skip dd -> C dz + A dd + E dc, 2742-2749; seeded by its
// own series) and an escape returns sqrt|z|^2 ln|z|^2 / (dx |dz/dc|) (2487-2488).
//
// Execution model as in the floatexp kernels: a persistent grid whose lanes refill
// from per-XCD atomic work queues (wave-aggregated fetch-add), samples handed out
// in 8x8 tiles, or taken from a list of sample indices for a later reference pass.
//
// The orbit in LDS: every step reads the next orbit entry, a dependent chain of one
// load per iteration that the latency of L2 would otherwise set.  When the launch
// asks for it, each workgroup (1024 lanes, one per CU) first copies the orbit's first
// `nstage` entries -- the whole orbit when it fits, else the prefix that every rebase
// restarts in -- into LDS, and the steps read them from there.
#include <climits>

#include "mx_dbl.h"
#include "mx_dev.h"

namespace {

struct Pd {
    double dzr, dzi, dcr, dci;
    double ddr, ddi;
    double zsr, zsi, conf_D2, conf_zr, conf_zi;
    double sdzr, sdzi;
    int j, k;
    int save_j, period_win, conf_P, conf_next, conf_count, conf_giveup;
    int sk, save_j2, twin2;
    int xv;        // 1: Xc = X[k-1] and Xn = X[k] are loaded
    double2 Xc, Xn;
};

struct Counters {
    unsigned long long steps, applies, skipped, rebases;
};

__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// Series seeds (the Horner loops of mandel_perturbation.cpp:2319-2330) with the
// explicit complex formulas of mx_dbl.h.
__device__ __forceinline__ MxC sa_horner(GPtr<const double2> C, int order, MxC delta, MxC dc) {
    MxC z = {0.0, 0.0};
    for (int x = order; x >= 0; --x) {
        const double2 a = C[x];
        z = mxc_add(z, MxC{a.x, a.y});
        z = mxc_mul(z, dc);
        z = mxc_div(z, delta);
    }
    return z;
}

template <bool EDE>
__device__ void init_pixel(const MxRefDev& R, Pd& p, double dcr, double dci) {
    p.dcr = dcr;
    p.dci = dci;
    const MxC dc = {dcr, dci}, delta = {R.sa_delta.x, R.sa_delta.y};
    const MxC dz = sa_horner(R.sa_A, R.sa_order, delta, dc);
    p.dzr = dz.r;
    p.dzi = dz.i;
    if (EDE) {
        const MxC dd = sa_horner(R.sa_B, R.sa_order, delta, dc);
        p.ddr = dd.r;
        p.ddi = dd.i;
    } else {
        p.ddr = p.ddi = 0.0;
    }
    p.j = R.sa_it + 1;
    p.k = p.j;
    p.zsr = 1e30;
    p.zsi = 1e30;
    p.save_j = p.j;
    p.period_win = 1;
    p.conf_P = p.conf_next = p.conf_count = p.conf_giveup = 0;
    p.conf_D2 = p.conf_zr = p.conf_zi = 0.0;
    p.sdzr = 1e300;
    p.sdzi = 1e300;
    p.sk = -1;
    p.save_j2 = p.j;
    p.twin2 = 1;
    p.xv = 0;
}

// tryBLA (mandel_perturbation.cpp:2714-2755) without the derivative carry: the
// largest aligned level whose radius admits |dz|, landing at least two entries
// before the end of the reference and not beyond the escape radius.
struct BlaRes {
    double nzr, nzi, nddr, nddi;
    int skip;
};

template <bool EDE>
__device__ __noinline__ BlaRes try_bla(GPtr<const double> bla_r2, GPtr<const double4> bla_coef,
                                       GPtr<const double4> bla_der, GPtr<const double2> zf,
                                       GPtr<const int2> lev, int nlev, int minlevel, int mx_ref_it, int s, double dzr,
                                       double dzi, double ddr, double ddi, double dcr, double dci, double esc2) {
    BlaRes r = {0.0, 0.0, 0.0, 0.0, 0};
    if (s < 1) return r;
    const double zmag2 = dzr * dzr + dzi * dzi;
    const int maxp = nlev - 1;
    int startp = (s == 1) ? maxp : __builtin_ctz((unsigned)(s - 1));
    if (startp > maxp) startp = maxp;
    if (startp < minlevel) return r;
    for (int p = startp; p >= minlevel; --p) {
        const int2 L = lev[p];
        const int i = (s - 1) >> p;
        if (i >= L.y) continue;
        const int q = L.x + i;
        const double r2 = bla_r2[q];
        if (r2 <= 0 || zmag2 >= r2) continue;
        const int l = 1 << p;
        const int land = s + l;
        if (land >= mx_ref_it - 1) continue;
        const double4 c = bla_coef[q];
        const double nzr = c.x * dzr - c.y * dzi + c.z * dcr - c.w * dci;
        const double nzi = c.x * dzi + c.y * dzr + c.z * dci + c.w * dcr;
        const double2 Z = zf[land];
        const double fr = Z.x + nzr, fi = Z.y + nzi;
        if (fr * fr + fi * fi > esc2) return r;
        if (EDE) {
            const double4 d = bla_der[q];
            r.nddr = (d.x * dzr - d.y * dzi) + (c.x * ddr - c.y * ddi) + (d.z * dcr - d.w * dci);
            r.nddi = (d.x * dzi + d.y * dzr) + (c.x * ddi + c.y * ddr) + (d.z * dci + d.w * dcr);
        }
        r.nzr = nzr;
        r.nzi = nzi;
        r.skip = l;
        return r;
    }
    return r;
}

// The k >= 1 step (mandel_perturbation.cpp:2444-2483), shared by the full and the lean
// path (this file is built without FP contraction: identical expressions, identical
// bits): dz' = 2 dz X + dz^2 + dc, and under EDE dd' = 2 (dd X + dz (D + dd)).
__device__ __forceinline__ void dz_step(double dzr, double dzi, double dcr, double dci, double xr, double xi, double& ndzr,
                                        double& ndzi) {
    const double dzr2 = dzr * dzr, dzi2 = dzi * dzi;
    ndzr = (dzr * xr - dzi * xi) * 2 + dzr2 - dzi2 + dcr;
    ndzi = (dzr * xi + dzi * xr) * 2 + dzr * dzi * 2 + dci;
}
__device__ __forceinline__ void dd_step(double ddr, double ddi, double dzr, double dzi, double xr, double xi, double2 D,
                                        double& nddr, double& nddi) {
    nddr = (ddr * xr - ddi * xi + dzr * (D.x + ddr) - dzi * (D.y + ddi)) * 2;
    nddi = (ddr * xi + ddi * xr + dzr * (D.y + ddi) + dzi * (D.x + ddr)) * 2;
}

// Orbit entries: the LDS copy of the first `n` entries, the rest from global memory.
struct OrbitSrc {
    const double2 MX_AS_LDS* lds;
    int n;
    GPtr<const double2> zf;
    __device__ __forceinline__ double2 operator[](int k) const { return k < n ? double2(lds[k]) : double2(zf[k]); }
};

// One iteration of the stepParallel loop body for one lane.  Returns 1 when the
// sample is finished (n, v set; n == INT_MIN: unresolved, *gp set).
template <bool INTERIOR, bool EDE>
__device__ __forceinline__ int trip(const MxRefDev& R, GPtr<const int2> lev, const MxTileDev& T, const OrbitSrc& Z, Pd& p,
                                    double& v, int& n, double& gp, Counters& c) {
    if (p.j >= T.mxit) {
        v = -2.0;
        n = -1;
        return 1;
    }
    if (INTERIOR && (T.flags & MX_F_INT_REP)) {
        if (p.k == p.sk && p.dzr == p.sdzr && p.dzi == p.sdzi) {
            v = -2.0;
            n = -2;
            return 1;
        }
        if (p.j - p.save_j2 >= p.twin2) {
            p.sdzr = p.dzr;
            p.sdzi = p.dzi;
            p.sk = p.k;
            p.save_j2 = p.j;
            p.twin2 += p.twin2;
        }
    }
    if ((T.flags & MX_F_BLA) && p.dzr * p.dzr + p.dzi * p.dzi < R.rmax2) {
        const BlaRes b = try_bla<EDE>(R.bla_r2, R.bla_coef, R.bla_der, R.zf, lev, R.nlev, R.minlevel, R.mx_ref_it, p.k - 1,
                                      p.dzr, p.dzi, p.ddr, p.ddi, p.dcr, p.dci, T.esc2);
        const int skip = b.skip;
        if (skip > 0) {
            p.dzr = b.nzr;
            p.dzi = b.nzi;
            if (EDE) {
                p.ddr = b.nddr;
                p.ddi = b.nddi;
            }
            p.k += skip;
            p.j += skip;
            p.xv = 0;
            c.applies++;
            c.skipped += (unsigned long long)skip;
            p.zsr = 1e30;
            p.zsi = 1e30;
            p.save_j = p.j;
            p.period_win = 1;
            p.conf_P = 0;
            p.conf_giveup = 0;
            return 0;
        }
    }
    if (!p.xv) {
        if (p.k != 0) p.Xc = Z[p.k - 1];
        p.Xn = Z[p.k];
    }
    // the entry after this step's, issued now so its latency overlaps the step; k can
    // equal mx_ref_it only on the first step after a series skip to the reference's
    // escape, and then the next step rebases, so the clamped entry goes unused
    const double2 Xpf = Z[min(p.k + 1, R.mx_ref_it)];
    double dr = 0.0, di = 0.0;
    if (EDE) {
        // dd' = 2 (dd X + dz (D + dd)) from the dz before this step (2444-2456)
        double t2;
        if (p.k == 0) {
            t2 = (p.dzr * (1 + p.ddr) - p.dzi * (1 + p.ddi)) * 2;
            p.ddi = (p.dzr * (1 + p.ddi) + p.dzi * (1 + p.ddr)) * 2;
        } else {
            dd_step(p.ddr, p.ddi, p.dzr, p.dzi, p.Xc.x, p.Xc.y, R.df[p.k - 1], t2, p.ddi);
        }
        p.ddr = t2;
        const double2 Dk = R.df[p.k];
        dr = p.ddr + Dk.x;
        di = p.ddi + Dk.y;
    }
    double tmp, dzi;
    if (p.k == 0) {
        const double dzr2 = p.dzr * p.dzr, dzi2 = p.dzi * p.dzi;
        tmp = dzr2 - dzi2 + p.dcr;
        dzi = p.dzr * p.dzi * 2 + p.dci;
    } else {
        dz_step(p.dzr, p.dzi, p.dcr, p.dci, p.Xc.x, p.Xc.y, tmp, dzi);
    }
    const double dzr = tmp;
    const double2 Xn = p.Xn;
    const double zr = dzr + Xn.x, zi = dzi + Xn.y;
    const double zrad = zr * zr + zi * zi;
    p.dzr = dzr;
    p.dzi = dzi;
    ++p.k;
    p.Xc = Xn; // X[k-1] and X[k] of the next step
    p.Xn = Xpf;
    p.xv = 1;
    c.steps++;
    if (zrad > T.esc2) {
        if (EDE) {
            const double dxf = ldexp(T.dxfe.m, T.dxfe.e); // mpf_get_ld(_dx)
            v = sqrt(zrad) / dxf * log(zrad) / sqrt(dr * dr + di * di);
        } else {
            const double LG = 0.69314718055994530942; // log(2) rounded to double
            v = p.j + 1 - log(log(zrad) / 2 / LG) / LG;
        }
        n = p.j + 1;
        return 1;
    }
    if (INTERIOR) {
        if (p.conf_P > 0) {
            p.conf_D2 *= 4 * zrad;
            if (p.conf_D2 > 1e18) p.conf_D2 = 1e18;
            if (p.j >= p.conf_next) {
                const double pr = zr - p.conf_zr, pi = zi - p.conf_zi;
                const bool ret = (pr * pr + pi * pi < T.interior_eps2 * zrad);
                if (ret && p.conf_D2 < 1.0) {
                    if (++p.conf_count >= T.interior_confirm) {
                        v = -2.0;
                        n = -2;
                        return 1;
                    }
                    p.conf_zr = zr;
                    p.conf_zi = zi;
                    p.conf_D2 = 1;
                    p.conf_next = p.j + p.conf_P;
                } else {
                    p.conf_P = 0;
                    ++p.conf_giveup;
                }
            }
        } else if ((p.j & 15) == 0 && p.conf_giveup < 3) {
            const double pr = zr - p.zsr, pi = zi - p.zsi;
            if (pr * pr + pi * pi < T.interior_eps2 * zrad) {
                p.conf_P = p.j - p.save_j;
                if (p.conf_P < 1) p.conf_P = 1;
                p.conf_zr = zr;
                p.conf_zi = zi;
                p.conf_D2 = 1;
                p.conf_next = p.j + p.conf_P;
                p.conf_count = 0;
            }
            if (p.j - p.save_j >= p.period_win) {
                p.zsr = zr;
                p.zsi = zi;
                p.save_j = p.j;
                p.period_win += p.period_win;
            }
        }
    }
    const double dz2 = dzr * dzr + dzi * dzi;
    // end of the reference: k >= mx_ref_it (a series skip that ends at the reference's
    // escape iteration starts at k == mx_ref_it, so k reaches mx_ref_it + 1 here)
    if (zrad < dz2 || p.k >= R.mx_ref_it) {
        if (dz2 / zrad > 10000000) {
            // glitch: the metric reads _z_m3 at the pixel's iteration, clamped to the
            // reference's last entry
            const int jm = p.j < R.mx_ref_it ? p.j : R.mx_ref_it;
            gp = zrad / R.z_m3[jm];
            n = INT_MIN;
            return 1;
        }
        p.dzr = zr;
        p.dzi = zi;
        if (EDE) {
            p.ddr = dr - 1;
            p.ddi = di;
        }
        p.k = 0;
        p.xv = 0;
        c.rebases++;
    }
    ++p.j;
    return 0;
}

// Lean steps: the trip's k >= 1 step for up to T.chunk iterations while it is provably
// the whole iteration, so results are bit-identical to trips alone.  A lane stops, its
// state untouched, before a step at which the trip would do more: the iteration
// limit, the end of the reference (k + 1 >= mx_ref_it, the rebase there), an
// exact-repeat save or the index back at the saved one, an active periodicity
// confirmation, a BLA try (|dz|^2 below the largest radius at an index whose level
// search is not empty: k == 2 or k - 2 a multiple of 2^minlevel); and after a step
// whose outcome is not plain: escape, a close return on a periodicity check,
// |z| < |dz|.  The periodicity check's save (every 16 iterations) is done here as the
// trip does it.  The wave leaves a run together once fewer than quorum/64 of the lanes
// that started it still step.
template <bool INTERIOR, bool EDE>
__device__ __forceinline__ void fast_steps(const MxRefDev& R, const MxTileDev& T, const OrbitSrc& Z, Pd& p, Counters& c) {
    int lim = min(T.chunk, T.mxit - p.j);
    lim = min(lim, R.mx_ref_it - 1 - p.k);
    if (p.k == 0 || !p.xv) lim = 0;
    const bool int_rep = INTERIOR && (T.flags & MX_F_INT_REP);
    if (int_rep) {
        lim = min(lim, p.save_j2 + p.twin2 - p.j);
        if (p.sk - p.k >= 0) lim = min(lim, p.sk - p.k);
    }
    if (INTERIOR && p.conf_P > 0) lim = 0;
    const bool can_bla = (T.flags & MX_F_BLA) != 0;
    const int amask = (1 << R.minlevel) - 1;
    const bool chk_ok = INTERIOR && (T.flags & MX_F_INTERIOR) && p.conf_giveup < 3;
    const int need = (__popcll(__ballot(lim > 0)) * T.quorum + 63) >> 6;
    if (lim <= 0) return;
    double dzr = p.dzr, dzi = p.dzi, ddr = p.ddr, ddi = p.ddi;
    double w2 = dzr * dzr + dzi * dzi;
    double2 Xc = p.Xc, Xn = p.Xn;
    double2 D = EDE ? double2(R.df[p.k - 1]) : make_double2(0.0, 0.0);
    int k = p.k, j = p.j, n = 0;
    while (n < lim) {
        if (can_bla && w2 < R.rmax2 && k >= 2 && (k == 2 || ((k - 2) & amask) == 0)) break;
        const double2 Xpf = Z[k + 1];
        double nddr = 0.0, nddi = 0.0;
        double2 Dk = D;
        if (EDE) {
            dd_step(ddr, ddi, dzr, dzi, Xc.x, Xc.y, D, nddr, nddi);
            Dk = R.df[k];
        }
        double ndzr, ndzi;
        dz_step(dzr, dzi, p.dcr, p.dci, Xc.x, Xc.y, ndzr, ndzi);
        const double zr = ndzr + Xn.x, zi = ndzi + Xn.y;
        const double zrad = zr * zr + zi * zi;
        const double dz2 = ndzr * ndzr + ndzi * ndzi;
        bool stop = zrad > T.esc2 || zrad < dz2;
        bool save = false;
        if (INTERIOR && chk_ok && (j & 15) == 0) {
            const double pr = zr - p.zsr, pi = zi - p.zsi;
            stop = stop || pr * pr + pi * pi < T.interior_eps2 * zrad;
            save = j - p.save_j >= p.period_win;
        }
        if (stop) break;
        if (INTERIOR && save) {
            p.zsr = zr;
            p.zsi = zi;
            p.save_j = j;
            p.period_win += p.period_win;
        }
        dzr = ndzr;
        dzi = ndzi;
        if (EDE) {
            ddr = nddr;
            ddi = nddi;
            D = Dk;
        }
        w2 = dz2;
        Xc = Xn;
        Xn = Xpf;
        ++k;
        ++j;
        ++n;
        if (__popcll(__ballot(1)) < need) break;
    }
    p.dzr = dzr;
    p.dzi = dzi;
    if (EDE) {
        p.ddr = ddr;
        p.ddi = ddi;
    }
    p.Xc = Xc;
    p.Xn = Xn;
    p.k = k;
    p.j = j;
    c.steps += (unsigned long long)n;
}

// Map a queue item to a sample index; false for the padding lanes of edge tiles.
__device__ __forceinline__ bool item_to_sample(const MxTileDev& T, unsigned item, int& sidx) {
    if (T.list) {
        sidx = T.list[item];
        return sidx != T.skip_sample;
    }
    const unsigned tile = item >> 6, l = item & 63;
    const int ty = (int)(tile / (unsigned)T.tiles_x), tx = (int)(tile % (unsigned)T.tiles_x);
    const int gi = ty * 8 + (int)(l >> 3), gj = tx * 8 + (int)(l & 7);
    if (gi >= T.ny || gj >= T.nx) return false;
    sidx = gi * T.stride * T.W + gj * T.stride;
    return sidx != T.skip_sample;
}

template <bool INTERIOR, bool EDE, bool STAGE>
__global__ __launch_bounds__(STAGE ? 1024 : 256) void mx_dbl_kernel(MxRefDev R, GPtr<const int2> lev, MxTileDev T, int nstage) {
    extern __shared__ double2 orbit_lds[];
    OrbitSrc Z{(const double2 MX_AS_LDS*)orbit_lds, 0, R.zf};
    if (STAGE) {
        for (int k = (int)threadIdx.x; k < nstage; k += (int)blockDim.x) orbit_lds[k] = R.zf[k];
        __syncthreads();
        Z.n = nstage;
    }
    const int lane = __lane_id();
    const unsigned long long lanemask_lt = (1ull << lane) - 1ull;
    int q = blockIdx.x % MX_NQUEUE;
    int exhausted = 0;
    bool alive = true, active = false;
    Pd p;
    Counters c = {0, 0, 0, 0};
    int sidx = 0;
    unsigned trips = 0;
    const unsigned nitems_all = T.list ? (unsigned)T.list_len : (unsigned)T.ntiles * 64u;
    while (true) {
        bool need = alive && !active;
        unsigned long long needmask = __ballot(need);
        while (needmask != 0ull && exhausted < MX_NQUEUE) {
            const unsigned i0 = (unsigned)(((unsigned long long)nitems_all * (unsigned)q) / MX_NQUEUE);
            const unsigned i1 = (unsigned)(((unsigned long long)nitems_all * (unsigned)(q + 1)) / MX_NQUEUE);
            const unsigned nitems = i1 - i0;
            const int k = __popcll(needmask);
            const int leader = __ffsll((long long)needmask) - 1;
            unsigned base = 0;
            if (lane == leader) base = atomicAdd(&T.queue[q], (unsigned)k);
            base = __shfl(base, leader);
            if (need) {
                const unsigned it = base + (unsigned)__popcll(needmask & lanemask_lt);
                if (it < nitems && item_to_sample(T, i0 + it, sidx)) {
                    init_pixel<EDE>(R, p, T.dcol_d[sidx % T.W], T.drow_d[sidx / T.W]);
                    need = false;
                    active = true;
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
            double v = 0.0, gp = 0.0;
            int n = 0;
            if (trip<INTERIOR, EDE>(R, lev, T, Z, p, v, n, gp, c)) {
                if (n == INT_MIN) {
                    const unsigned slot = atomicAdd(T.unres_count, 1u);
                    T.unres_idx[slot] = sidx;
                    T.unres_gp[slot] = gp;
                } else {
                    T.val[sidx] = v;
                    T.nesc[sidx] = n;
                    if (T.out) T.out[sidx] = n < 0 ? -2.0f : (float)v;
                }
                active = false;
            } else {
                fast_steps<INTERIOR, EDE>(R, T, Z, p, c);
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

} // namespace

namespace {

template <bool INTERIOR, bool EDE>
hipError_t launch_dbl(const MxRefDev& R, const int2* lev, const MxTileDev& T, int blocks, int threads, int nstage, hipStream_t stream) {
    if (nstage > 0) {
        const size_t bytes = (size_t)nstage * sizeof(double2);
        hipError_t e = hipFuncSetAttribute((const void*)mx_dbl_kernel<INTERIOR, EDE, true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)bytes);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL((mx_dbl_kernel<INTERIOR, EDE, true>), dim3(blocks), dim3(1024), bytes, stream, R, lev, T, nstage);
    } else {
        hipLaunchKernelGGL((mx_dbl_kernel<INTERIOR, EDE, false>), dim3(blocks), dim3(threads), 0, stream, R, lev, T, 0);
    }
    return hipGetLastError();
}

} // namespace

// nstage > 0: the orbit's first nstage entries staged in LDS, blocks of 1024 lanes
hipError_t mx_launch_dbl(const MxRefDev& R, const int2* lev, const MxTileDev& T, int blocks, int threads, hipStream_t stream, int nstage) {
    const bool ede = T.flags & MX_F_EDE, interior = T.flags & MX_F_INTERIOR;
    if (ede && interior) return launch_dbl<true, true>(R, lev, T, blocks, threads, nstage, stream);
    if (ede) return launch_dbl<false, true>(R, lev, T, blocks, threads, nstage, stream);
    if (interior) return launch_dbl<true, false>(R, lev, T, blocks, threads, nstage, stream);
    return launch_dbl<false, false>(R, lev, T, blocks, threads, nstage, stream);
}
