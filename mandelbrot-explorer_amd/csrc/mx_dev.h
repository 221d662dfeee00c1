// Device-side layout shared by the host launcher and the HIP kernels.
//
// HBM layout of one reference (uploaded once per reference, read by every pixel):
//   zf        double2[reflen]      X_{k+1} double shadow, k = 0..mx_ref_it (16 B/entry)
//   zfe_m     double2[reflen]      FloatExp mantissas of the same orbit (rebase, underflow)
//   zfe_e     int2[reflen]         FloatExp exponents
//   zf32      float2[reflen]       FP32 shadow (mixed form only)
//   bla_r2    double[total]        BLA validity radius^2, all levels concatenated
//   bla_coef  double4[total]       (ar, ai, br, bi) per entry
//   level p holds lev_size[p] entries starting at lev_off[p]; entry i of level p
//   starts at reference index 1 + i*2^p and skips 2^p iterations.
// One tile (frame or sub-frame) adds separable dc tables (W + H FloatExp values) and
// the per-sample outputs (n_esc int32, value double, compat float).
#pragma once

#include "mx_fe.h"

// A pointer into device (global) memory.  The layouts below are written by host code
// and read by the kernels from a job table in memory, where a plain pointer is
// generic: every access through it would be a flat_* instruction with a 64-bit VGPR
// address, which also waits on the LDS counter.  In the device compile this wrapper
// hands out address-space-1 lvalues, so the same accesses become global_* loads and
// stores with an SGPR base and a 32-bit offset.  It holds one plain pointer (same
// size and layout in host and device code) and converts to and from T*.
#if defined(__HIP_DEVICE_COMPILE__)
#define MX_AS_GLOBAL __attribute__((address_space(1)))
#define MX_AS_LDS __attribute__((address_space(3)))
#else
#define MX_AS_GLOBAL
#define MX_AS_LDS
#endif
template <class T>
struct GPtr {
    T* p;
    GPtr() = default;
    __host__ __device__ constexpr GPtr(T* q) : p(q) {}
    __host__ __device__ constexpr GPtr(decltype(nullptr)) : p(nullptr) {}
    __host__ __device__ operator T*() const { return p; }
    __host__ __device__ T MX_AS_GLOBAL& operator[](long long i) const { return ((T MX_AS_GLOBAL*)p)[i]; }
    __host__ __device__ T MX_AS_GLOBAL& operator*() const { return *(T MX_AS_GLOBAL*)p; }
    __host__ __device__ GPtr operator+(long long i) const { return GPtr(p + i); }
};

#define MX_MAXLEV 32
#define MX_NQUEUE 8 // one work queue per XCD

enum MxFlags {
    MX_F_EDE = 1,        // exterior distance estimate output (carry dz/dc)
    MX_F_INTERIOR = 2,   // periodicity-based interior detection
    MX_F_INT_REP = 4,    // exact state-repeat interior detection
    MX_F_BLA = 8,        // BLA skipping
    MX_F_NOBLA_LANE = 16 // the engine's no-BLA vector loop (see mx_tile_flags)
};

// Flags of a floatexp tile.  With BLA off and no derivative the engine runs its no-BLA
// vector loop (Mandel::solveRescaledSimd8NoBla, mandel_perturbation.cpp:3620-3816,
// selected by stepParallel at 2148-2149 and 2172-2181): no interior detection, and the
// step-underflow recovery taken when |w'|^2 is exactly 0, with z then formed from the
// FloatExp orbit.  Everything else runs pixelRescaled's loop, including a periodic
// reference without BLA (the engine builds periodic references only with BLA on).
inline int mx_tile_flags(bool ede, bool interior, bool int_rep, bool bla, bool periodic) {
    const bool nobla_lane = !bla && !ede && !periodic;
    return (ede ? MX_F_EDE : 0) | (interior && !nobla_lane ? MX_F_INTERIOR : 0) |
           (int_rep && interior && !nobla_lane ? MX_F_INT_REP : 0) | (bla ? MX_F_BLA : 0) | (nobla_lane ? MX_F_NOBLA_LANE : 0);
}

struct MxRefDev {
    GPtr<const double2> zf;
    GPtr<const double2> zfe_m;
    GPtr<const int2> zfe_e;
    GPtr<const float2> zf32;
    GPtr<const double> bla_r2;
    GPtr<const double4> bla_coef;
    GPtr<const float4> bla_coeff;    // mixed form: coefficient mantissas (ar, ai) and (br, bi), FP32
    GPtr<const int2> bla_coefe;      // mixed form: exponents of A and of B
    int mx_ref_it;
    int nlev;
    int minlevel;
    int period;                      // > 0: periodic reference, orbit read at (m-1) mod period
    int nent;                        // entries in zf/zfe/zf32: mx_ref_it+1, or period+2 (two wrap copies)
    double rmax2;
    // plain-double path
    GPtr<const double> z_m3;         // |X|^2 / 1e6 per entry (glitch metric)
    GPtr<const double2> sa_A;        // series coefficients A_0..A_29
    double2 sa_delta;
    int sa_it, sa_order;
    GPtr<const double2> df;          // reference derivative D (EDE)
    GPtr<const double4> bla_der;     // BLA derivative couplings (cr, ci, er, ei) (EDE)
    GPtr<const double2> sa_B;        // derivative-delta series (EDE)
};

struct MxTileDev {
    GPtr<const Fe> dcol;             // per column dc (real part)
    GPtr<const Fe> drow;             // per row dc (imag part)
    int W, H;                        // sample grid of the tile
    int stride;                      // 1: every sample; 8: the interior-gate probe grid
    int nx, ny;                      // samples in this launch (after stride)
    int tiles_x;                     // 8x8 lane tiles per launch row
    int ntiles;
    int mxit;
    int flags;
    double esc2;
    double interior_eps2;
    int interior_confirm;
    int skip_sample;                 // sample index not computed (the reference pixel), -1 none
    int chunk;                       // lean steps per trip between full-path checks (results do not depend on it)
    int quorum;                      // a lean run ends once fewer than quorum/64 of its lanes still step (neither)
    Fe dxfe;
    GPtr<double> val;                // W*H
    GPtr<int> nesc;                  // W*H
    GPtr<float> out;                 // W*H (reference-compatible float buffer), may be null
    GPtr<unsigned> queue;            // MX_NQUEUE counters, zeroed before launch
    GPtr<unsigned long long> stats;  // steps, bla applies, bla skipped iterations, rebases
    GPtr<const volatile int> cancel; // host-mapped; nonzero => lanes stop taking work
    // plain-double path
    GPtr<const double> dcol_d;       // per column dc (double)
    GPtr<const double> drow_d;       // per row dc (double)
    GPtr<const int> list;            // sample indices to compute (null: the grid)
    int list_len;                    // entries of `list`, or of `sdc`
    GPtr<const Fe> sdc;              // floatexp forms: explicit (re, im) dc per sample (null: the grid);
                                     // outputs are then indexed by list entry
    GPtr<int> unres_idx;             // unresolved samples of this launch
    GPtr<double> unres_gp;           // their glitch metric
    GPtr<unsigned> unres_count;
};

// One tile of a launch and the reference it is computed against.  A launch runs its
// jobs in order: each wave takes samples of job j from that job's queues and moves
// to job j+1 once those queues are drained and its lanes are done, so tiles of
// different size, depth and iteration limit share one persistent grid.
struct MxJob {
    MxRefDev R;
    GPtr<const int2> lev;
    MxTileDev T;
};

// Shallow direct path (views at scale <= 1e6): every sample iterates its own c in
// double.  Base pass: the W x H grid in 8x8 tiles (nitems = 64 tiles); adaptive pass:
// the sub*sub - 1 cells of each of the nlist flagged pixels in `list` (nitems =
// nlist (sub*sub - 1)), outputs indexed by item.
struct MxShallowDev {
    double c0r, c0i, dx, dy;         // truncated mpf grid values
    int W, H, sub, mxit;
    double esc2;
    GPtr<const int> list;            // flagged pixels (row * W + column), null: the base grid
    int nlist;
    int tiles_x;
    unsigned nitems;
    int chunk, quorum;
    GPtr<double> val;
    GPtr<int> nesc;
    GPtr<float> out;
    GPtr<unsigned> queue;            // MX_NQUEUE counters, zeroed before launch
    GPtr<unsigned long long> stats;  // iterations
    GPtr<const volatile int> cancel; // host-mapped; nonzero => lanes stop taking work
};
