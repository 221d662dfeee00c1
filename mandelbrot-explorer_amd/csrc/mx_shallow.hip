// Shallow direct path for gfx950: views at scale <= 1e6, where the engine iterates
// every pixel's own c in plain double instead of perturbing around a reference
// (Mandel::Compute's method 0, src/engine/mandel_perturbation.cpp:1444-1519; adaptive
// sub-samples 1933-1965).
//
// Per-sample algorithm: the engine's lane-refill loop solveShallowSimdList
// (mandel_perturbation.cpp:416-532), whose values equal the scalar floatPointCompute
// (330-399) for smooth and exterior-DE colouring: z = c, d = 2, dc = 1, then per
// iteration d' = 2 z d, dc' = 2 z dc + 1, z' = z^2 + c (all from the old z), escape
// when |z'|^2 > R^2, interior when |d'|^2 < 1e-9, -2 at max_iter.  IEEE FP64 in the
// engine's operation order; this file is built with -ffp-contract=off.
//
// c of a sample, in double as the engine forms it from the truncated mpf grid values
// c0, dx, dy: base pixel (i, j) at (c0r + dx j, c0i + dy i); adaptive sub-sample
// (xi, yi) of pixel (i, j) at ((c0r + dx j) + dx xi / sub, (c0i + dy i) + dy yi / sub).
//
// Execution model: a persistent grid whose lanes each own one sample.  Base pixels are
// handed out in 8x8 tiles (neighbouring pixels in one wave, similar escape counts) from
// one atomic queue per XCD; a wave steps its lanes in runs of up to `chunk`
// iterations and leaves a run early once fewer than quorum/64 of the lanes that
// started it are still iterating, then refills the finished lanes.
#include <climits>

#include "mx_dev.h"

namespace {

constexpr double LG2 = 0.69314718055994530942; // log(2.0), the double libm returns

struct Lane {
    double cr, ci, zr, zi, dr, di, er, ei;
    int it;
};

__device__ __forceinline__ unsigned long long wave_sum(unsigned long long x) {
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
    return x;
}

// Sample `item` of the launch: its output index and c.  False for the padding lanes
// of edge tiles.
__device__ __forceinline__ bool shallow_sample(const MxShallowDev& S, unsigned item, int& out, double& cr, double& ci) {
    if (S.list) {
        const int per = S.sub * S.sub - 1;
        if (item >= (unsigned)(S.nlist * per)) return false;
        const int pix = S.list[item / (unsigned)per];
        int t = (int)(item % (unsigned)per);
        const int h = S.sub / 2;
        if (t >= h * S.sub + h) ++t; // the centre cell is the base pixel
        const int xi = t / S.sub - h, yi = t % S.sub - h;
        const int i = pix / S.W, j = pix % S.W;
        cr = S.c0r + S.dx * j + S.dx * xi / S.sub;
        ci = S.c0i + S.dy * i + S.dy * yi / S.sub;
        out = (int)item;
        return true;
    }
    const unsigned tile = item >> 6, l = item & 63;
    const int i = (int)(tile / (unsigned)S.tiles_x) * 8 + (int)(l >> 3);
    const int j = (int)(tile % (unsigned)S.tiles_x) * 8 + (int)(l & 7);
    if (i >= S.H || j >= S.W) return false;
    cr = S.c0r + S.dx * j;
    ci = S.c0i + S.dy * i;
    out = i * S.W + j;
    return true;
}

template <bool EDE>
__global__ __launch_bounds__(256) void mx_shallow_kernel(MxShallowDev S) {
    const int lane = __lane_id();
    const unsigned long long lanemask_lt = (1ull << lane) - 1ull;
    int q = blockIdx.x % MX_NQUEUE;
    int exhausted = 0;
    bool alive = true, active = false;
    Lane p{};
    int oidx = 0;
    unsigned long long iters = 0;
    unsigned runs = 0;
    const unsigned nitems_all = S.nitems;
    while (true) {
        bool need = alive && !active;
        unsigned long long needmask = __ballot(need);
        while (needmask != 0ull && exhausted < MX_NQUEUE) {
            const unsigned i0 = (unsigned)((unsigned long long)nitems_all * q / MX_NQUEUE);
            const unsigned i1 = (unsigned)((unsigned long long)nitems_all * (q + 1) / MX_NQUEUE);
            const unsigned nitems = i1 - i0;
            const int k = __popcll(needmask);
            const int leader = __ffsll((long long)needmask) - 1;
            unsigned base = 0;
            if (lane == leader) base = atomicAdd(&S.queue[q], (unsigned)k);
            base = __shfl(base, leader);
            if (need) {
                const unsigned it = base + (unsigned)__popcll(needmask & lanemask_lt);
                if (it < nitems && shallow_sample(S, i0 + it, oidx, p.cr, p.ci)) {
                    p.zr = p.cr;
                    p.zi = p.ci;
                    p.dr = 2.0;
                    p.di = 0.0;
                    p.er = 1.0;
                    p.ei = 0.0;
                    p.it = 1;
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
        const unsigned long long am = __ballot(active);
        if (am == 0ull) break;
        if (((++runs) & 1023u) == 0u && S.cancel && *S.cancel) break;
        const int stay = (__popcll(am) * S.quorum + 63) >> 6;
        for (int k = 0; k < S.chunk; ++k) {
            if (active) {
                const double ndr = 2.0 * (p.dr * p.zr - p.di * p.zi);
                const double ndi = 2.0 * (p.dr * p.zi + p.di * p.zr);
                if (EDE) {
                    const double ner = 2.0 * (p.er * p.zr - p.ei * p.zi) + 1.0;
                    const double nei = 2.0 * (p.er * p.zi + p.ei * p.zr);
                    p.er = ner;
                    p.ei = nei;
                }
                const double nzr = p.zr * p.zr - p.zi * p.zi + p.cr;
                const double nzi = 2.0 * (p.zr * p.zi) + p.ci;
                p.zr = nzr;
                p.zi = nzi;
                p.dr = ndr;
                p.di = ndi;
                ++iters;
                const double zrad = nzr * nzr + nzi * nzi;
                const double dmag = ndr * ndr + ndi * ndi;
                int n = 0;
                double v = -2.0;
                if (zrad > S.esc2) {
                    n = p.it + 1;
                    if (EDE) {
                        const float r = (float)(sqrt(zrad) * log(zrad) / sqrt(p.er * p.er + p.ei * p.ei));
                        v = (double)r >= 0.0 ? (double)r / S.dx : (double)r;
                    } else {
                        v = (double)p.it + 1.0 - log(log(zrad) / 2.0 / LG2) / LG2;
                    }
                } else if (dmag < 1e-9) {
                    n = -2;
                } else if (p.it + 1 >= S.mxit) {
                    n = -1;
                } else {
                    ++p.it;
                }
                if (n != 0) {
                    S.val[oidx] = v;
                    S.nesc[oidx] = n;
                    S.out[oidx] = (float)v;
                    active = false;
                }
            }
            if (__popcll(__ballot(active)) < stay) break;
        }
    }
    const unsigned long long s0 = wave_sum(iters);
    if (lane == 0) atomicAdd(&S.stats[0], s0);
}

} // namespace

hipError_t mx_launch_shallow(const MxShallowDev& S, bool ede, int blocks, hipStream_t stream) {
    if (ede)
        hipLaunchKernelGGL(mx_shallow_kernel<true>, dim3(blocks), dim3(256), 0, stream, S);
    else
        hipLaunchKernelGGL(mx_shallow_kernel<false>, dim3(blocks), dim3(256), 0, stream, S);
    return hipGetLastError();
}
