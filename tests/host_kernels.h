// Test infrastructure: host builds of the floatexp kernels' per-sample code.
//
// host_fe64.cpp and host_fe32.cpp include the kernel sources unchanged; their
// __host__ __device__ functions (sample set-up, the full step with every event, the
// lean steps) are driven here one sample at a time on CPU threads, with the
// reference and tile laid out exactly as the device sees them.  This lets the tests
//   * compare the lean-step loop with the full step alone, bit for bit and counter
//     for counter (the lean step must change nothing but the speed);
//   * run the mixed FP32 form and its FP64 variant next to the plain form on frames
//     too large for the GPU tests to keep around, and classify where they differ.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "../include/mx.h"
#include "../mandelbrot-explorer_amd/csrc/mx_dev.h"
#include "../mandelbrot-explorer_amd/csrc/mx_tables.h"

extern "C" {
typedef struct {
    int32_t form;          // MX_FORM_FE64, MX_FORM_FE32 or MX_FORM_SC64
    int32_t ede, interior, int_rep, bla, bla_minlevel;
    double interior_eps2;
    int32_t interior_confirm;
    int32_t lean;          // 1: full step + lean steps, as the kernel runs; 0: full steps only
    int32_t chunk;         // lean steps per trip
    int32_t threads;
} mxh_opts;
}

namespace mxh {

// The device layout of one reference (mx_ref_upload) and one tile (mx_plan_create),
// built on the host from a prepared frame.
struct HostJob {
    std::vector<double2> zf, zfe_m;
    std::vector<int2> zfe_e;
    std::vector<float2> zf32;
    std::vector<float4> cf;
    std::vector<int2> ce;
    std::vector<int2> lev;
    std::vector<Fe> col, row;
    MxRefDev R;
    MxTileDev T;

    HostJob(const mx_frame_view& v, int W, int H, int max_iter, const mxh_opts& o) {
        const size_t len = (size_t)v.len;
        const size_t nent = v.period > 0 ? len + 2 : len;
        zf.resize(nent);
        zfe_m.resize(nent);
        zfe_e.resize(nent);
        for (size_t k = 0; k < nent; ++k) {
            const size_t s = k % len;
            zf[k] = make_double2(v.zf[2 * s], v.zf[2 * s + 1]);
            zfe_m[k] = make_double2(v.zfe_m[2 * s], v.zfe_m[2 * s + 1]);
            zfe_e[k] = make_int2(v.zfe_e[2 * s], v.zfe_e[2 * s + 1]);
        }
        zf32.resize(nent);
        mx::fe32_orbit(v.zf, len, nent, (float*)zf32.data());
        const int64_t total = v.bla_total;
        cf.resize((size_t)std::max<int64_t>(total, 1));
        ce.resize((size_t)std::max<int64_t>(total, 1));
        if (v.nlev > 0) mx::fe32_coefs(v.bla_coef, total, cf.data(), ce.data());
        lev.assign(MX_MAXLEV, make_int2(0, 0));
        for (int p = 0; p < v.nlev; ++p) lev[p] = make_int2(v.lev_off[p], v.lev_size[p]);
        std::memset(&R, 0, sizeof R);
        R.zf = zf.data();
        R.zfe_m = zfe_m.data();
        R.zfe_e = zfe_e.data();
        R.zf32 = zf32.data();
        R.bla_r2 = v.bla_r2;
        R.bla_coef = (const double4*)v.bla_coef;
        R.bla_coeff = cf.data();
        R.bla_coefe = ce.data();
        R.mx_ref_it = v.period > 0 ? v.period : v.len - 1;
        R.period = v.period > 0 ? v.period : 0;
        R.nent = (int)nent;
        R.nlev = v.nlev;
        R.minlevel = o.bla_minlevel;
        R.rmax2 = v.rmax2;
        col.resize(W);
        row.resize(H);
        for (int j = 0; j < W; ++j) col[j] = fe_make(v.dcol_m[j], v.dcol_e[j]);
        for (int i = 0; i < H; ++i) row[i] = fe_make(v.drow_m[i], v.drow_e[i]);
        std::memset(&T, 0, sizeof T);
        T.dcol = col.data();
        T.drow = row.data();
        T.W = W;
        T.H = H;
        T.stride = 1;
        T.mxit = max_iter;
        T.flags = mx_tile_flags(o.ede != 0, o.interior != 0, o.int_rep != 0, o.bla && v.nlev > 0, v.period > 0);
        T.esc2 = v.esc2;
        T.interior_eps2 = o.interior_eps2;
        T.interior_confirm = o.interior_confirm;
        T.dxfe = fe_make(v.dx_m, v.dx_e);
        T.skip_sample = -1;
        T.chunk = o.chunk > 0 ? std::min(o.chunk, 4096) : 32;
        T.quorum = 32;
    }
};

// Run fn(k) for k in [0, n) on `threads` CPU threads.
template <typename Fn>
void parallel_for(int n, int threads, Fn fn) {
    threads = std::max(1, std::min(threads, n));
    std::vector<std::thread> pool;
    std::atomic<int> next{0};
    for (int t = 0; t < threads; ++t)
        pool.emplace_back([&] {
            for (int k; (k = next.fetch_add(1)) < n;) fn(k);
        });
    for (auto& th : pool) th.join();
}

} // namespace mxh
