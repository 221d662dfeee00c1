// Test infrastructure: the mixed form's per-sample code (mx_fe32.hip, included as
// written; FP32 delta, or its FP64 variant) run on the host; see host_kernels.h.
#include "../mandelbrot-explorer_amd/csrc/mx_fe32.hip"

#include "host_kernels.h"

namespace {

template <typename F, bool EDE, bool INTERIOR>
void run_sample(const MxRefDev& R, const int2* lev, const MxTileDev& T, Fe dcr, Fe dci, bool lean, int& n, double& v,
                C32& c) {
    P32<F> p;
    init32(p, dcr, dci);
    const F esc2f = (F)T.esc2;
    while (!trip32<F, EDE, INTERIOR>(R, lev, T, p, v, n, c, esc2f))
        if (lean) fast_steps<F, EDE, INTERIOR>(R, T, p, c, esc2f, T.chunk);
}

template <typename F>
void run_one(const mxh::HostJob& J, bool ede, bool interior, bool lean, Fe dcr, Fe dci, int& ne, double& vv, C32& c) {
    if (ede && interior)
        run_sample<F, true, true>(J.R, J.lev.data(), J.T, dcr, dci, lean, ne, vv, c);
    else if (ede)
        run_sample<F, true, false>(J.R, J.lev.data(), J.T, dcr, dci, lean, ne, vv, c);
    else if (interior)
        run_sample<F, false, true>(J.R, J.lev.data(), J.T, dcr, dci, lean, ne, vv, c);
    else
        run_sample<F, false, false>(J.R, J.lev.data(), J.T, dcr, dci, lean, ne, vv, c);
}

} // namespace

// Samples idx[0..n) through the mixed form (o->form MX_FORM_FE32) or its FP64-delta
// variant (MX_FORM_SC64); outputs as mxh_run_fe64.
extern "C" int mxh_run_fe32(const mx_frame_view* v, int W, int H, int max_iter, const mxh_opts* o, const int32_t* idx,
                            int n, int32_t* nesc, double* val, uint64_t* stats) {
    if (!v || !o || !idx || !nesc || !val || W < 1 || H < 1 || n < 0) return MX_E_INVALID;
    if (o->form != MX_FORM_FE32 && o->form != MX_FORM_SC64) return MX_E_INVALID;
    const mxh::HostJob J(*v, W, H, max_iter, *o);
    const bool ede = o->ede != 0, interior = (J.T.flags & MX_F_INTERIOR) != 0, lean = o->lean != 0, f32 = o->form == MX_FORM_FE32;
    mxh::parallel_for(n, o->threads, [&](int k) {
        const int s = idx[k], i = s / W, j = s % W;
        C32 c = {0, 0, 0, 0};
        int ne = 0;
        double vv = 0.0;
        const Fe dcr = J.T.dcol[j], dci = J.T.drow[i];
        if (f32)
            run_one<float>(J, ede, interior, lean, dcr, dci, ne, vv, c);
        else
            run_one<double>(J, ede, interior, lean, dcr, dci, ne, vv, c);
        nesc[k] = ne;
        val[k] = vv;
        if (stats) {
            stats[4 * k] = c.steps;
            stats[4 * k + 1] = c.applies;
            stats[4 * k + 2] = c.skipped;
            stats[4 * k + 3] = c.rebases;
        }
    });
    return MX_OK;
}
