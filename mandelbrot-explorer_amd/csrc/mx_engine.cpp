// Device side of libmxhip.so: contexts, reference upload, tile plans and the frame
// driver that mirrors Mandel::Compute's pass structure
// (src/engine/mandel_perturbation.cpp:1656-1761): interior-gate probe on the
// 8x8-strided grid, then the main pass over every sample, then the reference
// pixel's own value.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "mx_internal.h"
#include "mx_tables.h"

namespace {

// one persistent launch over a device job table of a floatexp form
hipError_t launch_jobs(int form, const MxJob* jobs, int njobs, bool ede, bool interior, int blocks, int threads, hipStream_t stream) {
    if (form == MX_FORM_FE32) return mx_launch_fe32(jobs, njobs, ede, interior, blocks, threads, stream);
    if (form == MX_FORM_SC64) return mx_launch_sc64(jobs, njobs, ede, interior, blocks, threads, stream);
    return mx_launch_fe64(jobs, njobs, ede, interior, blocks, threads, stream);
}

double now_s() { return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count(); }

#define MX_HIP(call)                                                                 \
    do {                                                                             \
        hipError_t e_ = (call);                                                      \
        if (e_ != hipSuccess) {                                                      \
            if (ctx) ctx->set_err(std::string(#call) + ": " + hipGetErrorString(e_)); \
            return MX_E_HIP;                                                         \
        }                                                                            \
    } while (0)

size_t align_up(size_t x) { return (x + 255) & ~(size_t)255; }

__global__ void mx_reset_kernel(unsigned* queue, unsigned long long* stats, unsigned* unres_count, double* val, int* nesc, float* out,
                                size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < MX_NQUEUE) queue[i] = 0u;
    if (i == 0 && unres_count) *unres_count = 0u;
    if (i < 4) stats[i] = 0ull;
    for (size_t k = i; k < n; k += (size_t)gridDim.x * blockDim.x) {
        val[k] = -10.0;
        nesc[k] = INT_MIN;
        if (out) out[k] = -10.0f;
    }
}

// The scalar fields of a tile's device record (sizes, limits, flags, launch knobs);
// the caller sets its buffer pointers.  W x H: the grid, or n samples x 1.
void tile_fields(MxTileDev& T, const mx_tile& t, const mx_ref& ref, size_t W, size_t H, bool sdc, bool dbl) {
    std::memset(&T, 0, sizeof T);
    T.W = (int)W;
    T.H = (int)H;
    T.stride = sdc ? 1 : t.stride;
    T.nx = (T.W + T.stride - 1) / T.stride;
    T.ny = (T.H + T.stride - 1) / T.stride;
    T.tiles_x = (T.nx + 7) / 8;
    T.ntiles = sdc ? (int)((W * H + 63) / 64) : T.tiles_x * ((T.ny + 7) / 8);
    T.mxit = t.max_iter;
    T.flags = dbl ? (t.ede ? MX_F_EDE : 0) | (t.interior ? MX_F_INTERIOR : 0) | (t.int_rep ? MX_F_INT_REP : 0) | (t.bla ? MX_F_BLA : 0)
                  : mx_tile_flags(t.ede != 0, t.interior != 0, t.int_rep != 0, t.bla != 0, ref.dev.period > 0);
    T.esc2 = t.esc2;
    T.interior_eps2 = t.interior_eps2;
    T.interior_confirm = t.interior_confirm;
    T.dxfe = fe_make(t.dx_m, t.dx_e);
    T.skip_sample = t.skip_sample;
    T.chunk = t.chunk > 0 ? std::min(t.chunk, 4096) : 32;
    T.quorum = t.quorum > 0 ? std::min(t.quorum, 64) : 32;
}

// Fixed scratch area of mx_tile_render_noalloc, one copy per device (held until the
// process ends).  Host code never uses its name as an address: the device address
// comes from hipGetSymbolAddress at context creation (mx_ctx::noalloc_scratch).
// Layout: the job record, the work queues and counters, the dc tables as mantissas,
// exponents and packed FloatExp values, and the outputs (value, n_esc, engine float).
constexpr size_t NA_JOB = 0;
constexpr size_t NA_QUEUE = NA_JOB + ((sizeof(MxJob) + 255) & ~(size_t)255);
constexpr size_t NA_STATS = NA_QUEUE + 256;
constexpr size_t NA_DCM = NA_STATS + 256;
constexpr size_t NA_DCE = NA_DCM + 8 * (size_t)MX_NOALLOC_MAX_TABLE;
constexpr size_t NA_DCF = NA_DCE + 4 * (size_t)MX_NOALLOC_MAX_TABLE;
constexpr size_t NA_VAL = NA_DCF + sizeof(Fe) * (size_t)MX_NOALLOC_MAX_TABLE;
constexpr size_t NA_NESC = NA_VAL + 8 * (size_t)MX_NOALLOC_MAX_SAMPLES;
constexpr size_t NA_OUT = NA_NESC + 4 * (size_t)MX_NOALLOC_MAX_SAMPLES;
constexpr size_t NA_BYTES = NA_OUT + 4 * (size_t)MX_NOALLOC_MAX_SAMPLES;
static __device__ __align__(256) unsigned char scratch[NA_BYTES];

// dc tables from separate mantissa and exponent arrays into FloatExp records
__global__ void mx_pack_fe_kernel(const double* m, const int* e, Fe* out, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = fe_make(m[i], e[i]);
}

struct BusyGuard {
    mx_ctx* ctx;
    bool ok;
    explicit BusyGuard(mx_ctx* c) : ctx(c) {
        bool expected = false;
        ok = ctx->busy.compare_exchange_strong(expected, true);
    }
    ~BusyGuard() {
        if (ok) ctx->busy.store(false);
    }
};

} // namespace

extern "C" int mx_version(void) { return 1; }

extern "C" int mx_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

extern "C" int mx_ctx_create(int device, mx_ctx** out) {
    if (!out) return MX_E_INVALID;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return MX_E_NODEVICE;
    if (device < 0 || device >= n) return MX_E_INVALID;
    auto* ctx = new (std::nothrow) mx_ctx();
    if (!ctx) return MX_E_NOMEM;
    ctx->device = device;
    auto fail = [&](int rc) {
        mx_ctx_destroy(ctx);
        return rc;
    };
    if (hipSetDevice(device) != hipSuccess) return fail(MX_E_HIP);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return fail(MX_E_HIP);
    ctx->cu_count = prop.multiProcessorCount;
    ctx->lds_max = (int)prop.sharedMemPerBlock;
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) return fail(MX_E_HIP);
    if (hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking) != hipSuccess) return fail(MX_E_HIP);
    if (hipHostMalloc((void**)&ctx->cancel_host, sizeof(int), hipHostMallocMapped) != hipSuccess) return fail(MX_E_HIP);
    *ctx->cancel_host = 0;
    if (hipHostGetDevicePointer((void**)&ctx->cancel_dev, ctx->cancel_host, 0) != hipSuccess) return fail(MX_E_HIP);
    if (hipHostMalloc((void**)&ctx->abort_host, sizeof(int), hipHostMallocMapped) != hipSuccess) return fail(MX_E_HIP);
    *ctx->abort_host = 0;
    if (hipHostGetDevicePointer((void**)&ctx->abort_dev, ctx->abort_host, 0) != hipSuccess) return fail(MX_E_HIP);
    // this device's copy of the no-alloc scratch area (the device is current)
    if (hipGetSymbolAddress((void**)&ctx->noalloc_scratch, HIP_SYMBOL(scratch)) != hipSuccess) {
        (void)hipGetLastError();
        ctx->noalloc_scratch = nullptr;
        ctx->set_err("no scratch area for mx_tile_render_noalloc on this device");
    }
    *out = ctx;
    return MX_OK;
}

extern "C" void mx_ctx_destroy(mx_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) {
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipStreamDestroy(ctx->stream);
    }
    if (ctx->side) {
        (void)hipStreamSynchronize(ctx->side);
        (void)hipStreamDestroy(ctx->side);
    }
    if (ctx->cancel_host) (void)hipHostFree(ctx->cancel_host);
    if (ctx->abort_host) (void)hipHostFree(ctx->abort_host);
    delete ctx;
}

extern "C" int mx_ctx_cu_count(const mx_ctx* ctx) { return ctx ? ctx->cu_count : MX_E_INVALID; }

extern "C" const char* mx_last_error(const mx_ctx* ctx) { return ctx ? ctx->err.c_str() : "no context"; }

extern "C" void mx_cancel(mx_ctx* ctx) {
    if (!ctx || !ctx->cancel_host) return;
    __atomic_store_n(ctx->cancel_host, 1, __ATOMIC_RELEASE);
    // a pass on the side stream watches the abort flag: it stops as well
    if (ctx->abort_host) __atomic_store_n(ctx->abort_host, 1, __ATOMIC_RELEASE);
}

extern "C" void mx_reset_cancellation(mx_ctx* ctx) {
    if (!ctx || !ctx->cancel_host) return;
    if (!ctx->busy.load()) {
        __atomic_store_n(ctx->cancel_host, 0, __ATOMIC_RELEASE);
        if (ctx->abort_host) __atomic_store_n(ctx->abort_host, 0, __ATOMIC_RELEASE);
    }
}

extern "C" int mx_ref_upload(mx_ctx* ctx, const mx_ref_desc* d, mx_ref** out) {
    if (!ctx || !d || !out || d->len < 1 || !d->zf || !d->zfe_m || !d->zfe_e || d->nlev < 0 || d->nlev > MX_MAXLEV) return MX_E_INVALID;
    *out = nullptr;
    int64_t total = 0;
    for (int p = 0; p < d->nlev; ++p) {
        if (!d->lev_off || !d->lev_size || d->lev_off[p] != total || d->lev_size[p] < 0) return MX_E_INVALID;
        total += d->lev_size[p];
    }
    if (d->nlev > 0 && (!d->bla_r2 || !d->bla_coef)) return MX_E_INVALID;
    if (total > INT_MAX / 4 || (int64_t)d->len > INT_MAX / 2) return MX_E_INVALID;
    // the kernels address the orbit with 32-bit byte offsets (16 B per entry)
    if (d->len > (1 << 27)) {
        ctx->set_err("reference orbit longer than 2^27 entries");
        return MX_E_INVALID;
    }
    if (d->z_m3 && (!d->sa_coef || d->sa_order < 0 || d->sa_order >= MX_SA_N || d->sa_it < 0)) return MX_E_INVALID;
    if (d->df && (!d->z_m3 || !d->sa_coef_b || (d->nlev > 0 && !d->bla_der))) return MX_E_INVALID;
    if (d->period < 0 || (d->period > 0 && (d->period != d->len || d->len < 3 || d->z_m3))) {
        ctx->set_err("a periodic reference holds exactly one period (len == period >= 3) on a floatexp path");
        return MX_E_INVALID;
    }
    MX_HIP(hipSetDevice(ctx->device));
    const size_t len = (size_t)d->len;
    // a periodic orbit is stored with copies of its first two entries after the last,
    // so the kernels read X at rm + 1 and rm + 2 across the wrap without a modulo
    const size_t nent = d->period > 0 ? len + 2 : len;
    size_t o_zf = 0;
    size_t o_zfm = align_up(o_zf + 16 * nent);
    size_t o_zfe = align_up(o_zfm + 16 * nent);
    size_t o_zf32 = align_up(o_zfe + 8 * nent);
    size_t o_r2 = align_up(o_zf32 + 8 * nent);
    size_t o_coef = align_up(o_r2 + 8 * (size_t)total);
    size_t o_cf = align_up(o_coef + 32 * (size_t)total);
    size_t o_ce = align_up(o_cf + 16 * (size_t)total);
    size_t o_lev = align_up(o_ce + 8 * (size_t)total);
    size_t o_m3 = align_up(o_lev + 8 * MX_MAXLEV);
    size_t o_sa = align_up(o_m3 + (d->z_m3 ? 8 * len : 0));
    size_t o_df = align_up(o_sa + 16 * MX_SA_N);
    size_t o_der = align_up(o_df + (d->df ? 16 * len : 0));
    size_t o_sab = align_up(o_der + (d->df ? 32 * (size_t)total : 0));
    size_t bytes = align_up(o_sab + 16 * MX_SA_N);
    auto* ref = new (std::nothrow) mx_ref();
    if (!ref) return MX_E_NOMEM;
    ref->ctx = ctx;
    ref->bytes = bytes;
    if (hipMalloc(&ref->block, bytes) != hipSuccess) {
        delete ref;
        ctx->set_err("hipMalloc of the reference failed");
        return MX_E_NOMEM;
    }
    char* b = (char*)ref->block;
    // the mixed form's FP32 orbit and FP32-floatexp coefficients
    std::vector<float> zf32(2 * nent);
    mx::fe32_orbit(d->zf, len, nent, zf32.data());
    std::vector<float4> cf((size_t)total);
    std::vector<int2> ce((size_t)total);
    if (d->nlev > 0) mx::fe32_coefs(d->bla_coef, total, cf.data(), ce.data());
    std::vector<int2> lev(MX_MAXLEV, make_int2(0, 0));
    for (int p = 0; p < d->nlev; ++p) lev[p] = make_int2(d->lev_off[p], d->lev_size[p]);
    hipError_t e = hipSuccess;
    auto cp = [&](size_t off, const void* src, size_t n) {
        if (e == hipSuccess && n) e = hipMemcpyAsync(b + off, src, n, hipMemcpyHostToDevice, ctx->stream);
    };
    cp(o_zf, d->zf, 16 * len);
    cp(o_zfm, d->zfe_m, 16 * len);
    cp(o_zfe, d->zfe_e, 8 * len);
    if (nent > len) {
        cp(o_zf + 16 * len, d->zf, 16 * (nent - len));
        cp(o_zfm + 16 * len, d->zfe_m, 16 * (nent - len));
        cp(o_zfe + 8 * len, d->zfe_e, 8 * (nent - len));
    }
    cp(o_zf32, zf32.data(), 8 * nent);
    if (d->nlev > 0) {
        cp(o_r2, d->bla_r2, 8 * (size_t)total);
        cp(o_coef, d->bla_coef, 32 * (size_t)total);
        cp(o_cf, cf.data(), 16 * (size_t)total);
        cp(o_ce, ce.data(), 8 * (size_t)total);
    }
    cp(o_lev, lev.data(), 8 * MX_MAXLEV);
    if (d->z_m3) {
        cp(o_m3, d->z_m3, 8 * len);
        cp(o_sa, d->sa_coef, 16 * MX_SA_N);
    }
    if (d->df) {
        cp(o_df, d->df, 16 * len);
        if (d->nlev > 0) cp(o_der, d->bla_der, 32 * (size_t)total);
        cp(o_sab, d->sa_coef_b, 16 * MX_SA_N);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        ctx->set_err(std::string("reference upload: ") + hipGetErrorString(e));
        (void)hipFree(ref->block);
        delete ref;
        return MX_E_HIP;
    }
    MxRefDev& R = ref->dev;
    std::memset(&R, 0, sizeof R);
    R.zf = (const double2*)(b + o_zf);
    R.zfe_m = (const double2*)(b + o_zfm);
    R.zfe_e = (const int2*)(b + o_zfe);
    R.zf32 = (const float2*)(b + o_zf32);
    R.bla_r2 = (const double*)(b + o_r2);
    R.bla_coef = (const double4*)(b + o_coef);
    R.bla_coeff = (const float4*)(b + o_cf);
    R.bla_coefe = (const int2*)(b + o_ce);
    R.mx_ref_it = d->period > 0 ? d->period : d->len - 1;
    R.period = d->period > 0 ? d->period : 0;
    R.nent = (int)nent;
    R.nlev = d->nlev;
    R.rmax2 = d->rmax2;
    if (d->z_m3) {
        R.z_m3 = (const double*)(b + o_m3);
        R.sa_A = (const double2*)(b + o_sa);
        R.sa_delta = make_double2(d->sa_delta_re, d->sa_delta_im);
        R.sa_it = d->sa_it;
        R.sa_order = d->sa_order;
    }
    if (d->df) {
        R.df = (const double2*)(b + o_df);
        R.bla_der = (const double4*)(b + o_der);
        R.sa_B = (const double2*)(b + o_sab);
    }
    ref->lev = (int2*)(b + o_lev);
    *out = ref;
    return MX_OK;
}

extern "C" void mx_ref_release(mx_ref* ref) {
    if (!ref) return;
    if (ref->block) {
        (void)hipSetDevice(ref->ctx->device);
        (void)hipStreamSynchronize(ref->ctx->stream);
        (void)hipFree(ref->block);
    }
    delete ref;
}

extern "C" int mx_plan_create(mx_ctx* ctx, const mx_ref* ref, const mx_tile* t, mx_plan** out) {
    if (!ctx || !ref || !t || !out || t->max_iter < 1) return MX_E_INVALID;
    if (!t->sample_dc_m && (t->width < 1 || t->height < 1 || t->stride < 1)) return MX_E_INVALID;
    if (t->sample_dc_m && t->n_samples > INT_MAX / 2) return MX_E_INVALID;
    const bool dbl = t->form == MX_FORM_DOUBLE;
    if (t->form != MX_FORM_FE64 && t->form != MX_FORM_FE32 && t->form != MX_FORM_SC64 && !dbl) {
        ctx->set_err("unknown kernel form");
        return MX_E_INVALID;
    }
    const bool sdc = t->sample_dc_m != nullptr;
    if (sdc && (dbl || !t->sample_dc_e || t->n_samples < 1 || t->list)) {
        ctx->set_err("explicit samples need a floatexp form, both dc arrays and n_samples >= 1");
        return MX_E_INVALID;
    }
    if (!dbl && !sdc && (!t->dcol_m || !t->dcol_e || !t->drow_m || !t->drow_e)) return MX_E_INVALID;
    if (dbl && (!t->dcol_d || !t->drow_d || !ref->dev.z_m3 || (t->ede && !ref->dev.df))) {
        ctx->set_err("the plain-double form needs double dc tables and a reference with its series "
                     "(and its derivative under EDE)");
        return MX_E_INVALID;
    }
    if (t->list && (t->list_len < 0 || t->list_len > t->width * t->height)) return MX_E_INVALID;
    if (t->list)
        for (int32_t q = 0; q < t->list_len; ++q)
            if (t->list[q] < 0 || (int64_t)t->list[q] >= (int64_t)t->width * t->height) return MX_E_INVALID;
    if (t->bla && (ref->dev.nlev == 0 || t->bla_minlevel < 0)) return MX_E_INVALID;
    if (!t->sample_dc_m && (int64_t)t->width * t->height > (int64_t)INT_MAX / 2) return MX_E_INVALID;
    *out = nullptr;
    MX_HIP(hipSetDevice(ctx->device));
    auto* plan = new (std::nothrow) mx_plan();
    if (!plan) return MX_E_NOMEM;
    plan->ctx = ctx;
    plan->stream = ctx->stream;
    plan->ref = ref;
    plan->form = t->form;
    const size_t W = sdc ? (size_t)t->n_samples : (size_t)t->width, H = sdc ? 1 : (size_t)t->height, n = W * H;
    plan->n = n;
    const size_t nlist = t->list ? (size_t)t->list_len : 0;
    // unresolved samples: at most one per sample of the launch
    const size_t nunres = dbl ? (t->list ? nlist : n) : 0;
    size_t o_col = 0;
    size_t o_row = align_up(o_col + sizeof(Fe) * W);
    size_t o_val = align_up(o_row + sizeof(Fe) * H);
    size_t o_nesc = align_up(o_val + 8 * n);
    size_t o_out = align_up(o_nesc + 4 * n);
    size_t o_q = align_up(o_out + 4 * n);
    size_t o_st = align_up(o_q + 4 * MX_NQUEUE);
    size_t o_dcd = align_up(o_st + 8 * 4);
    size_t o_drd = align_up(o_dcd + (dbl ? 8 * W : 0));
    size_t o_list = align_up(o_drd + (dbl ? 8 * H : 0));
    size_t o_uidx = align_up(o_list + 4 * nlist);
    size_t o_ugp = align_up(o_uidx + 4 * nunres);
    size_t o_ucnt = align_up(o_ugp + 8 * nunres);
    size_t o_sdc = align_up(o_ucnt + 4);
    size_t o_job = align_up(o_sdc + (sdc ? sizeof(Fe) * 2 * n : 0));
    size_t bytes = align_up(o_job + sizeof(MxJob));
    if (hipMalloc(&plan->block, bytes) != hipSuccess) {
        delete plan;
        ctx->set_err("hipMalloc of the tile failed");
        return MX_E_NOMEM;
    }
    plan->bytes = bytes;
    char* b = (char*)plan->block;
    std::vector<Fe> col(W, Fe{0.0, 0}), row(H, Fe{0.0, 0});
    std::vector<Fe> sv(sdc ? 2 * n : 0);
    for (size_t q = 0; q < sv.size(); ++q) sv[q] = fe_make(t->sample_dc_m[q], t->sample_dc_e[q]);
    if (!dbl && !sdc) {
        for (size_t j = 0; j < W; ++j) col[j] = fe_make(t->dcol_m[j], t->dcol_e[j]);
        for (size_t i = 0; i < H; ++i) row[i] = fe_make(t->drow_m[i], t->drow_e[i]);
    }
    hipError_t e = hipMemcpyAsync(b + o_col, col.data(), sizeof(Fe) * W, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(b + o_row, row.data(), sizeof(Fe) * H, hipMemcpyHostToDevice, ctx->stream);
    if (dbl && e == hipSuccess) e = hipMemcpyAsync(b + o_dcd, t->dcol_d, 8 * W, hipMemcpyHostToDevice, ctx->stream);
    if (dbl && e == hipSuccess) e = hipMemcpyAsync(b + o_drd, t->drow_d, 8 * H, hipMemcpyHostToDevice, ctx->stream);
    if (nlist && e == hipSuccess) e = hipMemcpyAsync(b + o_list, t->list, 4 * nlist, hipMemcpyHostToDevice, ctx->stream);
    if (sdc && e == hipSuccess) e = hipMemcpyAsync(b + o_sdc, sv.data(), sizeof(Fe) * sv.size(), hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess) e = hipEventCreate(&plan->ev0);
    if (e == hipSuccess) e = hipEventCreate(&plan->ev1);
    if (e != hipSuccess) {
        ctx->set_err(std::string("tile set-up: ") + hipGetErrorString(e));
        mx_plan_destroy(plan);
        return MX_E_HIP;
    }
    MxTileDev& T = plan->tile;
    tile_fields(T, *t, *ref, W, H, sdc, dbl);
    T.dcol = (const Fe*)(b + o_col);
    T.drow = (const Fe*)(b + o_row);
    T.val = (double*)(b + o_val);
    T.nesc = (int*)(b + o_nesc);
    T.out = (float*)(b + o_out);
    T.queue = (unsigned*)(b + o_q);
    T.stats = (unsigned long long*)(b + o_st);
    T.cancel = ctx->cancel_dev;
    if (dbl) {
        T.dcol_d = (const double*)(b + o_dcd);
        T.drow_d = (const double*)(b + o_drd);
        T.unres_idx = (int*)(b + o_uidx);
        T.unres_gp = (double*)(b + o_ugp);
        T.unres_count = (unsigned*)(b + o_ucnt);
    }
    if (nlist) {
        T.list = (const int*)(b + o_list);
        T.list_len = (int)nlist;
    }
    if (sdc) {
        T.sdc = (const Fe*)(b + o_sdc);
        T.list_len = (int)n;
        T.skip_sample = -1;
    }
    plan->threads = 256;
    plan->minlevel = t->bla_minlevel;
    plan->ede = t->ede != 0;
    plan->grid_request = t->grid_blocks;
    plan->interior = (T.flags & MX_F_INTERIOR) != 0;
    if (!dbl) {
        // the plan's launch is a one-job table resident beside its tile
        MxJob J;
        J.R = ref->dev;
        J.R.minlevel = t->bla_minlevel;
        J.lev = ref->lev;
        J.T = T;
        plan->job = J;
        plan->job_dev = (MxJob*)(b + o_job);
        e = hipMemcpy(plan->job_dev, &J, sizeof J, hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            ctx->set_err(std::string("tile set-up: ") + hipGetErrorString(e));
            mx_plan_destroy(plan);
            return MX_E_HIP;
        }
    }
    // persistent grid: 8 waves per SIMD (32 per CU); never more workgroups than
    // there are 8x8 tiles to hand out
    const int waves = ctx->cu_count * 32;
    const int work_tiles = t->list ? (int)((nlist + 63) / 64) : T.ntiles;
    plan->work_tiles = work_tiles;
    plan->blocks = std::max(1, std::min(t->grid_blocks > 0 ? t->grid_blocks : waves / 4, work_tiles));
    if (dbl && (t->lds_stage > 0 || (t->lds_stage == 0 && !(t->ede && plan->interior)))) {
        // the orbit in LDS: one workgroup of 16 waves per CU
        plan->nstage = std::min(ref->dev.nent, ctx->lds_max / (int)sizeof(double2));
        if (plan->nstage > 0) {
            plan->threads = 1024;
            plan->blocks = std::max(1, std::min(t->grid_blocks > 0 ? t->grid_blocks : ctx->cu_count, (work_tiles + 15) / 16));
        }
    }
    *out = plan;
    return MX_OK;
}

extern "C" int mx_plan_launch(mx_plan* plan) {
    if (!plan) return MX_E_INVALID;
    mx_ctx* ctx = plan->ctx;
    MX_HIP(hipSetDevice(ctx->device));
    MxTileDev& T = plan->tile;
    MX_HIP(hipEventRecord(plan->ev0, plan->stream));
    hipLaunchKernelGGL(mx_reset_kernel, dim3(1024), dim3(256), 0, plan->stream, T.queue, T.stats, T.unres_count, T.val, T.nesc, T.out, plan->n);
    MX_HIP(hipGetLastError());
    if (plan->form == MX_FORM_DOUBLE) {
        MxRefDev R = plan->ref->dev;
        R.minlevel = plan->minlevel;
        MX_HIP(mx_launch_dbl(R, plan->ref->lev, T, plan->blocks, plan->threads, plan->stream, plan->nstage));
    } else {
        MX_HIP(launch_jobs(plan->form, plan->job_dev, 1, plan->ede, plan->interior, plan->blocks, plan->threads, plan->stream));
    }
    MX_HIP(hipEventRecord(plan->ev1, plan->stream));
    return MX_OK;
}

extern "C" int mx_plan_wait(mx_plan* plan, double* kernel_ms) {
    if (!plan) return MX_E_INVALID;
    mx_ctx* ctx = plan->ctx;
    MX_HIP(hipStreamSynchronize(plan->stream));
    if (kernel_ms) {
        float ms = 0.f;
        MX_HIP(hipEventElapsedTime(&ms, plan->ev0, plan->ev1));
        *kernel_ms = ms;
    }
    if (__atomic_load_n(ctx->cancel_host, __ATOMIC_ACQUIRE)) {
        ctx->set_err("cancelled");
        return MX_E_CANCELLED;
    }
    return MX_OK;
}

extern "C" int mx_plan_read(mx_plan* plan, int32_t* n_esc, double* value, float* out, uint64_t* stats4) {
    if (!plan) return MX_E_INVALID;
    mx_ctx* ctx = plan->ctx;
    MxTileDev& T = plan->tile;
    if (n_esc) MX_HIP(hipMemcpyAsync(n_esc, T.nesc, 4 * plan->n, hipMemcpyDeviceToHost, plan->stream));
    if (value) MX_HIP(hipMemcpyAsync(value, T.val, 8 * plan->n, hipMemcpyDeviceToHost, plan->stream));
    if (out) MX_HIP(hipMemcpyAsync(out, T.out, 4 * plan->n, hipMemcpyDeviceToHost, plan->stream));
    if (stats4) MX_HIP(hipMemcpyAsync(stats4, T.stats, 32, hipMemcpyDeviceToHost, plan->stream));
    MX_HIP(hipStreamSynchronize(plan->stream));
    return MX_OK;
}

extern "C" int mx_plan_unresolved(mx_plan* plan, int32_t* idx, double* glitch, int32_t cap, int32_t* count) {
    if (!plan || !count || cap < 0) return MX_E_INVALID;
    mx_ctx* ctx = plan->ctx;
    const MxTileDev& T = plan->tile;
    *count = 0;
    if (!T.unres_count) return MX_OK;
    unsigned c = 0;
    MX_HIP(hipMemcpyAsync(&c, T.unres_count, 4, hipMemcpyDeviceToHost, plan->stream));
    MX_HIP(hipStreamSynchronize(plan->stream));
    *count = (int32_t)c;
    const size_t k = std::min((size_t)c, (size_t)cap);
    if (k && idx) MX_HIP(hipMemcpyAsync(idx, T.unres_idx, 4 * k, hipMemcpyDeviceToHost, plan->stream));
    if (k && glitch) MX_HIP(hipMemcpyAsync(glitch, T.unres_gp, 8 * k, hipMemcpyDeviceToHost, plan->stream));
    MX_HIP(hipStreamSynchronize(plan->stream));
    return MX_OK;
}

extern "C" void mx_plan_destroy(mx_plan* plan) {
    if (!plan) return;
    (void)hipSetDevice(plan->ctx->device);
    if (plan->stream) (void)hipStreamSynchronize(plan->stream);
    if (plan->ev0) (void)hipEventDestroy(plan->ev0);
    if (plan->ev1) (void)hipEventDestroy(plan->ev1);
    if (plan->block) (void)hipFree(plan->block);
    delete plan;
}

extern "C" int mx_batch_create(mx_ctx* ctx, mx_plan* const* plans, int32_t n, mx_batch** out) {
    if (!ctx || !plans || n < 1 || !out) return MX_E_INVALID;
    *out = nullptr;
    const mx_plan* p0 = plans[0];
    if (!p0) return MX_E_INVALID;
    int64_t work = 0;
    for (int32_t k = 0; k < n; ++k) {
        const mx_plan* p = plans[k];
        if (!p || p->ctx != ctx || !p->job_dev) {
            ctx->set_err("a batch takes floatexp-form plans of this context");
            return MX_E_INVALID;
        }
        if (p->form != p0->form || p->ede != p0->ede) {
            ctx->set_err("the plans of a batch share one form and the same EDE flag");
            return MX_E_INVALID;
        }
        for (int32_t m = 0; m < k; ++m)
            if (plans[m] == p) {
                ctx->set_err("a plan appears twice in the batch");
                return MX_E_INVALID;
            }
        work += p->work_tiles;
    }
    MX_HIP(hipSetDevice(ctx->device));
    auto* b = new (std::nothrow) mx_batch();
    if (!b) return MX_E_NOMEM;
    b->ctx = ctx;
    b->plans.assign(plans, plans + n);
    b->form = p0->form;
    b->ede = p0->ede;
    // the interior-detection kernel variant when any plan detects interiors (a job
    // without the flag skips those tests at run time)
    b->interior = false;
    for (int32_t k = 0; k < n; ++k) b->interior = b->interior || plans[k]->interior;
    std::vector<MxJob> jobs((size_t)n);
    for (int32_t k = 0; k < n; ++k) jobs[(size_t)k] = plans[k]->job;
    hipError_t e = hipMalloc(&b->jobs_dev, sizeof(MxJob) * (size_t)n);
    if (e == hipSuccess) e = hipMemcpy(b->jobs_dev, jobs.data(), sizeof(MxJob) * (size_t)n, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipEventCreate(&b->ev0);
    if (e == hipSuccess) e = hipEventCreate(&b->ev1);
    if (e != hipSuccess) {
        ctx->set_err(std::string("batch set-up: ") + hipGetErrorString(e));
        mx_batch_destroy(b);
        return MX_E_HIP;
    }
    b->threads = 256;
    // the first plan's grid size request, if any, sizes the batch's grid
    const int want = p0->grid_request > 0 ? p0->grid_request : ctx->cu_count * 32 / 4;
    b->blocks = (int)std::max<int64_t>(1, std::min<int64_t>(want, work));
    *out = b;
    return MX_OK;
}

extern "C" int mx_batch_launch(mx_batch* b) {
    if (!b) return MX_E_INVALID;
    mx_ctx* ctx = b->ctx;
    MX_HIP(hipSetDevice(ctx->device));
    MX_HIP(hipEventRecord(b->ev0, ctx->stream));
    for (mx_plan* p : b->plans) {
        const MxTileDev& T = p->tile;
        hipLaunchKernelGGL(mx_reset_kernel, dim3(1024), dim3(256), 0, ctx->stream, T.queue, T.stats, T.unres_count, T.val, T.nesc, T.out, p->n);
        MX_HIP(hipGetLastError());
    }
    MX_HIP(launch_jobs(b->form, b->jobs_dev, (int)b->plans.size(), b->ede, b->interior, b->blocks, b->threads, ctx->stream));
    MX_HIP(hipEventRecord(b->ev1, ctx->stream));
    return MX_OK;
}

extern "C" int mx_batch_wait(mx_batch* b, double* kernel_ms) {
    if (!b) return MX_E_INVALID;
    mx_ctx* ctx = b->ctx;
    MX_HIP(hipStreamSynchronize(ctx->stream));
    if (kernel_ms) {
        float ms = 0.f;
        MX_HIP(hipEventElapsedTime(&ms, b->ev0, b->ev1));
        *kernel_ms = ms;
    }
    if (__atomic_load_n(ctx->cancel_host, __ATOMIC_ACQUIRE)) {
        ctx->set_err("cancelled");
        return MX_E_CANCELLED;
    }
    return MX_OK;
}

extern "C" void mx_batch_destroy(mx_batch* b) {
    if (!b) return;
    (void)hipSetDevice(b->ctx->device);
    (void)hipStreamSynchronize(b->ctx->stream);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->jobs_dev) (void)hipFree(b->jobs_dev);
    delete b;
}

extern "C" int mx_graph_create(mx_ctx* ctx, mx_plan* const* plans, int32_t n, mx_graph** out) {
    if (!ctx || !plans || n < 1 || !out) return MX_E_INVALID;
    *out = nullptr;
    for (int32_t k = 0; k < n; ++k)
        if (!plans[k] || plans[k]->ctx != ctx || !plans[k]->job_dev) {
            ctx->set_err("a graph takes floatexp-form plans of this context");
            return MX_E_INVALID;
        }
    MX_HIP(hipSetDevice(ctx->device));
    auto* g = new (std::nothrow) mx_graph();
    if (!g) return MX_E_NOMEM;
    g->ctx = ctx;
    g->plans.assign(plans, plans + n);
    // every plan's reset (a kernel: no memset node) and launch, in order, on the
    // context's stream
    hipError_t e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal);
    hipError_t el = hipSuccess;
    if (e == hipSuccess) {
        for (mx_plan* p : g->plans) {
            const MxTileDev& T = p->tile;
            hipLaunchKernelGGL(mx_reset_kernel, dim3(1024), dim3(256), 0, ctx->stream, T.queue, T.stats, T.unres_count, T.val, T.nesc, T.out, p->n);
            el = hipGetLastError();
            if (el != hipSuccess) break;
            el = launch_jobs(p->form, p->job_dev, 1, p->ede, p->interior, p->blocks, p->threads, ctx->stream);
            if (el != hipSuccess) break;
        }
        e = hipStreamEndCapture(ctx->stream, &g->graph);
        if (e == hipSuccess) e = el;
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
    if (e == hipSuccess) e = hipEventCreate(&g->ev0);
    if (e == hipSuccess) e = hipEventCreate(&g->ev1);
    if (e != hipSuccess) {
        ctx->set_err(std::string("graph capture: ") + hipGetErrorString(e));
        mx_graph_destroy(g);
        return MX_E_HIP;
    }
    *out = g;
    return MX_OK;
}

extern "C" int mx_graph_launch(mx_graph* g) {
    if (!g) return MX_E_INVALID;
    mx_ctx* ctx = g->ctx;
    MX_HIP(hipSetDevice(ctx->device));
    MX_HIP(hipEventRecord(g->ev0, ctx->stream));
    MX_HIP(hipGraphLaunch(g->exec, ctx->stream));
    MX_HIP(hipEventRecord(g->ev1, ctx->stream));
    return MX_OK;
}

extern "C" int mx_graph_wait(mx_graph* g, double* ms) {
    if (!g) return MX_E_INVALID;
    mx_ctx* ctx = g->ctx;
    MX_HIP(hipStreamSynchronize(ctx->stream));
    if (ms) {
        float f = 0.f;
        MX_HIP(hipEventElapsedTime(&f, g->ev0, g->ev1));
        *ms = f;
    }
    if (__atomic_load_n(ctx->cancel_host, __ATOMIC_ACQUIRE)) {
        ctx->set_err("cancelled");
        return MX_E_CANCELLED;
    }
    return MX_OK;
}

extern "C" void mx_graph_destroy(mx_graph* g) {
    if (!g) return;
    (void)hipSetDevice(g->ctx->device);
    (void)hipStreamSynchronize(g->ctx->stream);
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    if (g->ev0) (void)hipEventDestroy(g->ev0);
    if (g->ev1) (void)hipEventDestroy(g->ev1);
    delete g;
}

extern "C" int mx_tile_render(mx_ctx* ctx, const mx_ref* ref, const mx_tile* tile, int32_t* n_esc, double* value, float* out, uint64_t* stats4) {
    if (!ctx) return MX_E_INVALID;
    BusyGuard busy(ctx);
    if (!busy.ok) return MX_E_BUSY;
    mx_plan* plan = nullptr;
    int rc = mx_plan_create(ctx, ref, tile, &plan);
    if (rc != MX_OK) return rc;
    rc = mx_plan_launch(plan);
    if (rc == MX_OK) rc = mx_plan_wait(plan, nullptr);
    if (rc == MX_OK || rc == MX_E_CANCELLED) {
        int rc2 = mx_plan_read(plan, n_esc, value, out, stats4);
        if (rc == MX_OK) rc = rc2;
    }
    mx_plan_destroy(plan);
    return rc;
}

extern "C" int mx_tile_render_noalloc(mx_ctx* ctx, const mx_ref* ref, const mx_tile* t, int32_t* n_esc, double* value, float* out,
                                      uint64_t* stats4) {
    if (!ctx || !ref || !t || ref->ctx != ctx) return MX_E_INVALID;
    BusyGuard busy(ctx);
    if (!busy.ok) return MX_E_BUSY;
    if (t->form != MX_FORM_FE64 && t->form != MX_FORM_FE32 && t->form != MX_FORM_SC64) {
        ctx->set_err("mx_tile_render_noalloc takes a floatexp form");
        return MX_E_INVALID;
    }
    if (t->list || t->sample_dc_m || t->width < 1 || t->height < 1 || t->stride < 1 || t->max_iter < 1 || !t->dcol_m ||
        !t->dcol_e || !t->drow_m || !t->drow_e || (t->bla && (ref->dev.nlev == 0 || t->bla_minlevel < 0))) {
        ctx->set_err("mx_tile_render_noalloc takes a grid tile with its dc tables");
        return MX_E_INVALID;
    }
    const size_t W = (size_t)t->width, H = (size_t)t->height, n = W * H;
    if (n > (size_t)MX_NOALLOC_MAX_SAMPLES || W + H > (size_t)MX_NOALLOC_MAX_TABLE) {
        ctx->set_err("tile beyond the no-alloc scratch limits (" + std::to_string(MX_NOALLOC_MAX_SAMPLES) + " samples, " +
                     std::to_string(MX_NOALLOC_MAX_TABLE) + " columns + rows)");
        return MX_E_INVALID;
    }
    if (!ctx->noalloc_scratch) {
        ctx->set_err("no scratch area for mx_tile_render_noalloc on this device");
        return MX_E_HIP;
    }
    if (__atomic_load_n(ctx->cancel_host, __ATOMIC_ACQUIRE)) {
        ctx->set_err("cancelled");
        return MX_E_CANCELLED;
    }
    unsigned char* b = ctx->noalloc_scratch;
    MX_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    // dc tables: columns then rows, packed into FloatExp records on the device
    MX_HIP(hipMemcpyAsync(b + NA_DCM, t->dcol_m, 8 * W, hipMemcpyHostToDevice, st));
    MX_HIP(hipMemcpyAsync(b + NA_DCM + 8 * W, t->drow_m, 8 * H, hipMemcpyHostToDevice, st));
    MX_HIP(hipMemcpyAsync(b + NA_DCE, t->dcol_e, 4 * W, hipMemcpyHostToDevice, st));
    MX_HIP(hipMemcpyAsync(b + NA_DCE + 4 * W, t->drow_e, 4 * H, hipMemcpyHostToDevice, st));
    const int ntab = (int)(W + H);
    hipLaunchKernelGGL(mx_pack_fe_kernel, dim3((ntab + 255) / 256), dim3(256), 0, st, (const double*)(b + NA_DCM), (const int*)(b + NA_DCE),
                       (Fe*)(b + NA_DCF), ntab);
    MX_HIP(hipGetLastError());
    MxJob J;
    J.R = ref->dev;
    J.R.minlevel = t->bla_minlevel;
    J.lev = ref->lev;
    MxTileDev& T = J.T;
    tile_fields(T, *t, *ref, W, H, false, false);
    T.dcol = (const Fe*)(b + NA_DCF);
    T.drow = (const Fe*)(b + NA_DCF) + W;
    T.val = (double*)(b + NA_VAL);
    T.nesc = (int*)(b + NA_NESC);
    T.out = (float*)(b + NA_OUT);
    T.queue = (unsigned*)(b + NA_QUEUE);
    T.stats = (unsigned long long*)(b + NA_STATS);
    T.cancel = ctx->cancel_dev;
    MX_HIP(hipMemcpyAsync(b + NA_JOB, &J, sizeof J, hipMemcpyHostToDevice, st));
    // the previous call's counters and outputs are cleared by a kernel
    hipLaunchKernelGGL(mx_reset_kernel, dim3(1024), dim3(256), 0, st, T.queue, T.stats, (unsigned*)nullptr, T.val, T.nesc, T.out, n);
    MX_HIP(hipGetLastError());
    const int blocks = std::max(1, std::min(t->grid_blocks > 0 ? t->grid_blocks : ctx->cu_count * 8, T.ntiles));
    const bool interior = (T.flags & MX_F_INTERIOR) != 0;
    MX_HIP(launch_jobs(t->form, (const MxJob*)(b + NA_JOB), 1, t->ede != 0, interior, blocks, 256, st));
    if (n_esc) MX_HIP(hipMemcpyAsync(n_esc, T.nesc, 4 * n, hipMemcpyDeviceToHost, st));
    if (value) MX_HIP(hipMemcpyAsync(value, T.val, 8 * n, hipMemcpyDeviceToHost, st));
    if (out) MX_HIP(hipMemcpyAsync(out, T.out, 4 * n, hipMemcpyDeviceToHost, st));
    if (stats4) MX_HIP(hipMemcpyAsync(stats4, T.stats, 32, hipMemcpyDeviceToHost, st));
    MX_HIP(hipStreamSynchronize(st));
    if (__atomic_load_n(ctx->cancel_host, __ATOMIC_ACQUIRE)) {
        ctx->set_err("cancelled");
        return MX_E_CANCELLED;
    }
    return MX_OK;
}

namespace {

// cell of sample (i, j, u, v) in the (H sub) x (W sub) output (getIndex,
// mandel_perturbation.cpp:4334-4340)
struct CellMap {
    int W, S;
    size_t operator()(int i, int j, int u, int v) const {
        return ((size_t)i * S + S / 2 + u) * (size_t)W * S + ((size_t)j * S + S / 2 + v);
    }
};

// Pixels (row-major) that adaptive supersampling refines, from the base values in
// the output buffer (mandel_perturbation.cpp:1796-1883): under EDE a sign change
// against a neighbour (the interior/exterior edge); otherwise log of the largest
// difference to a neighbour above log(max_iter) / 8 (MANDEL_SS_K = 8), with the
// float logarithm of the float difference.
std::vector<int32_t> ss_flags(const mx::Frame& F, const float* it, const CellMap& cm) {
    std::vector<int32_t> out;
    const double thresh = std::log((double)F.max_iter) / 8.0;
    for (int i = 0; i < F.H; ++i)
        for (int j = 0; j < F.W; ++j) {
            const float cv = it[cm(i, j, 0, 0)];
            bool need = false;
            if (F.ede) {
                for (int xi = -1; xi <= 1 && !need; ++xi)
                    for (int yi = -1; yi <= 1; ++yi) {
                        if (xi == 0 && yi == 0) continue;
                        const int ny = i + yi, nx = j + xi;
                        if (nx >= 0 && nx < F.W && ny >= 0 && ny < F.H && cv * it[cm(ny, nx, 0, 0)] < 0) {
                            need = true;
                            break;
                        }
                    }
            } else {
                float diff = 0.0f;
                for (int xi = -1; xi <= 1; ++xi)
                    for (int yi = -1; yi <= 1; ++yi) {
                        if (xi == 0 && yi == 0) continue;
                        const int ny = i + yi, nx = j + xi;
                        if (nx >= 0 && nx < F.W && ny >= 0 && ny < F.H) diff = std::max(diff, std::fabs(cv - it[cm(ny, nx, 0, 0)]));
                    }
                need = std::log(diff) > thresh;
            }
            if (need) out.push_back(i * F.W + j);
        }
    return out;
}

int upload_frame_ref(mx_ctx* ctx, const mx::Frame& F, mx_ref** ref) {
    mx_frame_view v;
    F.fill_view(v);
    mx_ref_desc d;
    std::memset(&d, 0, sizeof d);
    d.len = v.len;
    d.period = v.period;
    d.zf = v.zf;
    d.zfe_m = v.zfe_m;
    d.zfe_e = v.zfe_e;
    d.nlev = v.nlev;
    d.lev_off = v.lev_off;
    d.lev_size = v.lev_size;
    d.bla_r2 = v.bla_r2;
    d.bla_coef = v.bla_coef;
    d.rmax2 = v.rmax2;
    d.z_m3 = v.z_m3;
    d.sa_it = v.sa_it;
    d.sa_order = v.sa_order;
    d.sa_coef = v.sa_coef;
    d.sa_delta_re = v.sa_delta_re;
    d.sa_delta_im = v.sa_delta_im;
    std::vector<double> df;
    if (v.dfr) {
        df.resize(2 * (size_t)v.len);
        for (int32_t k = 0; k < v.len; ++k) {
            df[2 * (size_t)k] = v.dfr[k];
            df[2 * (size_t)k + 1] = v.dfi[k];
        }
        d.df = df.data();
        d.bla_der = v.bla_der;
        d.sa_coef_b = v.sa_coef_b;
    }
    return mx_ref_upload(ctx, &d, ref);
}

// Creates and launches one pass of the double kernel against `ref` over the grid
// (list == nullptr) or a list of samples.  side: on the context's side stream, stopped
// by its abort flag instead of the cancellation flag.
int dbl_pass_begin(mx_ctx* ctx, const std::vector<double>& dcol, const std::vector<double>& drow, const mx_ref* ref,
                   mx_tile t, const std::vector<int32_t>* list, int stride, bool side, mx_plan** out) {
    t.stride = stride;
    t.dcol_d = dcol.data();
    t.drow_d = drow.data();
    t.list = list ? list->data() : nullptr;
    t.list_len = list ? (int32_t)list->size() : 0;
    mx_plan* plan = nullptr;
    int rc = mx_plan_create(ctx, ref, &t, &plan);
    if (rc != MX_OK) return rc;
    if (side) {
        plan->stream = ctx->side;
        plan->tile.cancel = ctx->abort_dev;
    }
    rc = mx_plan_launch(plan);
    if (rc != MX_OK) {
        mx_plan_destroy(plan);
        return rc;
    }
    *out = plan;
    return MX_OK;
}

// Stops a side-stream pass whose results are not wanted and releases it.
void dbl_pass_discard(mx_ctx* ctx, mx_plan* plan) {
    if (!plan) return;
    __atomic_store_n(ctx->abort_host, 1, __ATOMIC_RELEASE);
    (void)hipStreamSynchronize(plan->stream);
    mx_plan_destroy(plan);
    __atomic_store_n(ctx->abort_host, 0, __ATOMIC_RELEASE);
}

// Waits for a launched pass; results merged into nesc/val for the samples it resolved,
// the unresolved ones returned with their glitch metric.  The plan stays alive.
int dbl_pass_collect(mx_plan* plan, std::vector<int32_t>& nesc, std::vector<double>& val, std::vector<int32_t>& un_idx,
                     std::vector<double>& un_gp, uint64_t st[4], double* kernel_ms) {
    double ms = 0.0;
    const size_t n = plan->n;
    std::vector<int32_t> pn(n);
    std::vector<double> pv(n);
    uint64_t ps[4] = {0, 0, 0, 0};
    int rc = mx_plan_wait(plan, &ms);
    if (rc == MX_OK) rc = mx_plan_read(plan, pn.data(), pv.data(), nullptr, ps);
    int32_t cnt = 0;
    if (rc == MX_OK) rc = mx_plan_unresolved(plan, nullptr, nullptr, 0, &cnt);
    if (rc == MX_OK) {
        un_idx.resize((size_t)cnt);
        un_gp.resize((size_t)cnt);
        rc = mx_plan_unresolved(plan, un_idx.data(), un_gp.data(), cnt, &cnt);
    }
    if (rc != MX_OK) return rc;
    for (size_t k = 0; k < n; ++k)
        if (pn[k] != INT_MIN) {
            nesc[k] = pn[k];
            val[k] = pv[k];
        }
    for (int q = 0; q < 4; ++q) st[q] += ps[q];
    if (kernel_ms) *kernel_ms += ms;
    // row-major order, as the engine's std::set of pending pixels
    std::vector<size_t> ord(un_idx.size());
    for (size_t q = 0; q < ord.size(); ++q) ord[q] = q;
    std::sort(ord.begin(), ord.end(), [&](size_t a, size_t b) { return un_idx[a] < un_idx[b]; });
    std::vector<int32_t> si(ord.size());
    std::vector<double> sg(ord.size());
    for (size_t q = 0; q < ord.size(); ++q) {
        si[q] = un_idx[ord[q]];
        sg[q] = un_gp[ord[q]];
    }
    un_idx.swap(si);
    un_gp.swap(sg);
    return MX_OK;
}

// One pass, launched and collected on the context's stream.
int dbl_pass(mx_ctx* ctx, const std::vector<double>& dcol, const std::vector<double>& drow, const mx_ref* ref, mx_tile t,
             const std::vector<int32_t>* list, int stride, std::vector<int32_t>& nesc, std::vector<double>& val,
             std::vector<int32_t>& un_idx, std::vector<double>& un_gp, uint64_t st[4], double* kernel_ms) {
    mx_plan* plan = nullptr;
    int rc = dbl_pass_begin(ctx, dcol, drow, ref, t, list, stride, false, &plan);
    if (rc != MX_OK) return rc;
    rc = dbl_pass_collect(plan, nesc, val, un_idx, un_gp, st, kernel_ms);
    mx_plan_destroy(plan);
    return rc;
}

// Mandel::Compute's plain-double path (mandel_perturbation.cpp:1668-1761): probe
// gate on the 8x8-strided grid, then passes until every pixel is resolved; after a
// pass that leaves glitched pixels the next reference is the pending pixel with the
// smallest glitch metric (the first one in row-major order among equals; the first
// pending pixel when none has a finite metric below 1e100), and the next pass runs
// over the pending pixels only.
int compute_dbl(mx_ctx* ctx, mx::Frame& F, const mx_request* req, mx_frame_info* info, double t0) {
    const double t_up = now_s();
    mx_ref* ref = nullptr;
    int rc = upload_frame_ref(ctx, F, &ref);
    if (rc != MX_OK) return rc;
    double upload_s = now_s() - t_up;
    const int ref_it0 = F.ref_it, sa_it0 = F.sa_it;
    const bool ref_escaped0 = F.ref_escaped;
    const int bla_levels0 = (int)F.lev_off.size();
    mx_tile t;
    std::memset(&t, 0, sizeof t);
    t.width = F.W;
    t.height = F.H;
    t.max_iter = F.max_iter;
    t.esc2 = F.esc2;
    t.form = MX_FORM_DOUBLE;
    t.ede = F.ede;
    t.dx_m = F.dxfe.m;
    t.dx_e = F.dxfe.e;
    t.interior = F.interior;
    t.int_rep = F.interior;
    t.bla = F.use_bla && !F.lev_off.empty();
    t.bla_minlevel = F.bla_minlevel;
    t.interior_eps2 = 3e-9 * 3e-9;
    t.interior_confirm = 4;
    t.skip_sample = -1;
    const size_t n = (size_t)F.W * F.H;
    std::vector<int32_t> nesc(n, INT_MIN);
    std::vector<double> val(n, -10.0);
    std::vector<int32_t> un_idx;
    std::vector<double> un_gp;
    uint64_t st[4] = {0, 0, 0, 0};
    double probe_ms = 0.0, kernel_ms = 0.0;
    bool interior_used = F.interior;
    // the reference pixel's value comes from its own orbit
    size_t refpix = (size_t)F.ref_i * F.W + F.ref_j;
    // the first pass as it runs when the probe finds no interior sample (interior
    // detection off), started on the side stream beside the probe: the probe and the
    // first pass are each bound by their longest chain of dependent steps, so side by
    // side the frame waits for the longer one only.  Identical launch, identical
    // results; stopped and discarded when the probe does find an interior sample.
    mx_plan* spec = nullptr;
    if (F.interior) {
        mx_tile ts = t;
        ts.interior = 0;
        ts.int_rep = 0;
        ts.skip_sample = (int32_t)refpix;
        rc = dbl_pass_begin(ctx, F.dcol_d, F.drow_d, ref, ts, nullptr, 1, true, &spec);
        if (rc != MX_OK) {
            mx_ref_release(ref);
            return rc;
        }
        std::vector<int32_t> pn(n, INT_MIN);
        std::vector<double> pv(n, -10.0);
        uint64_t pst[4] = {0, 0, 0, 0};
        mx_plan* probe = nullptr;
        rc = dbl_pass_begin(ctx, F.dcol_d, F.drow_d, ref, t, nullptr, 8, false, &probe);
        if (rc == MX_OK) rc = dbl_pass_collect(probe, pn, pv, un_idx, un_gp, pst, &probe_ms);
        int nint = 0;
        for (int i = 0; rc == MX_OK && i < F.H; i += 8)
            for (int j = 0; j < F.W; j += 8) {
                const int32_t v = pn[(size_t)i * F.W + j];
                if (v == -1 || v == -2) ++nint;
            }
        if (rc != MX_OK || nint > 0) {
            // stop the side pass before anything waits on the whole device
            dbl_pass_discard(ctx, spec);
            spec = nullptr;
        }
        mx_plan_destroy(probe);
        if (rc != MX_OK) {
            mx_ref_release(ref);
            return rc;
        }
        if (nint == 0) {
            interior_used = false;
            t.interior = 0;
            t.int_rep = 0;
        }
    }
    nesc[refpix] = F.ref_pixel_n;
    val[refpix] = F.ref_pixel_value;
    t.skip_sample = (int32_t)refpix;
    int passes = 0;
    std::vector<int32_t> list;
    bool first = true;
    while (true) {
        if (spec) {
            rc = dbl_pass_collect(spec, nesc, val, un_idx, un_gp, st, &kernel_ms);
            mx_plan_destroy(spec);
            spec = nullptr;
        } else
            rc = dbl_pass(ctx, F.dcol_d, F.drow_d, ref, t, first ? nullptr : &list, 1, nesc, val, un_idx, un_gp, st, &kernel_ms);
        first = false;
        ++passes;
        if (rc != MX_OK) break;
        if (un_idx.empty()) break;
        if (__atomic_load_n(ctx->cancel_host, __ATOMIC_ACQUIRE)) {
            ctx->set_err("cancelled");
            rc = MX_E_CANCELLED;
            break;
        }
        size_t best = 0;
        double mn = 1e100;
        bool found = false;
        for (size_t q = 0; q < un_idx.size(); ++q)
            if (un_gp[q] >= 0 && un_gp[q] < mn) {
                mn = un_gp[q];
                best = q;
                found = true;
            }
        if (!found) best = 0;
        const int32_t p = un_idx[best];
        mx_ref_release(ref);
        ref = nullptr;
        F.set_pixel_reference(p / F.W, p % F.W);
        if (F.halted()) {
            ctx->set_err("cancelled");
            rc = MX_E_CANCELLED;
            break;
        }
        refpix = (size_t)p;
        nesc[refpix] = F.ref_pixel_n;
        val[refpix] = F.ref_pixel_value;
        list.clear();
        for (size_t q = 0; q < un_idx.size(); ++q)
            if (q != best) list.push_back(un_idx[q]);
        if (list.empty()) break;
        const double tu = now_s();
        rc = upload_frame_ref(ctx, F, &ref);
        upload_s += now_s() - tu;
        if (rc != MX_OK) break;
        t.bla = F.use_bla && !F.lev_off.empty();
        t.skip_sample = (int32_t)refpix;
    }
    if (rc != MX_OK) {
        if (ref) mx_ref_release(ref);
        return rc;
    }
    const CellMap cm{F.W, F.sub};
    uint64_t nominal = 0;
    for (int i = 0; i < F.H; ++i)
        for (int j = 0; j < F.W; ++j) {
            const size_t k = (size_t)i * F.W + j, c = cm(i, j, 0, 0);
            const int32_t ne = nesc[k];
            req->iterations[c] = ne == INT_MIN ? -10.0f : ne < 0 ? -2.0f : (float)val[k];
            if (req->values) req->values[c] = val[k];
            if (req->n_esc) req->n_esc[c] = ne;
            nominal += ne > 0 ? (uint64_t)ne : (uint64_t)F.max_iter;
        }
    int ss_pixels = 0;
    double ss_ms = 0.0;
    if (F.ss) {
        // adaptive supersampling (mandel_perturbation.cpp:1789-2005): the flagged pixels'
        // other cells against the current reference, the first pass with its BLA table;
        // a glitched sub-sample becomes the next reference (BLA off from then on) until
        // every sub-sample resolves.  The cells form an (H sub) x (W sub) grid whose
        // dc tables are relative to the reference's cell.
        const std::vector<int32_t> flagged = ss_flags(F, req->iterations, cm);
        ss_pixels = (int)flagged.size();
        const int S = F.sub, h = S / 2, Wc = F.W * S, Hc = F.H * S;
        std::vector<int32_t> cells;
        for (int32_t pix : flagged) {
            const int i = pix / F.W, j = pix % F.W;
            for (int yi = -h; yi <= h; ++yi)
                for (int xi = -h; xi <= h; ++xi)
                    if (xi || yi) cells.push_back((int32_t)cm(i, j, yi, xi));
        }
        // the engine's std::set order of sub-samples (i, j, yi, xi), for ties
        auto set_key = [&](int32_t c) {
            const int r = c / Wc, q = c % Wc;
            return (((int64_t)(r / S) * F.W + q / S) * S + r % S) * S + q % S;
        };
        if (!cells.empty() && !ref) {
            const double tu = now_s();
            rc = upload_frame_ref(ctx, F, &ref);
            upload_s += now_s() - tu;
        }
        mx_tile ts = t;
        ts.width = Wc;
        ts.height = Hc;
        ts.skip_sample = -1;
        ts.bla = F.use_bla && !F.lev_off.empty();
        std::vector<double> ccol, crow;
        F.cell_dc(ccol, crow);
        std::vector<int32_t> cn;
        std::vector<double> cv;
        if (!cells.empty()) {
            cn.assign((size_t)Wc * Hc, INT_MIN);
            cv.assign((size_t)Wc * Hc, -10.0);
        }
        std::vector<int32_t> pending;
        pending.insert(pending.end(), cells.begin(), cells.end());
        while (rc == MX_OK && !pending.empty()) {
            rc = dbl_pass(ctx, ccol, crow, ref, ts, &pending, 1, cn, cv, un_idx, un_gp, st, &ss_ms);
            if (rc != MX_OK || un_idx.empty()) break;
            if (__atomic_load_n(ctx->cancel_host, __ATOMIC_ACQUIRE)) {
                ctx->set_err("cancelled");
                rc = MX_E_CANCELLED;
                break;
            }
            size_t best = 0;
            double mn = 1e100;
            bool found = false;
            for (size_t q = 0; q < un_idx.size(); ++q) {
                const bool less = un_gp[q] >= 0 && (un_gp[q] < mn || (found && un_gp[q] == mn && set_key(un_idx[q]) < set_key(un_idx[best])));
                if (less) {
                    mn = un_gp[q];
                    best = q;
                    found = true;
                }
            }
            if (!found)
                for (size_t q = 1; q < un_idx.size(); ++q)
                    if (set_key(un_idx[q]) < set_key(un_idx[best])) best = q;
            const int32_t bc = un_idx[best];
            const int br = bc / Wc, bq = bc % Wc;
            mx_ref_release(ref);
            ref = nullptr;
            F.set_pixel_reference(br / S, bq / S, br % S - h, bq % S - h, false);
            if (F.halted()) {
                ctx->set_err("cancelled");
                rc = MX_E_CANCELLED;
                break;
            }
            cn[(size_t)bc] = F.ref_pixel_n;
            cv[(size_t)bc] = F.ref_pixel_value;
            pending.clear();
            for (size_t q = 0; q < un_idx.size(); ++q)
                if (q != best) pending.push_back(un_idx[q]);
            if (pending.empty()) break;
            const double tu = now_s();
            rc = upload_frame_ref(ctx, F, &ref);
            upload_s += now_s() - tu;
            F.cell_dc(ccol, crow);
            ts.bla = 0;
        }
        if (rc == MX_OK)
            for (int32_t c : cells) {
                const int32_t ne = cn[(size_t)c];
                if (ne == INT_MIN) continue;
                req->iterations[c] = ne < 0 ? -2.0f : (float)cv[(size_t)c];
                if (req->values) req->values[c] = cv[(size_t)c];
                if (req->n_esc) req->n_esc[c] = ne;
                nominal += ne > 0 ? (uint64_t)ne : (uint64_t)F.max_iter;
            }
    }
    if (ref) mx_ref_release(ref);
    if (rc != MX_OK) return rc;
    if (info) {
        std::memset(info, 0, sizeof *info);
        info->path = 1;
        info->ref_it = ref_it0;
        info->ref_escaped = ref_escaped0;
        info->ref_virtual = 0;
        info->ref_x = F.W / 2;
        info->ref_y = F.H / 2;
        info->interior_used = interior_used;
        info->precision_bits = (int32_t)F.prec;
        info->bla_levels = bla_levels0;
        info->ref_seconds = F.ref_seconds;
        info->bla_seconds = F.bla_seconds;
        info->upload_seconds = upload_s;
        info->probe_ms = probe_ms;
        info->kernel_ms = kernel_ms;
        info->total_seconds = now_s() - t0;
        info->steps = st[0];
        info->bla_applies = st[1];
        info->bla_skipped = st[2];
        info->rebases = st[3];
        info->nominal_iterations = nominal;
        info->passes = passes;
        info->sa_it = sa_it0;
        info->ss_pixels = ss_pixels;
        info->ss_ms = ss_ms;
    }
    return MX_OK;
}

// The cutoff-sensitive frame's brute-force pass (finalize, mandel_perturbation.cpp:
// 2173-2183): every computed cell whose float value is negative or above max_iter - 16
// (the classification the double-precision deltas cannot be trusted with when the
// reference itself escapes at the cutoff) is recomputed in arbitrary precision at its
// own c (accuratePixelCompute, 896-918), on the host's cores as the engine does.
int cutoff_fallback(mx::Frame& F, const mx_request* req, const CellMap& cm, const std::vector<size_t>& ss_cells, int& count) {
    struct Job {
        size_t c;
        int i, j, u, v;
    };
    std::vector<Job> jobs;
    auto consider = [&](size_t c, int i, int j, int u, int v) {
        const float x = req->iterations[c];
        if (x < 0 || x > F.max_iter - 16) jobs.push_back({c, i, j, u, v});
    };
    for (int i = 0; i < F.H; ++i)
        for (int j = 0; j < F.W; ++j) consider(cm(i, j, 0, 0), i, j, 0, 0);
    const int S = F.sub, h = S / 2;
    const size_t rowlen = (size_t)F.W * S;
    for (size_t c : ss_cells) {
        const size_t row = c / rowlen, col = c % rowlen;
        consider(c, (int)(row / S), (int)(col / S), (int)(row % S) - h, (int)(col % S) - h);
    }
    count = (int)jobs.size();
    if (jobs.empty()) return MX_OK;
    std::atomic<size_t> next{0};
    std::atomic<bool> halted{false};
    auto work = [&]() {
        for (size_t q; (q = next.fetch_add(1)) < jobs.size();) {
            const Job& jb = jobs[q];
            int n = 0;
            const double r = F.pixel_value(jb.i, jb.j, jb.u, jb.v, n);
            if (n == INT_MIN) {
                halted.store(true);
                return;
            }
            req->iterations[jb.c] = n < 0 ? -2.0f : (float)r;
            if (req->values) req->values[jb.c] = n < 0 ? -2.0 : r;
            if (req->n_esc) req->n_esc[jb.c] = n;
        }
    };
    const size_t nt = std::min<size_t>(jobs.size(), std::max(1u, std::thread::hardware_concurrency()));
    std::vector<std::thread> pool;
    for (size_t k = 1; k < nt; ++k) pool.emplace_back(work);
    work();
    for (auto& t : pool) t.join();
    return halted.load() ? MX_E_CANCELLED : MX_OK;
}

// One launch of the shallow kernel over `nitems` items (the base grid, or the cells of
// the flagged pixels in `list`); outputs per item.
int run_shallow(mx_ctx* ctx, MxShallowDev S, bool ede, const std::vector<int32_t>* list, size_t nout, std::vector<int32_t>& nesc,
                std::vector<double>& val, uint64_t& iters, double& ms) {
    const size_t o_val = 0, o_nesc = align_up(8 * nout), o_out = align_up(o_nesc + 4 * nout), o_q = align_up(o_out + 4 * nout);
    const size_t o_st = align_up(o_q + 4 * MX_NQUEUE), o_list = align_up(o_st + 8 * 4);
    const size_t bytes = align_up(o_list + (list ? 4 * list->size() : 0));
    char* b = nullptr;
    if (hipMalloc(&b, bytes) != hipSuccess) {
        ctx->set_err("hipMalloc of the shallow frame failed");
        return MX_E_NOMEM;
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    S.val = (double*)(b + o_val);
    S.nesc = (int*)(b + o_nesc);
    S.out = (float*)(b + o_out);
    S.queue = (unsigned*)(b + o_q);
    S.stats = (unsigned long long*)(b + o_st);
    S.list = list ? (const int*)(b + o_list) : nullptr;
    S.nlist = list ? (int)list->size() : 0;
    S.cancel = ctx->cancel_dev;
    auto run = [&]() -> hipError_t {
        hipError_t e = hipEventCreate(&e0);
        if (e == hipSuccess) e = hipEventCreate(&e1);
        if (e == hipSuccess && list)
            e = hipMemcpyAsync(b + o_list, list->data(), 4 * list->size(), hipMemcpyHostToDevice, ctx->stream);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(mx_reset_kernel, dim3(1024), dim3(256), 0, ctx->stream, S.queue, S.stats, (unsigned*)nullptr, S.val, S.nesc,
                           S.out, nout);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        if ((e = hipEventRecord(e0, ctx->stream)) != hipSuccess) return e;
        const int work = (int)std::min<size_t>((S.nitems + 63) / 64, (size_t)INT_MAX);
        const int blocks = std::max(1, std::min(ctx->cu_count * 8, work));
        if ((e = mx_launch_shallow(S, ede, blocks, ctx->stream)) != hipSuccess) return e;
        if ((e = hipEventRecord(e1, ctx->stream)) != hipSuccess) return e;
        if ((e = hipMemcpyAsync(nesc.data(), S.nesc, 4 * nout, hipMemcpyDeviceToHost, ctx->stream)) != hipSuccess) return e;
        if ((e = hipMemcpyAsync(val.data(), S.val, 8 * nout, hipMemcpyDeviceToHost, ctx->stream)) != hipSuccess) return e;
        unsigned long long st = 0;
        if ((e = hipMemcpyAsync(&st, S.stats, 8, hipMemcpyDeviceToHost, ctx->stream)) != hipSuccess) return e;
        if ((e = hipStreamSynchronize(ctx->stream)) != hipSuccess) return e;
        iters = st;
        float f = 0.f;
        e = hipEventElapsedTime(&f, e0, e1);
        ms = f;
        return e;
    };
    const hipError_t e = run();
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(b);
    if (e != hipSuccess) {
        ctx->set_err(std::string("shallow path: ") + hipGetErrorString(e));
        return MX_E_HIP;
    }
    if (__atomic_load_n(ctx->cancel_host, __ATOMIC_ACQUIRE)) {
        ctx->set_err("cancelled");
        return MX_E_CANCELLED;
    }
    return MX_OK;
}

// Shallow direct path (scale <= 1e6; Mandel::Compute's method 0,
// mandel_perturbation.cpp:1444-1519): every pixel iterates its own c in double on the
// device, then, with adaptive supersampling, the other cells of each flagged pixel
// (1933-1965).  No reference orbit, series, BLA or interior probe.
int compute_shallow(mx_ctx* ctx, mx::Frame& F, const mx_request* req, mx_frame_info* info, double t0) {
    MxShallowDev S;
    std::memset(&S, 0, sizeof S);
    S.c0r = F.sh_c0r;
    S.c0i = F.sh_c0i;
    S.dx = F.sh_dx;
    S.dy = F.sh_dy;
    S.W = F.W;
    S.H = F.H;
    S.sub = F.sub;
    S.mxit = F.max_iter;
    S.esc2 = F.esc2;
    S.tiles_x = (F.W + 7) / 8;
    S.nitems = (unsigned)S.tiles_x * (unsigned)((F.H + 7) / 8) * 64u;
    S.chunk = 64;
    S.quorum = 32;
    const size_t n = (size_t)F.W * F.H;
    std::vector<int32_t> nesc(n);
    std::vector<double> val(n);
    uint64_t iters = 0;
    double kernel_ms = 0.0;
    int rc = run_shallow(ctx, S, F.ede, nullptr, n, nesc, val, iters, kernel_ms);
    if (rc != MX_OK) return rc;
    const CellMap cm{F.W, F.sub};
    uint64_t nominal = 0;
    for (int i = 0; i < F.H; ++i)
        for (int j = 0; j < F.W; ++j) {
            const size_t k = (size_t)i * F.W + j, c = cm(i, j, 0, 0);
            req->iterations[c] = (float)val[k];
            if (req->values) req->values[c] = val[k];
            if (req->n_esc) req->n_esc[c] = nesc[k];
            nominal += nesc[k] > 0 ? (uint64_t)nesc[k] : (uint64_t)F.max_iter;
        }
    int ss_pixels = 0;
    double ss_ms = 0.0;
    if (F.ss) {
        const std::vector<int32_t> flagged = ss_flags(F, req->iterations, cm);
        ss_pixels = (int)flagged.size();
        const int SS = F.sub, h = SS / 2, per = SS * SS - 1;
        const size_t m = flagged.size() * (size_t)per;
        if (m > 0) {
            if (m > (size_t)UINT_MAX) return MX_E_INVALID;
            MxShallowDev T = S;
            T.nitems = (unsigned)m;
            std::vector<int32_t> sn(m);
            std::vector<double> sv(m);
            uint64_t it2 = 0;
            rc = run_shallow(ctx, T, F.ede, &flagged, m, sn, sv, it2, ss_ms);
            if (rc != MX_OK) return rc;
            iters += it2;
            for (size_t q = 0; q < m; ++q) {
                const int pix = flagged[q / per];
                int t = (int)(q % per);
                if (t >= h * SS + h) ++t;
                const size_t c = cm(pix / F.W, pix % F.W, t % SS - h, t / SS - h);
                req->iterations[c] = (float)sv[q];
                if (req->values) req->values[c] = sv[q];
                if (req->n_esc) req->n_esc[c] = sn[q];
                nominal += sn[q] > 0 ? (uint64_t)sn[q] : (uint64_t)F.max_iter;
            }
        }
    }
    if (info) {
        std::memset(info, 0, sizeof *info);
        info->path = 0;
        info->ss_pixels = ss_pixels;
        info->ss_ms = ss_ms;
        info->precision_bits = (int32_t)F.prec;
        info->kernel_ms = kernel_ms;
        info->total_seconds = now_s() - t0;
        info->steps = iters;
        info->nominal_iterations = nominal;
        info->passes = 1;
        info->ref_virtual = 1;
    }
    return MX_OK;
}

} // namespace

extern "C" int mx_compute(mx_ctx* ctx, const mx_request* req, mx_frame_info* info) {
    if (!ctx || !req || !req->iterations) return MX_E_INVALID;
    BusyGuard busy(ctx);
    if (!busy.ok) return MX_E_BUSY;
    if (__atomic_load_n(ctx->cancel_host, __ATOMIC_ACQUIRE)) return MX_E_CANCELLED;
    const double t0 = now_s();
    mx_frame frame;
    mx::Frame& F = frame.impl;
    F.halt = ctx->cancel_host;
    F.check_cutoff = true;
    std::string err;
    int rc = F.prepare(*req, err);
    if (rc != MX_OK) {
        ctx->set_err(err);
        return rc;
    }
    if (F.shallow) return compute_shallow(ctx, F, req, info, t0);
    if (F.dbl) return compute_dbl(ctx, F, req, info, t0);
    if (req->form != MX_FORM_FE64 && req->form != MX_FORM_FE32 && req->form != MX_FORM_SC64) {
        ctx->set_err("form must be MX_FORM_FE64, MX_FORM_FE32 or MX_FORM_SC64 on the floatexp path");
        return MX_E_INVALID;
    }
    mx_frame_view v;
    F.fill_view(v);
    const double t_up = now_s();
    mx_ref_desc d;
    std::memset(&d, 0, sizeof d);
    d.len = v.len;
    d.period = v.period;
    d.zf = v.zf;
    d.zfe_m = v.zfe_m;
    d.zfe_e = v.zfe_e;
    d.nlev = v.nlev;
    d.lev_off = v.lev_off;
    d.lev_size = v.lev_size;
    d.bla_r2 = v.bla_r2;
    d.bla_coef = v.bla_coef;
    d.rmax2 = v.rmax2;
    mx_ref* ref = nullptr;
    rc = mx_ref_upload(ctx, &d, &ref);
    if (rc != MX_OK) return rc;
    const double upload_s = now_s() - t_up;
    mx_tile t;
    std::memset(&t, 0, sizeof t);
    t.width = F.W;
    t.height = F.H;
    t.stride = 1;
    t.max_iter = F.max_iter;
    t.esc2 = F.esc2;
    t.form = req->form;
    t.ede = F.ede;
    t.interior = F.interior;
    t.int_rep = F.interior;
    t.bla = F.use_bla && v.nlev > 0;
    t.bla_minlevel = F.bla_minlevel;
    t.interior_eps2 = 3e-9 * 3e-9;
    t.interior_confirm = 4;
    t.dcol_m = v.dcol_m;
    t.dcol_e = v.dcol_e;
    t.drow_m = v.drow_m;
    t.drow_e = v.drow_e;
    t.dx_m = v.dx_m;
    t.dx_e = v.dx_e;
    t.skip_sample = -1;
    const size_t n = (size_t)F.W * F.H;
    std::vector<int32_t> nesc(n);
    std::vector<double> val(n);
    uint64_t st[4] = {0, 0, 0, 0};
    double probe_ms = 0.0, kernel_ms = 0.0;
    bool interior_used = F.interior;
    // interior auto-gate probe on the 8x8-strided grid (mandel_perturbation.cpp:1668-1700)
    if (F.interior) {
        mx_tile pt = t;
        pt.stride = 8;
        mx_plan* plan = nullptr;
        rc = mx_plan_create(ctx, ref, &pt, &plan);
        if (rc == MX_OK) rc = mx_plan_launch(plan);
        if (rc == MX_OK) rc = mx_plan_wait(plan, &probe_ms);
        if (rc == MX_OK) rc = mx_plan_read(plan, nesc.data(), nullptr, nullptr, nullptr);
        mx_plan_destroy(plan);
        if (rc != MX_OK) {
            mx_ref_release(ref);
            return rc;
        }
        int nint = 0;
        for (int i = 0; i < F.H; i += 8)
            for (int j = 0; j < F.W; j += 8)
                if (nesc[(size_t)i * F.W + j] < 0 && nesc[(size_t)i * F.W + j] != INT_MIN) ++nint;
        if (nint == 0) {
            interior_used = false;
            t.interior = 0;
            t.int_rep = 0;
        }
    }
    // main pass: every sample but the reference pixel, whose value comes from the
    // reference orbit itself (mandel_perturbation.cpp:1724-1726)
    if (!F.ref_virtual) t.skip_sample = (F.H / 2) * F.W + (F.W / 2);
    {
        mx_plan* plan = nullptr;
        rc = mx_plan_create(ctx, ref, &t, &plan);
        if (rc == MX_OK) rc = mx_plan_launch(plan);
        if (rc == MX_OK) rc = mx_plan_wait(plan, &kernel_ms);
        if (rc == MX_OK) rc = mx_plan_read(plan, nesc.data(), val.data(), nullptr, st);
        mx_plan_destroy(plan);
    }
    mx_ref_release(ref);
    if (rc != MX_OK) return rc;
    if (!F.ref_virtual) {
        F.ref_pixel_wait();
        const size_t c = (size_t)(F.H / 2) * F.W + (F.W / 2);
        nesc[c] = F.ref_pixel_n;
        val[c] = F.ref_pixel_value;
    }
    // results into the (H sub) x (W sub) buffer, pixel (i, j) at its centre cell
    const CellMap cm{F.W, F.sub};
    uint64_t nominal = 0;
    for (int i = 0; i < F.H; ++i)
        for (int j = 0; j < F.W; ++j) {
            const size_t k = (size_t)i * F.W + j, c = cm(i, j, 0, 0);
            const int32_t ne = nesc[k];
            req->iterations[c] = ne < 0 ? -2.0f : (float)val[k];
            if (req->values) req->values[c] = val[k];
            if (req->n_esc) req->n_esc[c] = ne;
            nominal += ne > 0 ? (uint64_t)ne : (uint64_t)F.max_iter;
        }
    int ss_pixels = 0;
    double ss_ms = 0.0;
    std::vector<size_t> ss_cells;
    if (F.ss) {
        // adaptive supersampling (mandel_perturbation.cpp:1789-1981): flag the pixels
        // that differ from a neighbour, then every other cell of each flagged pixel in
        // one pass against the same reference and BLA table
        const std::vector<int32_t> flagged = ss_flags(F, req->iterations, cm);
        ss_pixels = (int)flagged.size();
        const int S = F.sub, h = S / 2;
        const size_t m = flagged.size() * (size_t)(S * S - 1);
        if (m > 0) {
            std::vector<double> dm(2 * m);
            std::vector<int32_t> de(2 * m);
            std::vector<size_t> cells(m);
            size_t q = 0;
            for (int32_t pix : flagged) {
                const int i = pix / F.W, j = pix % F.W;
                for (int yi = -h; yi <= h; ++yi)
                    for (int xi = -h; xi <= h; ++xi) {
                        if (xi == 0 && yi == 0) continue;
                        // stepParallel's dc of a sub-sample (mandel_perturbation.cpp:2160-2163)
                        const double px = j + (double)xi / S, py = i + (double)yi / S;
                        const Fe a = fe_mul_d(F.dxfe, px - F.ref_x), b = fe_mul_d(F.dyfe, py - F.ref_y);
                        dm[2 * q] = a.m;
                        de[2 * q] = a.e;
                        dm[2 * q + 1] = b.m;
                        de[2 * q + 1] = b.e;
                        cells[q] = cm(i, j, yi, xi);
                        ++q;
                    }
            }
            mx_tile ts = t;
            ts.skip_sample = -1;
            ts.n_samples = (int32_t)m;
            ts.sample_dc_m = dm.data();
            ts.sample_dc_e = de.data();
            std::vector<int32_t> sn(m);
            std::vector<double> svl(m);
            uint64_t st2[4] = {0, 0, 0, 0};
            rc = mx_ref_upload(ctx, &d, &ref);
            if (rc != MX_OK) return rc;
            mx_plan* plan = nullptr;
            rc = mx_plan_create(ctx, ref, &ts, &plan);
            if (rc == MX_OK) rc = mx_plan_launch(plan);
            if (rc == MX_OK) rc = mx_plan_wait(plan, &ss_ms);
            if (rc == MX_OK) rc = mx_plan_read(plan, sn.data(), svl.data(), nullptr, st2);
            mx_plan_destroy(plan);
            mx_ref_release(ref);
            if (rc != MX_OK) return rc;
            for (size_t k = 0; k < m; ++k) {
                const size_t c = cells[k];
                req->iterations[c] = sn[k] < 0 ? -2.0f : (float)svl[k];
                if (req->values) req->values[c] = svl[k];
                if (req->n_esc) req->n_esc[c] = sn[k];
                nominal += sn[k] > 0 ? (uint64_t)sn[k] : (uint64_t)F.max_iter;
            }
            for (int k = 0; k < 4; ++k) st[k] += st2[k];
            ss_cells = std::move(cells);
        }
    }
    int gmp_fallback = 0;
    if (F.cutoff_sensitive) {
        rc = cutoff_fallback(F, req, cm, ss_cells, gmp_fallback);
        if (rc != MX_OK) {
            ctx->set_err("cancelled");
            return rc;
        }
    }
    if (info) {
        std::memset(info, 0, sizeof *info);
        info->gmp_fallback = gmp_fallback;
        info->ss_pixels = ss_pixels;
        info->ss_ms = ss_ms;
        info->path = 2;
        info->ref_it = F.ref_it;
        info->ref_escaped = F.ref_escaped;
        info->ref_virtual = F.ref_virtual;
        info->ref_x = F.ref_x;
        info->ref_y = F.ref_y;
        info->interior_used = interior_used;
        info->cutoff_sensitive = F.cutoff_sensitive;
        info->precision_bits = (int32_t)F.prec;
        info->bla_levels = v.nlev;
        info->ref_seconds = F.ref_seconds;
        info->bla_seconds = F.bla_seconds;
        info->upload_seconds = upload_s;
        info->probe_ms = probe_ms;
        info->kernel_ms = kernel_ms;
        info->total_seconds = now_s() - t0;
        info->steps = st[0];
        info->bla_applies = st[1];
        info->bla_skipped = st[2];
        info->rebases = st[3];
        info->nominal_iterations = nominal;
        info->passes = 1;
        info->escalated = F.escalated;
        info->period = F.period;
        info->subpixel = F.subpixel;
    }
    return MX_OK;
}
