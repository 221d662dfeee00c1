// HipComputeBackend: the IComputeBackend plugin over the C ABI (include/mx.h), the
// slot the D3D11 backend occupies in the engine (compute_backend_d3d11.cpp:209-635,
// compute_backend.cpp:277-290).  Validation, busy and cancellation follow
// CpuComputeBackend (compute_backend.cpp:44-128, 263-272).
#include <strings.h>

#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/mx.h"
#include "compute_backend.h"

namespace {

std::string g_create_error;

// decimal string of an mpf value with all its digits ("0.DDDDe<exp>" or "0")
std::string mpf_decimal(mpf_srcptr x) {
    mp_exp_t e = 0;
    char* d = mpf_get_str(nullptr, &e, 10, 0, x);
    std::string s(d);
    void (*freefn)(void*, size_t);
    mp_get_memory_functions(nullptr, nullptr, &freefn);
    freefn(d, std::strlen(d) + 1);
    if (s.empty()) return "0";
    std::string out;
    if (s[0] == '-') {
        out = "-";
        s.erase(0, 1);
    }
    out += "0." + s + "e" + std::to_string((long)e);
    return out;
}

class HipComputeBackend final : public IComputeBackend {
  public:
    explicit HipComputeBackend(mx_ctx* ctx) : _ctx(ctx) {
        _info.name = "HIP";
        _info.detail = "MI355X (gfx950) perturbation kernels";
        _info.hardwareAccelerated = true;
        _info.fallback = false;
    }
    ~HipComputeBackend() override { mx_ctx_destroy(_ctx); }

    const ComputeBackendInfo& info() const override { return _info; }
    bool lastComputeUsedGpuPath() const override { return _lastGpu.load(std::memory_order_acquire); }

    bool compute(const ComputeRequest& request) override {
        _lastGpu.store(false, std::memory_order_release);
        if (!valid(request)) return false;
        if (_cancelRequested.load(std::memory_order_acquire)) return false;
        {
            std::lock_guard<std::mutex> g(_mutex);
            if (_active) return false; // busy
            _active = true;
        }
        const std::string cre = mpf_decimal(request.centerRe), cim = mpf_decimal(request.centerIm);
        // the engine reads the scale as a decimal magnitude; its precision follows the
        // scale's digit count (render.cpp:170), so pass it as an integer string when it
        // is one
        std::string sc = mpf_decimal(request.scale);
        {
            mpf_t t;
            mpf_init2(t, mpf_get_prec(request.scale));
            mpf_floor(t, request.scale);
            if (mpf_cmp(t, request.scale) == 0) {
                char* d = nullptr;
                gmp_asprintf(&d, "%.0Ff", t);
                sc = d;
                void (*freefn)(void*, size_t);
                mp_get_memory_functions(nullptr, nullptr, &freefn);
                freefn(d, std::strlen(d) + 1);
            }
            mpf_clear(t);
        }
        mx_request r;
        std::memset(&r, 0, sizeof r);
        r.center_re = cre.c_str();
        r.center_im = cim.c_str();
        r.scale = sc.c_str();
        r.width = request.width;
        r.height = request.height;
        r.full_height = request.fullHeight;
        r.row_base = request.rowBase;
        r.sub = request.sub;
        r.max_iterations = request.maxIterations;
        r.coloring_method = request.coloringMethod & ~MX_SUPER_SAMPLING;
        r.escape_radius = 0.0f;
        if (const char* e = std::getenv("MANDEL_BAILOUT")) { // Mandel::Mandel, mandel_perturbation.cpp:232-233
            const double b = std::atof(e);
            if (b > 0) r.escape_radius = (float)b;
        }
        // precision from the scale's digit count (render.cpp:170); MANDEL_PREC=<digits>
        // overrides it as in render.cpp:174
        r.precision_bits = 0;
        if (const char* e = std::getenv("MANDEL_PREC")) r.precision_bits = (int)(std::atol(e) * 3.3219281 + 64.0);
        r.form = MX_FORM_FE64;
        if (const char* f = std::getenv("MANDEL_HIP_FORM")) r.form = std::atoi(f);
        r.use_bla = -1;
        r.interior = -1;
        if (const char* e = std::getenv("MANDEL_BLA")) r.use_bla = std::atoi(e) != 0;
        if (const char* e = std::getenv("MANDEL_INTERIOR")) r.interior = std::atoi(e) != 0;
        if (const char* e = std::getenv("MANDEL_BLA_MINSKIP")) r.bla_minskip = std::atoi(e);
        // MANDEL_PERIODIC=0: no nucleus search (buildReferenceOrbit, mandel_perturbation.cpp:
        // 1240-1250); the search runs by default above 1e280 with BLA on
        if (const char* e = std::getenv("MANDEL_PERIODIC"))
            if (std::atoi(e) == 0) r.options |= MX_OPT_NO_PERIODIC;
        r.iterations = request.iterations;
        if (request.progress) request.progress->store(0.0f, std::memory_order_relaxed);
        mx_frame_info info;
        const int rc = mx_compute(_ctx, &r, &info);
        {
            std::lock_guard<std::mutex> g(_mutex);
            _active = false;
        }
        if (rc != MX_OK) return false;
        _lastGpu.store(true, std::memory_order_release);
        if (request.progress) request.progress->store(1.0f, std::memory_order_relaxed);
        return !_cancelRequested.load(std::memory_order_acquire);
    }

    void cancel() override {
        _cancelRequested.store(true, std::memory_order_release);
        mx_cancel(_ctx);
    }

    void resetCancellation() override {
        std::lock_guard<std::mutex> g(_mutex);
        if (!_active) {
            _cancelRequested.store(false, std::memory_order_release);
            mx_reset_cancellation(_ctx);
        }
    }

  private:
    // compute_backend.cpp:263-272 for the Mandelbrot mode
    static bool valid(const ComputeRequest& q) {
        const int fullHeight = q.fullHeight > 0 ? q.fullHeight : q.height;
        if (!q.centerRe || !q.centerIm || !q.scale || mpf_sgn(q.scale) <= 0 || q.width < 2 || q.height < 1 || fullHeight < 2 ||
            q.rowBase < 0 || q.rowBase > fullHeight - q.height || q.sub < 1 || q.maxIterations < 1 || !q.iterations)
            return false;
        return q.mode == ComputeMode::Mandelbrot;
    }

    mx_ctx* _ctx;
    ComputeBackendInfo _info;
    std::mutex _mutex;
    bool _active = false;
    std::atomic_bool _cancelRequested{false};
    std::atomic_bool _lastGpu{false};
};

} // namespace

std::unique_ptr<IComputeBackend> createComputeBackend(const char* requested) {
    g_create_error.clear();
    const bool hip = requested && (strcasecmp(requested, "hip") == 0 || strcasecmp(requested, "gpu") == 0 || strcasecmp(requested, "auto") == 0);
    if (!hip) {
        g_create_error = "unknown backend name";
        return nullptr;
    }
    mx_ctx* ctx = nullptr;
    const int rc = mx_ctx_create(0, &ctx);
    if (rc != MX_OK) {
        g_create_error = rc == MX_E_NODEVICE ? "no HIP device" : "HIP device initialisation failed";
        return nullptr;
    }
    return std::make_unique<HipComputeBackend>(ctx);
}

const char* computeBackendError() { return g_create_error.c_str(); }
