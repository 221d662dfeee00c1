// HipComputeBackend: the IComputeBackend plugin over the C ABI (include/mx.h), the
// slot the D3D11 backend occupies in the engine (compute_backend_d3d11.cpp:209-635,
// compute_backend.cpp:277-290).  Validation, busy and cancellation follow
// CpuComputeBackend (compute_backend.cpp:44-128, 263-272); requests outside this path
// (plain-double adaptive supersampling frames, Julia and Expression modes,
// other colorings) and device failures go to the CPU backend it was given, as the
// D3D11 backend does (compute_backend_d3d11.cpp:289-315).
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/mx.h"
#include "hip_compute_backend.h"

namespace {

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
    HipComputeBackend(mx_ctx* ctx, std::unique_ptr<IComputeBackend> cpu) : _ctx(ctx), _cpu(std::move(cpu)) {
        _info.name = "HIP";
        _info.detail = "MI355X (gfx950) perturbation kernels";
        if (_cpu) _info.detail += " + " + _cpu->info().name + " fallback";
        _info.hardwareAccelerated = true;
        _info.fallback = false;
    }
    ~HipComputeBackend() override { mx_ctx_destroy(_ctx); }

    const ComputeBackendInfo& info() const override { return _info; }
    bool lastComputeUsedGpuPath() const override { return _lastGpu.load(std::memory_order_acquire); }
    bool lastComputeUsedCustomDeepPath() const override { return ranCpu() && _cpu->lastComputeUsedCustomDeepPath(); }
    bool lastComputeUsedGenericDeepPath() const override { return ranCpu() && _cpu->lastComputeUsedGenericDeepPath(); }
    GenericDeepInfo lastGenericDeepInfo() const override { return ranCpu() ? _cpu->lastGenericDeepInfo() : GenericDeepInfo{}; }

    bool compute(const ComputeRequest& request) override {
        _lastGpu.store(false, std::memory_order_release);
        _lastCpu.store(false, std::memory_order_release);
        if (!valid(request)) return false;
        if (_cancelRequested.load(std::memory_order_acquire)) return false;
        {
            std::lock_guard<std::mutex> g(_mutex);
            if (_active) return false; // busy
            _active = true;
        }
        struct Release {
            HipComputeBackend* b;
            ~Release() {
                std::lock_guard<std::mutex> g(b->_mutex);
                b->_active = false;
            }
        } release{this};
        if (!eligible(request)) return computeCpu(request);
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
        r.coloring_method = request.coloringMethod;
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
        // MANDEL_PERIODIC=1 runs it at every scale; MANDEL_FE=1 takes the floatexp path
        // at every scale above 1e6 (mandel_perturbation.cpp:1573-1576)
        if (const char* e = std::getenv("MANDEL_PERIODIC")) r.options |= std::atoi(e) == 0 ? MX_OPT_NO_PERIODIC : std::atoi(e) == 1 ? MX_OPT_PERIODIC : 0;
        if (const char* e = std::getenv("MANDEL_FE")) {
            // MANDEL_FE=0 above 1e280 asks for the plain-double path past its range: the
            // engine's CPU backend runs that
            if (std::atoi(e) == 0) {
                if (mpf_cmp_d(request.scale, 1e280) > 0) return computeCpu(request);
            } else {
                r.options |= MX_OPT_FLOATEXP;
            }
        }
        r.iterations = request.iterations;
        if (request.progress) request.progress->store(0.0f, std::memory_order_relaxed);
        mx_frame_info info;
        const int rc = mx_compute(_ctx, &r, &info);
        if (rc == MX_OK) {
            _lastGpu.store(true, std::memory_order_release);
            const bool ok = !_cancelRequested.load(std::memory_order_acquire);
            if (ok && request.progress) request.progress->store(1.0f, std::memory_order_relaxed);
            return ok;
        }
        if (rc == MX_E_INVALID || rc == MX_E_CANCELLED || _cancelRequested.load(std::memory_order_acquire)) return false;
        // outside this path (MX_E_UNSUPPORTED) or a device failure: the CPU backend
        return computeCpu(request);
    }

    void cancel() override {
        _cancelRequested.store(true, std::memory_order_release);
        mx_cancel(_ctx);
        if (_cpu) _cpu->cancel();
    }

    void resetCancellation() override {
        {
            std::lock_guard<std::mutex> g(_mutex);
            if (!_active) {
                _cancelRequested.store(false, std::memory_order_release);
                mx_reset_cancellation(_ctx);
            }
        }
        if (_cpu) _cpu->resetCancellation();
    }

  private:
    // compute_backend.cpp:263-266: the fields every mode needs (the mode-specific checks
    // belong to the backend that serves the mode)
    static bool valid(const ComputeRequest& q) {
        const int fullHeight = q.fullHeight > 0 ? q.fullHeight : q.height;
        return q.centerRe && q.centerIm && q.scale && mpf_sgn(q.scale) > 0 && q.width >= 2 && q.height >= 1 && fullHeight >= 2 &&
               q.rowBase >= 0 && q.rowBase <= fullHeight - q.height && q.sub >= 1 && q.maxIterations >= 1 && q.iterations;
    }

    // the Mandelbrot mode with the smooth or distance-estimate output (optionally
    // adaptive supersampling); whether the depth is on this path mx_compute decides
    static bool eligible(const ComputeRequest& q) {
        return q.mode == ComputeMode::Mandelbrot && (q.coloringMethod & ~(MX_SUPER_SAMPLING | MX_EXTERIOR_DIST_EST)) == 0;
    }

    bool computeCpu(const ComputeRequest& request) {
        if (!_cpu) return false;
        _lastCpu.store(true, std::memory_order_release);
        return _cpu->compute(request);
    }

    bool ranCpu() const { return _cpu && _lastCpu.load(std::memory_order_acquire); }

    mx_ctx* _ctx;
    std::unique_ptr<IComputeBackend> _cpu;
    ComputeBackendInfo _info;
    std::mutex _mutex;
    bool _active = false;
    std::atomic_bool _cancelRequested{false};
    std::atomic_bool _lastGpu{false};
    std::atomic_bool _lastCpu{false};
};

} // namespace

std::unique_ptr<IComputeBackend> createHipComputeBackend(std::unique_ptr<IComputeBackend> cpuFallback, std::string* error) {
    mx_ctx* ctx = nullptr;
    const int rc = mx_ctx_create(0, &ctx);
    if (rc != MX_OK) {
        if (error) *error = rc == MX_E_NODEVICE ? "no HIP device" : "HIP device initialisation failed";
        return nullptr;
    }
    return std::make_unique<HipComputeBackend>(ctx, std::move(cpuFallback));
}
