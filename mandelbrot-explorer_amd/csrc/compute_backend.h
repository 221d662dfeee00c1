// The engine's compute-backend plugin interface for the Mandelbrot mode
// (src/engine/compute_backend.h:17-115), restated for this library: same type and
// member names, same meaning of every field and of compute()'s bool result.  The
// Julia and Expression modes and the Expression fields are not on this path and are
// absent; a request naming them is rejected.
#pragma once

#include <gmp.h>

#include <atomic>
#include <memory>
#include <string>

enum class ComputeMode { Mandelbrot, Julia, Expression };

struct ComputeRequest {
    ComputeMode mode = ComputeMode::Mandelbrot;
    void* cpuEngine = nullptr; // the engine's Mandel*; unused by the HIP backend

    mpf_ptr centerRe = nullptr;
    mpf_ptr centerIm = nullptr;
    mpf_ptr scale = nullptr;
    mpf_ptr fixedCRe = nullptr;
    mpf_ptr fixedCIm = nullptr;

    int width = 0;
    int height = 0;
    int fullHeight = 0;
    int rowBase = 0;
    int sub = 1;
    int maxIterations = 0;
    int coloringMethod = 0;
    float* iterations = nullptr; // width*height*sub*sub engine floats, caller-owned
    float* normal = nullptr;
    std::atomic<float>* progress = nullptr;
};

struct ComputeBackendInfo {
    std::string name;
    std::string detail;
    bool hardwareAccelerated = false;
    bool fallback = false;
};

class IComputeBackend {
  public:
    virtual ~IComputeBackend() = default;
    virtual const ComputeBackendInfo& info() const = 0;
    virtual bool lastComputeUsedGpuPath() const = 0;
    virtual bool compute(const ComputeRequest& request) = 0;
    virtual void cancel() = 0;
    virtual void resetCancellation() = 0;
};

// "hip", "gpu" or "auto" -> the HIP backend on device 0; nullptr when no HIP device is
// present or the name is unknown (this library has no CPU engine to fall back to).
std::unique_ptr<IComputeBackend> createComputeBackend(const char* requested);
// Reason the last createComputeBackend() returned nullptr.
const char* computeBackendError();
