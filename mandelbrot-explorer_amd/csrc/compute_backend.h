// The engine's compute-backend plugin interface (src/engine/compute_backend.h:17-115),
// restated for builds of this library outside the engine tree: same type and member
// names, the same members in the same order, the same virtual functions in the same
// order (so the same vtable), the same meaning of every field and of compute()'s bool
// result.  One difference: ComputeRequest stops after `progress`; the engine's
// Expression-mode fields follow it there (compute_backend.h:43-50), so every field
// read here sits at the engine's offset.  Inside the engine tree the plugin compiles
// against the engine's own header instead (hip_compute_backend.cpp).
#pragma once

#include <gmp.h>

#include <array>
#include <atomic>
#include <cstdint>
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

// Report of the engine's generic certified deep path (Expression mode); the HIP
// backend never takes that path and reports the default, or its CPU fallback's.
struct GenericDeepInfo {
    bool used = false;
    bool settled = false;
    bool success = false;
    bool cancelled = false;
    bool taylorAccepted = false;
    bool preflightAttempted = false;
    bool preflightRejectedFast = false;
    bool specializedPiecewiseMpfr = false;
    bool piecewiseBigFixed = false;
    std::string status;
    std::string error;
    std::string phase;
    std::string predictedPath;
    float progress = 0.0f;
    uint64_t pixelCount = 0;
    uint64_t preflightSampleCount = 0;
    uint64_t preflightFallbackCount = 0;
    uint64_t preflightAvoidedFastPixelCount = 0;
    uint64_t preflightIterationCount = 0;
    uint64_t preflightOperationCount = 0;
    uint64_t preflightFoldOperationCount = 0;
    std::array<uint64_t, 16> preflightFirstUncertainHistogram{};
    uint64_t fastPixelCount = 0;
    uint64_t fallbackPixelCount = 0;
    uint64_t taylorPixelCoverage = 0;
    uint64_t totalIterationCount = 0;
    uint64_t specializedPiecewiseMpfrPixelCount = 0;
    uint64_t specializedPiecewiseMpfrIterationCount = 0;
    uint64_t specializedPiecewiseMpfrPeriodicPixelCount = 0;
    uint64_t genericMpfrPeriodicPixelCount = 0;
    uint64_t piecewiseBigFixedPixelCount = 0;
    uint64_t piecewiseBigFixedIterationCount = 0;
    uint64_t selectedPrecision = 0;
    uint64_t fallbackPrecision = 0;
    double totalSeconds = 0.0;
    double referenceSeconds = 0.0;
    double preflightSeconds = 0.0;
    double taylorSeconds = 0.0;
    double fastSeconds = 0.0;
    double fallbackSeconds = 0.0;
};

class IComputeBackend {
  public:
    virtual ~IComputeBackend() = default;
    virtual const ComputeBackendInfo& info() const = 0;
    virtual bool lastComputeUsedGpuPath() const = 0;
    virtual bool lastComputeUsedCustomDeepPath() const { return false; }
    virtual bool lastComputeUsedGenericDeepPath() const { return false; }
    virtual GenericDeepInfo lastGenericDeepInfo() const { return {}; }
    virtual bool compute(const ComputeRequest& request) = 0;
    virtual void cancel() = 0;
    virtual void resetCancellation() = 0;
};
