// Test infrastructure: prints the layout the plugin relies on (offsets of the
// ComputeRequest fields it reads, the sizes of the info structs, and the vtable slot of
// every IComputeBackend virtual), for whichever compute_backend.h hip_compute_backend.h
// selects.  Built once against the engine's header and once against the restatement,
// the two outputs must be equal.
#include <cstddef>
#include <cstdio>
#include <cstring>

#include "../mandelbrot-explorer_amd/csrc/hip_compute_backend.h"

// Itanium C++ ABI: a pointer to a virtual member function holds 1 + its vtable offset
template <typename T>
static long slot(T pm) {
    long v = 0;
    std::memcpy(&v, &pm, sizeof v);
    return v;
}

int main() {
#ifdef MANDEL_COMPUTE_BACKEND_H
    std::printf("header engine\n"); // the engine's include guard
#else
    std::printf("header restated\n");
#endif
#define OFF(f) std::printf("ComputeRequest.%s %zu\n", #f, offsetof(ComputeRequest, f))
    OFF(mode);
    OFF(cpuEngine);
    OFF(centerRe);
    OFF(centerIm);
    OFF(scale);
    OFF(fixedCRe);
    OFF(fixedCIm);
    OFF(width);
    OFF(height);
    OFF(fullHeight);
    OFF(rowBase);
    OFF(sub);
    OFF(maxIterations);
    OFF(coloringMethod);
    OFF(iterations);
    OFF(normal);
    OFF(progress);
    std::printf("sizeof ComputeBackendInfo %zu\n", sizeof(ComputeBackendInfo));
    std::printf("sizeof GenericDeepInfo %zu\n", sizeof(GenericDeepInfo));
    std::printf("slot info %ld\n", slot(&IComputeBackend::info));
    std::printf("slot lastComputeUsedGpuPath %ld\n", slot(&IComputeBackend::lastComputeUsedGpuPath));
    std::printf("slot lastComputeUsedCustomDeepPath %ld\n", slot(&IComputeBackend::lastComputeUsedCustomDeepPath));
    std::printf("slot lastComputeUsedGenericDeepPath %ld\n", slot(&IComputeBackend::lastComputeUsedGenericDeepPath));
    std::printf("slot lastGenericDeepInfo %ld\n", slot(&IComputeBackend::lastGenericDeepInfo));
    std::printf("slot compute %ld\n", slot(&IComputeBackend::compute));
    std::printf("slot cancel %ld\n", slot(&IComputeBackend::cancel));
    std::printf("slot resetCancellation %ld\n", slot(&IComputeBackend::resetCancellation));
    return 0;
}
