// Factory of the HIP compute backend, the counterpart of createD3D11ComputeBackend
// (src/engine/compute_backend_d3d11.h, called from compute_backend.cpp:281-286): the
// engine's createComputeBackend hands it the CPU backend that serves what the HIP path
// does not; with no usable device it returns nullptr and says why in *error (the
// engine then keeps its CPU backend, marked as a fallback).
//
// Inside the engine tree (src/engine on the include path) the plugin is built against
// the engine's own compute_backend.h; elsewhere against this library's restatement.
#pragma once

#if __has_include(<compute_backend.h>)
#include <compute_backend.h>
#else
#include "compute_backend.h"
#endif

#include <memory>
#include <string>

std::unique_ptr<IComputeBackend> createHipComputeBackend(std::unique_ptr<IComputeBackend> cpuFallback, std::string* error);
