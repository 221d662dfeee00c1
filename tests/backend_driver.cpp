// Exercises the C++ plugin interface (IComputeBackend over libmxhip.so) the way the
// engine's verify tool does in runBackendCase (src/tools/verify.cpp:9668-10069).
//   backend_driver names            unknown names are refused (no GPU needed)
//   backend_driver checks           invalid requests, cancellation (GPU)
//   backend_driver frame CRE CIM SCALE W H MXIT CMETHOD OUT.f32   one frame (GPU)
#include <gmp.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../mandelbrot-explorer_amd/csrc/compute_backend.h"

static int fail(const char* what) {
    std::fprintf(stderr, "FAIL: %s\n", what);
    return 1;
}

struct Mpf {
    mpf_t v;
    Mpf(const char* s, unsigned long prec) {
        mpf_init2(v, prec);
        mpf_set_str(v, s, 10);
    }
    ~Mpf() { mpf_clear(v); }
};

int main(int argc, char** argv) {
    if (argc < 2) return fail("usage");
    const std::string mode = argv[1];
    if (mode == "names") {
        if (createComputeBackend("d3d11-warp")) return fail("unknown name accepted");
        if (std::strcmp(computeBackendError(), "unknown backend name") != 0) return fail("error text");
        if (createComputeBackend(nullptr)) return fail("null name accepted");
        std::printf("names ok\n");
        return 0;
    }
    auto be = createComputeBackend("hip");
    if (!be) {
        std::fprintf(stderr, "no backend: %s\n", computeBackendError());
        return 2;
    }
    if (!be->info().hardwareAccelerated || be->info().fallback) return fail("info");
    if (mode == "checks") {
        Mpf cre("0", 4000), cim("1", 4000), sc("1e1000", 4000);
        std::vector<float> it(64 * 48, -10.f);
        ComputeRequest q;
        q.centerRe = cre.v;
        q.centerIm = cim.v;
        q.scale = sc.v;
        q.width = 64;
        q.height = 48;
        q.maxIterations = 10000;
        q.iterations = it.data();
        // invalid requests return false (verify.cpp:9803-9816)
        ComputeRequest bad = q;
        bad.width = 1;
        if (be->compute(bad)) return fail("width 1 accepted");
        bad = q;
        bad.rowBase = 100;
        if (be->compute(bad)) return fail("row base outside the strip accepted");
        bad = q;
        bad.iterations = nullptr;
        if (be->compute(bad)) return fail("null buffer accepted");
        bad = q;
        bad.mode = ComputeMode::Julia;
        if (be->compute(bad)) return fail("Julia accepted");
        // a cancelled backend refuses work until reset (verify.cpp:9817-9818)
        be->cancel();
        if (be->compute(q)) return fail("compute after cancel");
        be->resetCancellation();
        if (!be->compute(q)) return fail("compute after reset");
        if (!be->lastComputeUsedGpuPath()) return fail("gpu path flag");
        for (float v : it)
            if (!(v > 2600.f - 6.f && v < 2700.f)) return fail("c=i frame values");
        // strip of a taller frame: rows [8, 24) of a 64-row image
        std::vector<float> strip(64 * 16, -10.f);
        ComputeRequest s = q;
        s.height = 16;
        s.fullHeight = 64;
        s.rowBase = 8;
        s.iterations = strip.data();
        if (!be->compute(s)) return fail("strip");
        std::printf("checks ok\n");
        return 0;
    }
    if (mode == "frame" && argc == 10) {
        Mpf cre(argv[2], 8000), cim(argv[3], 8000), sc(argv[4], 8000);
        const int W = std::atoi(argv[5]), H = std::atoi(argv[6]);
        std::vector<float> it((size_t)W * H, -10.f);
        ComputeRequest q;
        q.centerRe = cre.v;
        q.centerIm = cim.v;
        q.scale = sc.v;
        q.width = W;
        q.height = H;
        q.maxIterations = std::atoi(argv[7]);
        q.coloringMethod = std::atoi(argv[8]);
        q.iterations = it.data();
        if (!be->compute(q)) return fail("compute");
        FILE* f = std::fopen(argv[9], "wb");
        if (!f) return fail("open output");
        std::fwrite(it.data(), sizeof(float), it.size(), f);
        std::fclose(f);
        std::printf("frame ok\n");
        return 0;
    }
    return fail("unknown mode");
}
