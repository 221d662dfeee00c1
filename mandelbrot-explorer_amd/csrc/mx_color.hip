// On-device colouring and box downsampling of a frame's float buffer: the final pass
// of tools/render.cpp (src/tools/render.cpp:252-302), so that a frame computed on the
// GPU leaves it as 3 bytes per pixel instead of 4 bytes per sample.
//
// Per output pixel: each of its ss x ss samples is mapped through the palette
// (getColor, render.cpp:69-85: interior and empty samples black; smooth counts through
// color_func's log-power curve, distance estimates through tanh), taken to linear light
// (toLin, c^2.2), averaged, and returned to gamma (toGam, l^(1/2.2)), rounded half up to
// a byte.  The palette is render.cpp's colorMapInit (21-30): six knots through
// mono_cubic_interpolate (engine/interpolate.cpp:7-65), 2048 entries per channel.
#include <hip/hip_runtime.h>

#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "mx_internal.h"

namespace {

constexpr int COLP = 2048;

// mono_cubic_interpolate (interpolate.cpp:7-65): monotone cubic Hermite through the
// knots, interior slopes the weighted harmonic mean of the neighbouring secants (the
// end slopes are the end secants), sampled at i / mp_sz; all in float, as there
void mono_cubic(const float* xs, const float* ys, int n, float* out, int m) {
    std::vector<float> dxs(n - 1), ms(n - 1), c1(n), c2(n - 1), c3(n - 1);
    for (int i = 0; i < n - 1; ++i) {
        dxs[i] = xs[i + 1] - xs[i];
        ms[i] = (ys[i + 1] - ys[i]) / dxs[i];
    }
    c1[0] = ms[0];
    for (int i = 0; i < n - 2; ++i) {
        const float a = ms[i], b = ms[i + 1];
        if (a * b <= 0) {
            c1[i + 1] = 0;
        } else {
            const float common = dxs[i] + dxs[i + 1];
            c1[i + 1] = 3 * common / ((common + dxs[i + 1]) / a + (common + dxs[i]) / b);
        }
    }
    c1[n - 1] = ms[n - 2];
    for (int i = 0; i < n - 1; ++i) {
        const float inv = 1 / dxs[i], common = c1[i] + c1[i + 1] - ms[i] - ms[i];
        c2[i] = (ms[i] - c1[i] - common) * inv;
        c3[i] = common * inv * inv;
    }
    for (int k = 0; k < m; ++k) {
        const float x = 1.f * k / m;
        float y;
        if (x == xs[n - 1]) {
            y = ys[n - 1];
        } else {
            // the interval holding x by bisection over the first n-1 knots; a knot
            // itself returns its value
            int lo = 0, hi = n - 2, hit = -1;
            while (lo <= hi) {
                const int mid = (lo + hi) / 2;
                if (xs[mid] < x)
                    lo = mid + 1;
                else if (xs[mid] > x)
                    hi = mid - 1;
                else {
                    hit = mid;
                    break;
                }
            }
            if (hit >= 0) {
                y = ys[hit];
            } else {
                const int i = hi > 0 ? hi : 0;
                const float d = x - xs[i], d2 = d * d;
                y = ys[i] + c1[i] * d + c2[i] * d2 + c3[i] * d * d2;
            }
        }
        out[k] = y;
    }
}

// render.cpp's palette: three channels of COLP floats (0..255)
void build_palette(float* pal) {
    static const float pos[6] = {0.0f, 0.16f, 0.42f, 0.6425f, 0.8575f, 1.0f};
    static const float col[3][6] = {{0, 32, 237, 255, 0, 0}, {70, 107, 255, 170, 2, 70}, {100, 203, 255, 0, 0, 100}};
    for (int c = 0; c < 3; ++c) mono_cubic(pos, col[c], 6, pal + c * COLP, COLP);
}

// (int)(f * COLP) % COLP with the conversion x86 performs for a float out of int range
// (the integer indefinite, INT_MIN, whose remainder is 0); NaN maps there too
__device__ __forceinline__ int palette_index(float f) {
    const float v = f * (float)COLP;
    int x = (v < 2147483648.0f && v > -2147483904.0f) ? (int)v : INT_MIN;
    x %= COLP;
    if (x < 0) x += COLP;
    return x;
}

template <bool EDE>
__global__ __launch_bounds__(256) void mx_color_kernel(const float* __restrict__ it, int W, int H, int ss, float density,
                                                       const float* __restrict__ pal, uint8_t* __restrict__ rgb) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
    if (j >= W || i >= H) return;
    const size_t Ws = (size_t)W * ss;
    float lr = 0.f, lg = 0.f, lb = 0.f;
    for (int a = 0; a < ss; ++a)
        for (int b = 0; b < ss; ++b) {
            const float v = it[((size_t)i * ss + a) * Ws + (size_t)j * ss + b];
            float r = 0.f, g = 0.f, bb = 0.f;
            if (v >= 0.f) {
                float f;
                if (EDE) {
                    f = tanhf(v * density / 3600.0f * 5.0f);
                } else {
                    const float l = logf(v + 2);
                    f = powf(l, l * l * density / 3600.0f);
                }
                const int x = palette_index(f);
                r = pal[x];
                g = pal[COLP + x];
                bb = pal[2 * COLP + x];
            }
            // toLin (render.cpp:87-89), summed in float
            lr += powf(r / 255.f, 2.2f);
            lg += powf(g / 255.f, 2.2f);
            lb += powf(bb / 255.f, 2.2f);
        }
    const int n = ss * ss;
    auto gam = [](float l) { return 255.f * powf(l < 0 ? 0 : l, 1.f / 2.2f); };
    uint8_t* p = rgb + ((size_t)i * W + j) * 3;
    p[0] = (uint8_t)(gam(lr / n) + 0.5f);
    p[1] = (uint8_t)(gam(lg / n) + 0.5f);
    p[2] = (uint8_t)(gam(lb / n) + 0.5f);
}

// Colour a frame's samples already on the device into rgb (device), W x H pixels of
// ss x ss samples each; the palette is uploaded for the call.
int colorize_device(mx_ctx* ctx, const float* dev_iter, int W, int H, int ss, bool ede, float density, uint8_t* dev_rgb) {
    std::vector<float> pal(3 * COLP);
    build_palette(pal.data());
    float* dpal = nullptr;
    hipError_t e = hipMalloc((void**)&dpal, sizeof(float) * pal.size());
    if (e == hipSuccess) e = hipMemcpyAsync(dpal, pal.data(), sizeof(float) * pal.size(), hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) {
        const dim3 grid((unsigned)((W + 255) / 256), (unsigned)H);
        if (ede)
            hipLaunchKernelGGL((mx_color_kernel<true>), grid, dim3(256), 0, ctx->stream, dev_iter, W, H, ss, density, dpal, dev_rgb);
        else
            hipLaunchKernelGGL((mx_color_kernel<false>), grid, dim3(256), 0, ctx->stream, dev_iter, W, H, ss, density, dpal, dev_rgb);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (dpal) (void)hipFree(dpal);
    if (e != hipSuccess) {
        ctx->set_err(std::string("colorize: ") + hipGetErrorString(e));
        return MX_E_HIP;
    }
    return MX_OK;
}

// rgb (host) from a device float frame; the device rgb buffer is allocated for the call
int colorize_to_host(mx_ctx* ctx, const float* dev_iter, int W, int H, int ss, bool ede, float density, uint8_t* rgb) {
    const size_t np = (size_t)W * H;
    uint8_t* drgb = nullptr;
    if (hipMalloc((void**)&drgb, np * 3) != hipSuccess) {
        ctx->set_err("colorize: hipMalloc failed");
        return MX_E_NOMEM;
    }
    int rc = colorize_device(ctx, dev_iter, W, H, ss, ede, density, drgb);
    if (rc == MX_OK) {
        hipError_t e = hipMemcpyAsync(rgb, drgb, np * 3, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) {
            ctx->set_err(std::string("colorize: ") + hipGetErrorString(e));
            rc = MX_E_HIP;
        }
    }
    (void)hipFree(drgb);
    return rc;
}

} // namespace

extern "C" int mx_palette(float* out) {
    if (!out) return MX_E_INVALID;
    build_palette(out);
    return MX_OK;
}

extern "C" int mx_colorize(mx_ctx* ctx, const float* iterations, int32_t width, int32_t height, int32_t ss, int32_t coloring_method,
                           float density, uint8_t* rgb) {
    if (!ctx || !iterations || !rgb || width < 1 || height < 1 || height > 65535 || ss < 1 || ss > 16 || !(density > 0.0f))
        return MX_E_INVALID;
    if ((int64_t)width * ss * height * ss > ((int64_t)1 << 32)) return MX_E_INVALID;
    if (hipSetDevice(ctx->device) != hipSuccess) return MX_E_HIP;
    const size_t ns = (size_t)width * ss * height * ss;
    float* dit = nullptr;
    if (hipMalloc((void**)&dit, ns * sizeof(float)) != hipSuccess) {
        ctx->set_err("colorize: hipMalloc failed");
        return MX_E_NOMEM;
    }
    int rc = MX_OK;
    hipError_t e = hipMemcpyAsync(dit, iterations, ns * sizeof(float), hipMemcpyHostToDevice, ctx->stream);
    if (e != hipSuccess) {
        ctx->set_err(std::string("colorize: ") + hipGetErrorString(e));
        rc = MX_E_HIP;
    }
    if (rc == MX_OK) rc = colorize_to_host(ctx, dit, width, height, ss, (coloring_method & MX_EXTERIOR_DIST_EST) != 0, density, rgb);
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(dit);
    return rc;
}

extern "C" int mx_plan_colorize(mx_plan* plan, int32_t ss, int32_t coloring_method, float density, uint8_t* rgb) {
    if (!plan || !rgb || ss < 1 || ss > 16 || !(density > 0.0f)) return MX_E_INVALID;
    const MxTileDev& T = plan->tile;
    mx_ctx* ctx = plan->ctx;
    if (T.sdc || T.list || T.stride != 1 || !T.out || T.W % ss != 0 || T.H % ss != 0 || T.H / ss > 65535) {
        ctx->set_err("colorize needs a full grid tile (stride 1, no sample list) whose sides are multiples of ss");
        return MX_E_INVALID;
    }
    if (hipSetDevice(ctx->device) != hipSuccess) return MX_E_HIP;
    return colorize_to_host(ctx, T.out, T.W / ss, T.H / ss, ss, (coloring_method & MX_EXTERIOR_DIST_EST) != 0, density, rgb);
}
