// denoise.hip — the Vulkan renderer's post-filters on float RGB images:
// GaussDenoiser (src/graphic/denoiser/gauss_denoiser.cpp:31-47 weights,
// res/shader/denoise_gauss_comp.comp:17-33 filter) and MedianDenoiser
// (res/shader/denoise_median_comp.comp:22-46, a 3x3 median per channel).
//
// Both are HBM-bound stencils: a 16x16 tile of output pixels reads its
// (16 + 2 off)^2 input footprint into LDS once, then every output pixel sums
// (Gauss) or sorts (median) from LDS.  The arithmetic is the shaders':
// out = sum over x (outer) and y (inner) of in(p + (x, y)) * g[x + off] *
// g[y + off], taps outside the image skipped; the median network of the
// shader on the 9 taps, taps outside the image 0.
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>

#include "rt_internal.hpp"

namespace rt {
namespace {

constexpr int DN_TILE = 16;
constexpr int MAX_KERNEL = 19;  // odd sizes up to the shader's 20-entry weight array

struct Weights {
    float g[MAX_KERNEL + 1];
};

// A SPAN x SPAN pixel footprint whose top-left pixel is (x0, y0) into LDS
// (3 floats per pixel, row-major), one pixel per lane and step; pixels
// outside the image read 0.  (Flat float-row loads, and staging the output
// tile through LDS for row stores, both measured slower.)
template <int SPAN>
__device__ __forceinline__ void load_tile(const float* in, float* tile, int x0, int y0, uint32_t W, uint32_t H) {
    for (int i = threadIdx.x; i < SPAN * SPAN; i += blockDim.x) {
        const int tx = i % SPAN, ty = i / SPAN;
        const int gx = x0 + tx, gy = y0 + ty;
        const bool in_img = gx >= 0 && gx < (int)W && gy >= 0 && gy < (int)H;
        const size_t src = in_img ? 3 * ((size_t)gy * W + gx) : 0;
        for (int c = 0; c < 3; ++c) tile[3 * i + c] = in_img ? in[src + c] : 0.0f;
    }
}

// OFF = kernel_size / 2, a compile-time constant so the tap loops unroll;
// tiles away from the image border skip the per-tap bounds test (every tap
// is inside, the same terms in the same order).
template <int OFF>
__global__ void __launch_bounds__(DN_TILE* DN_TILE) k_gauss(const float* in, float* out, uint32_t W, uint32_t H,
                                                             Weights w) {
    constexpr int SPAN = DN_TILE + 2 * OFF;
    __shared__ float tile[SPAN * SPAN * 3];
    const int x0 = (int)blockIdx.x * DN_TILE - OFF, y0 = (int)blockIdx.y * DN_TILE - OFF;
    load_tile<SPAN>(in, tile, x0, y0, W, H);
    __syncthreads();
    const int lx = threadIdx.x % DN_TILE, ly = threadIdx.x / DN_TILE;
    const int px = (int)blockIdx.x * DN_TILE + lx, py = (int)blockIdx.y * DN_TILE + ly;
    const bool interior = x0 >= 0 && y0 >= 0 && x0 + SPAN <= (int)W && y0 + SPAN <= (int)H;  // block-uniform
    float r = 0.0f, g = 0.0f, b = 0.0f;
    if (px >= (int)W || py >= (int)H) {
        // outside the image: nothing to compute, the store skips it
    } else if (interior) {
#pragma unroll
        for (int x = -OFF; x <= OFF; ++x) {
#pragma unroll
            for (int y = -OFF; y <= OFF; ++y) {
                const float* t = tile + 3 * ((ly + OFF + y) * SPAN + (lx + OFF + x));
                const float wx = w.g[x + OFF], wy = w.g[y + OFF];
                r = r + t[0] * wx * wy;
                g = g + t[1] * wx * wy;
                b = b + t[2] * wx * wy;
            }
        }
    } else {
#pragma unroll
        for (int x = -OFF; x <= OFF; ++x) {
#pragma unroll
            for (int y = -OFF; y <= OFF; ++y) {
                const int qx = px + x, qy = py + y;
                if (qx < 0 || qx >= (int)W || qy < 0 || qy >= (int)H) continue;
                const float* t = tile + 3 * ((ly + OFF + y) * SPAN + (lx + OFF + x));
                const float wx = w.g[x + OFF], wy = w.g[y + OFF];
                r = r + t[0] * wx * wy;
                g = g + t[1] * wx * wy;
                b = b + t[2] * wx * wy;
            }
        }
    }
    if (px < (int)W && py < (int)H) {
        float* o = out + 3 * ((size_t)py * W + px);
        o[0] = r; o[1] = g; o[2] = b;
    }
}

// the shader's compare-exchange: a = min(a, b), b = max(old a, b)
__device__ __forceinline__ void s2(float& a, float& b) {
    const float t = a;
    a = fminf(a, b);
    b = fmaxf(t, b);
}

__global__ void __launch_bounds__(DN_TILE* DN_TILE) k_median(const float* in, float* out, uint32_t W, uint32_t H) {
    constexpr int span = DN_TILE + 2;
    __shared__ float tile[span * span * 3];
    load_tile<span>(in, tile, (int)blockIdx.x * DN_TILE - 1, (int)blockIdx.y * DN_TILE - 1, W, H);
    __syncthreads();
    const int lx = threadIdx.x % DN_TILE, ly = threadIdx.x / DN_TILE;
    const int px = (int)blockIdx.x * DN_TILE + lx, py = (int)blockIdx.y * DN_TILE + ly;
    if (px >= (int)W || py >= (int)H) return;
    float* o = out + 3 * ((size_t)py * W + px);
    for (int c = 0; c < 3; ++c) {
        // v[(dx + 1) * 3 + (dy + 1)]
        float v[9];
        for (int dx = -1; dx <= 1; ++dx)
            for (int dy = -1; dy <= 1; ++dy) v[(dx + 1) * 3 + (dy + 1)] = tile[3 * ((ly + 1 + dy) * span + (lx + 1 + dx)) + c];
        // mnmx6(0..5)
        s2(v[0], v[3]); s2(v[1], v[4]); s2(v[2], v[5]);
        s2(v[0], v[1]); s2(v[0], v[2]);
        s2(v[4], v[5]); s2(v[3], v[5]);
        // mnmx5(1, 2, 3, 4, 6)
        s2(v[1], v[2]); s2(v[3], v[4]);
        s2(v[1], v[3]); s2(v[1], v[6]);
        s2(v[4], v[6]); s2(v[2], v[6]);
        // mnmx4(2, 3, 4, 7)
        s2(v[2], v[3]); s2(v[4], v[7]); s2(v[2], v[4]); s2(v[3], v[7]);
        // mnmx3(3, 4, 8)
        s2(v[4], v[8]); s2(v[3], v[8]); s2(v[3], v[4]);
        o[c] = v[4];
    }
}

int stage_image(rt_ctx* c, const float* in, float* out, size_t bytes, int ptr_kind, const float** din, float** dout) {
    if (ptr_kind == RT_PTR_DEVICE) {
        *din = in;
        *dout = out;
        return RT_OK;
    }
    int rc;
    if ((rc = ensure(c, c->pt_image, 2 * bytes))) return rc;
    float* base = (float*)c->pt_image.ptr;
    RT_HIP(c, hipMemcpyAsync(base, in, bytes, hipMemcpyHostToDevice, c->stream));
    *din = base;
    *dout = base + bytes / 4;
    return RT_OK;
}

int check_image(rt_ctx* c, const float* in, float* out, uint32_t w, uint32_t h, int ptr_kind, const char* fn) {
    if (!in || !out) return fail(c, RT_E_INVALID, std::string(fn) + ": NULL image");
    if (w == 0 || h == 0) return fail(c, RT_E_INVALID, std::string(fn) + ": empty image");
    if ((uint64_t)w * h > (1ull << 28)) return fail(c, RT_E_UNSUPPORTED, std::string(fn) + ": image above 2^28 pixels");
    if (ptr_kind != RT_PTR_HOST && ptr_kind != RT_PTR_DEVICE) return fail(c, RT_E_INVALID, std::string(fn) + ": bad ptr_kind");
    return RT_OK;
}

}  // namespace
}  // namespace rt

using namespace rt;

extern "C" {

int rt_denoise_gauss_weights(uint32_t kernel_size, float sigma, float* weights) {
    if (!weights || kernel_size == 0 || kernel_size > MAX_KERNEL || kernel_size % 2 == 0) return RT_E_INVALID;
    // GaussDenoiser::updateUniforms: centred on index kernel_size - 1, normalised by 1 / sum
    float sum = 0.0f;
    for (uint32_t i = 0; i < kernel_size; ++i) {
        float a = ((float)i - ((float)kernel_size - 1.0f));
        a = a * a;
        const float b = (2.0f * sigma * sigma);
        const float r = (float)std::exp((double)(-a / b));
        weights[i] = r;
        sum += r;
    }
    sum = 1.0f / sum;
    for (uint32_t i = 0; i < kernel_size; ++i) weights[i] *= sum;
    return RT_OK;
}

int rt_denoise_gauss(rt_ctx* c, const float* in, float* out, uint32_t width, uint32_t height, uint32_t kernel_size,
                     float sigma, int ptr_kind) {
    if (!c) return RT_E_INVALID;
    int rc;
    if ((rc = check_image(c, in, out, width, height, ptr_kind, "rt_denoise_gauss"))) return rc;
    Weights w{};
    if (rt_denoise_gauss_weights(kernel_size, sigma, w.g) != RT_OK || !(sigma > 0.0f))
        return fail(c, RT_E_INVALID, "rt_denoise_gauss: kernel_size must be odd, 1..19, and sigma > 0");
    (void)hipSetDevice(c->device);
    const size_t bytes = 12ull * width * height;
    const float* din;
    float* dout;
    if ((rc = stage_image(c, in, out, bytes, ptr_kind, &din, &dout))) return rc;
    const dim3 grid((width + DN_TILE - 1) / DN_TILE, (height + DN_TILE - 1) / DN_TILE);
    RT_HIP(c, hipEventRecord(c->ev[2], c->stream));
    switch (kernel_size / 2) {
#define RT_GAUSS_CASE(o)                                                                                  \
    case o:                                                                                               \
        hipLaunchKernelGGL(k_gauss<o>, grid, dim3(DN_TILE * DN_TILE), 0, c->stream, din, dout, width, height, w); \
        break;
        RT_GAUSS_CASE(0) RT_GAUSS_CASE(1) RT_GAUSS_CASE(2) RT_GAUSS_CASE(3) RT_GAUSS_CASE(4)
        RT_GAUSS_CASE(5) RT_GAUSS_CASE(6) RT_GAUSS_CASE(7) RT_GAUSS_CASE(8) RT_GAUSS_CASE(9)
#undef RT_GAUSS_CASE
        default:
            return fail(c, RT_E_INVALID, "rt_denoise_gauss: kernel_size above 19");
    }
    RT_HIP(c, hipGetLastError());
    RT_HIP(c, hipEventRecord(c->ev[3], c->stream));
    c->last_op = 5;
    if (ptr_kind == RT_PTR_HOST) {
        RT_HIP(c, hipMemcpyAsync(out, dout, bytes, hipMemcpyDeviceToHost, c->stream));
        RT_HIP(c, hipStreamSynchronize(c->stream));
    }
    return RT_OK;
}

int rt_denoise_median(rt_ctx* c, const float* in, float* out, uint32_t width, uint32_t height, int ptr_kind) {
    if (!c) return RT_E_INVALID;
    int rc;
    if ((rc = check_image(c, in, out, width, height, ptr_kind, "rt_denoise_median"))) return rc;
    (void)hipSetDevice(c->device);
    const size_t bytes = 12ull * width * height;
    const float* din;
    float* dout;
    if ((rc = stage_image(c, in, out, bytes, ptr_kind, &din, &dout))) return rc;
    const dim3 grid((width + DN_TILE - 1) / DN_TILE, (height + DN_TILE - 1) / DN_TILE);
    RT_HIP(c, hipEventRecord(c->ev[2], c->stream));
    hipLaunchKernelGGL(k_median, grid, dim3(DN_TILE * DN_TILE), 0, c->stream, din, dout, width, height);
    RT_HIP(c, hipGetLastError());
    RT_HIP(c, hipEventRecord(c->ev[3], c->stream));
    c->last_op = 5;
    if (ptr_kind == RT_PTR_HOST) {
        RT_HIP(c, hipMemcpyAsync(out, dout, bytes, hipMemcpyDeviceToHost, c->stream));
        RT_HIP(c, hipStreamSynchronize(c->stream));
    }
    return RT_OK;
}

}  // extern "C"
