// denoise.hip — the Vulkan renderer's post-filters on float RGB images:
// GaussDenoiser (src/graphic/denoiser/gauss_denoiser.cpp:31-47 weights,
// res/shader/denoise_gauss_comp.comp:17-33 filter) and MedianDenoiser
// (res/shader/denoise_median_comp.comp:22-46, a 3x3 median per channel).
//
// Each workgroup computes a 64-pixel-wide tile of output pixels from its
// input footprint staged in LDS once; each thread computes R vertically
// adjacent pixels of one column, so a column of taps is read from LDS once
// for all R.  The arithmetic is the shaders': out = sum over x (outer) and y
// (inner) of (in(p + (x, y)) * g[x + off]) * g[y + off], taps outside the
// image skipped (the product in(p + (x, y)) * g[x + off] depends on the tap
// and x only, so it is formed once per tap and x and shared by the pixels
// that use it: the same roundings); the median network of the shader on the
// 9 taps, taps outside the image 0.
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>

#include "rt_internal.hpp"

namespace rt {
namespace {

constexpr int DN_W = 64;               // tile width: one column per thread
constexpr int DN_G = 4;                // row groups (threads per column)
constexpr int DN_NT = DN_W * DN_G;     // threads per workgroup
// Rows per thread, R (the tile is DN_W x DN_G R): sharing a column of taps
// over more rows saves LDS reads and products but costs registers, so it
// pays for wide kernels only (measured at 3840 x 2160: Gauss 5 x 5 best at
// R = 2, 19 x 19 at R = 4; the median at R = 4).
constexpr int gauss_rows(int off) { return off >= 4 ? 4 : 2; }
constexpr int MEDIAN_ROWS = 4;
constexpr int MAX_KERNEL = 19;  // odd sizes up to the shader's 20-entry weight array

struct Weights {
    float g[MAX_KERNEL + 1];
};

// The SW x SH pixel footprint whose top-left pixel is (x0, y0) into LDS, 3
// floats per pixel, row-major.  Thread f < 3 SW owns float f of every
// footprint row (one image column and channel), so its column test is made
// once and each row costs one load at a running offset, all of them in
// flight together; pixels outside the image read 0.  Image offsets fit in
// 32 bits (at most 3 * 2^28 floats).
template <int SW, int SH>
__device__ __forceinline__ void load_tile(const float* in, float* tile, int x0, int y0, uint32_t W, uint32_t H) {
    constexpr int RF = 3 * SW;  // floats per footprint row
    static_assert(RF <= DN_NT, "one pass of the workgroup loads a footprint row");
    const int f = (int)threadIdx.x;
    if (f >= RF) return;
    const int gx = x0 + f / 3;
    const bool col_in = gx >= 0 && gx < (int)W;
    const int row = 3 * (int)W;
    int off = y0 * row + 3 * x0 + f;
#pragma unroll
    for (int ty = 0; ty < SH; ++ty, off += row) {
        const int gy = y0 + ty;
        tile[ty * RF + f] = (col_in && gy >= 0 && gy < (int)H) ? in[off] : 0.0f;
    }
}

// A thread's DN_R pixels of one tile: OFF = kernel_size / 2, a compile-time
// constant so the tap loops unroll; EDGE: the footprint reaches past the
// image, so taps outside it are skipped (interior tiles test nothing: every
// tap is inside, the same terms in the same order).
template <int OFF, bool EDGE>
__device__ __forceinline__ void gauss_pixels(const float* tile, float* out, const Weights& w, int lx, int lr, int px,
                                             int py, uint32_t W, uint32_t H) {
    constexpr int DN_R = gauss_rows(OFF);
    constexpr int SW = DN_W + 2 * OFF, NR = DN_R + 2 * OFF;
    float acc[DN_R][3];
#pragma unroll
    for (int p = 0; p < DN_R; ++p) acc[p][0] = acc[p][1] = acc[p][2] = 0.0f;
#pragma unroll
    for (int x = -OFF; x <= OFF; ++x) {
        const float wx = w.g[x + OFF];
        const bool col_in = !EDGE || (px + x >= 0 && px + x < (int)W);
        // tap rows lr .. lr + NR - 1 of column lx + OFF + x, times wx
        float tw[NR][3];
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            const float* t = tile + 3 * ((lr + r) * SW + (lx + OFF + x));
            tw[r][0] = t[0] * wx;
            tw[r][1] = t[1] * wx;
            tw[r][2] = t[2] * wx;
        }
#pragma unroll
        for (int y = -OFF; y <= OFF; ++y) {
            const float wy = w.g[y + OFF];
#pragma unroll
            for (int p = 0; p < DN_R; ++p) {
                const int r = p + y + OFF;
                if (!EDGE || (col_in && py + p + y >= 0 && py + p + y < (int)H)) {
                    acc[p][0] = acc[p][0] + tw[r][0] * wy;
                    acc[p][1] = acc[p][1] + tw[r][1] * wy;
                    acc[p][2] = acc[p][2] + tw[r][2] * wy;
                }
            }
        }
    }
#pragma unroll
    for (int p = 0; p < DN_R; ++p) {
        if (!EDGE || py + p < (int)H) {
            float* o = out + 3 * ((size_t)(py + p) * W + px);
            o[0] = acc[p][0]; o[1] = acc[p][1]; o[2] = acc[p][2];
        }
    }
}

template <int OFF>
__global__ void __launch_bounds__(DN_NT) k_gauss(const float* in, float* out, uint32_t W, uint32_t H, Weights w) {
    constexpr int DN_R = gauss_rows(OFF), DN_H = DN_G * DN_R;
    constexpr int SW = DN_W + 2 * OFF, SH = DN_H + 2 * OFF;
    __shared__ float tile[SW * SH * 3];
    const int x0 = (int)blockIdx.x * DN_W - OFF, y0 = (int)blockIdx.y * DN_H - OFF;
    load_tile<SW, SH>(in, tile, x0, y0, W, H);
    __syncthreads();
    const int lx = threadIdx.x % DN_W, lr = (threadIdx.x / DN_W) * DN_R;  // column, first row in the tile
    const int px = (int)blockIdx.x * DN_W + lx, py = (int)blockIdx.y * DN_H + lr;
    if (px >= (int)W || py >= (int)H) return;
    if (x0 >= 0 && y0 >= 0 && x0 + SW <= (int)W && y0 + SH <= (int)H)  // block-uniform
        gauss_pixels<OFF, false>(tile, out, w, lx, lr, px, py, W, H);
    else
        gauss_pixels<OFF, true>(tile, out, w, lx, lr, px, py, W, H);
}

// the shader's compare-exchange: a = min(a, b), b = max(old a, b)
__device__ __forceinline__ void s2(float& a, float& b) {
    const float t = a;
    a = fminf(a, b);
    b = fmaxf(t, b);
}

__global__ void __launch_bounds__(DN_NT) k_median(const float* in, float* out, uint32_t W, uint32_t H) {
    constexpr int DN_R = MEDIAN_ROWS, DN_H = DN_G * DN_R;
    constexpr int SW = DN_W + 2, SH = DN_H + 2, NR = DN_R + 2;
    __shared__ float tile[SW * SH * 3];
    load_tile<SW, SH>(in, tile, (int)blockIdx.x * DN_W - 1, (int)blockIdx.y * DN_H - 1, W, H);
    __syncthreads();
    const int lx = threadIdx.x % DN_W, lr = (threadIdx.x / DN_W) * DN_R;
    const int px = (int)blockIdx.x * DN_W + lx, py = (int)blockIdx.y * DN_H + lr;
    if (px >= (int)W || py >= (int)H) return;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        // the three tap columns of this thread's pixels, rows lr .. lr + NR - 1
        float col[3][NR];
#pragma unroll
        for (int dx = 0; dx < 3; ++dx)
#pragma unroll
            for (int r = 0; r < NR; ++r) col[dx][r] = tile[3 * ((lr + r) * SW + (lx + dx)) + c];
#pragma unroll
        for (int p = 0; p < DN_R; ++p) {
            if (py + p >= (int)H) break;
            // v[(dx + 1) * 3 + (dy + 1)]
            float v[9];
#pragma unroll
            for (int dx = 0; dx < 3; ++dx)
#pragma unroll
                for (int dy = 0; dy < 3; ++dy) v[dx * 3 + dy] = col[dx][p + dy];
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
            out[3 * ((size_t)(py + p) * W + px) + c] = v[4];
        }
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
    RT_HIP(c, hipEventRecord(c->ev[2], c->stream));
    switch (kernel_size / 2) {
#define RT_GAUSS_CASE(o)                                                                                       \
    case o:                                                                                                    \
        hipLaunchKernelGGL(k_gauss<o>,                                                                         \
                           dim3((width + DN_W - 1) / DN_W, (height + DN_G * gauss_rows(o) - 1) / (DN_G * gauss_rows(o))), \
                           dim3(DN_NT), 0, c->stream, din, dout, width, height, w);                            \
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
    const dim3 grid((width + DN_W - 1) / DN_W, (height + DN_G * MEDIAN_ROWS - 1) / (DN_G * MEDIAN_ROWS));
    RT_HIP(c, hipEventRecord(c->ev[2], c->stream));
    hipLaunchKernelGGL(k_median, grid, dim3(DN_NT), 0, c->stream, din, dout, width, height);
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
