// ddgi.hip — irradiance probes (the Vulkan renderer's Majercik2019 pipeline,
// src/graphic/renderer/majercik2019_renderer.cpp:35-47) as a cache of the
// path tracer: probe rays, their light samples, the atlas update and the
// lookup at the cached render's query vertices.
//
// One update: k_ddgi_rays writes probes * rays rays (spherical Fibonacci
// directions, majercik2019_probe_raygen.rgen:8-28), the closest-hit kernel
// traces them, the path tracer's k_ddgi_shade turns each hit into radiance
// and queues its light samples, the any-hit kernel tests those, k_ddgi_resolve
// adds the unoccluded ones, and k_ddgi_update rebuilds every atlas texel
// (majercik2019_shading_update_comp.comp:25-94) from its probe's rays.
#include <hip/hip_runtime.h>

#include <cmath>
#include <new>
#include <string>

#include "ddgi_field.hpp"
#include "rt_internal.hpp"

using namespace rt;

struct rt_ddgi {
    rt_ctx* ctx = nullptr;
    rt_ddgi_config cfg{};
    DdgiGrid grid{};
    uint32_t P = 0, R = 0, W = 0, H = 0;
    uint64_t updates = 0;
    float rot[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    // atl: the field, one 32-byte record per texel (irradiance RGB + 0, depth moments + 0, 0); the API's
    // planar atlases are unpacked from it and packed into it
    DevBuf atl, surf, rays, hits, sh_ray, sh_w, sh_occ, io_in, io_out;
};

namespace {

constexpr int DG_BLOCK = 256;
constexpr uint32_t MAX_RAYS = 1024;  // rays per probe (the update keeps a probe's surfels in LDS)

struct Rot {
    float m[9];
};

int dg_fail(rt_ddgi* g, int code, const std::string& msg) { return fail(g->ctx, code, msg); }

__global__ void __launch_bounds__(DG_BLOCK) k_ddgi_rays(DdgiGrid g, Rot rot, uint32_t n, rt_ray* rays, float* surf) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const uint32_t p = j / g.rays, i = j % g.rays;
    const f3 d = rot_apply(rot.m, fib_dir(i, g.rays));
    const f3 o = probe_pos(g, p);
    rt_ray r;
    r.ox = o.x; r.oy = o.y; r.oz = o.z; r.tmin = 0.0f;
    r.dx = d.x; r.dy = d.y; r.dz = d.z; r.tmax = INFINITY;
    rays[j] = r;
    float* s = surf + 8 * (size_t)j;
    s[0] = d.x; s[1] = d.y; s[2] = d.z;
}

// radiance += (sum of the unoccluded light samples' weights) / count, in
// sample order
__global__ void __launch_bounds__(DG_BLOCK) k_ddgi_resolve(uint32_t n, uint32_t count, const float* sh_w,
                                                          const uint8_t* occ, float* surf) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    float ar = 0.0f, ag = 0.0f, ab = 0.0f;
    for (uint32_t k = 0; k < count; ++k) {
        const size_t q = (size_t)j * count + k;
        if (occ[q]) continue;
        ar = ar + sh_w[3 * q];
        ag = ag + sh_w[3 * q + 1];
        ab = ab + sh_w[3 * q + 2];
    }
    const float fc = (float)count;
    float* s = surf + 8 * (size_t)j;
    s[3] = s[3] + ar / fc;
    s[4] = s[4] + ag / fc;
    s[5] = s[5] + ab / fc;
}

// w^s for a cosine w in [0, 1] and a sharpness s >= 0, as GLSL defines pow:
// exp2(s * log2(w)), here on the hardware's base-2 logarithm and exponential
// (v_log_f32, v_exp_f32).  s = 0 gives 1 and w = 0 gives 0, as powf does.
// A subnormal w is scaled by 2^64 before the logarithm (the instruction
// takes normal inputs), so a small s still weights such a ray as powf would.
__device__ inline float depth_weight(float w, float s) {
    if (s == 0.0f) return 1.0f;
    if (w == 0.0f) return 0.0f;
    const bool sub = w < 1.17549435e-38f;
    const float lw = __builtin_amdgcn_logf(sub ? w * 18446744073709551616.0f : w) - (sub ? 64.0f : 0.0f);
    return __builtin_amdgcn_exp2f(s * lw);
}

// ppb probes per workgroup (their surfels side by side in LDS), one thread
// per atlas texel; each texel sums its probe's rays in ray order.  Border
// texels compute the interior texel they mirror, so they hold the same
// value.  first: no previous field to blend.
__global__ void __launch_bounds__(DG_BLOCK) k_ddgi_update(DdgiGrid g, const float* surf, float4* atl, uint32_t first,
                                                         uint32_t ppb, uint32_t P) {
    // a surfel is two float4: (dir.xyz, r), (g, b, distance, hit)
    __shared__ float4 ls[2 * MAX_RAYS];
    const uint32_t p0 = blockIdx.x * ppb;
    const uint32_t R = g.rays;
    const uint32_t np = min(ppb, P - p0);
    const float4* src = reinterpret_cast<const float4*>(surf) + (size_t)p0 * 2 * R;
    for (uint32_t k = threadIdx.x; k < 2 * R * np; k += blockDim.x) ls[k] = src[k];
    __syncthreads();
    const uint32_t ts = g.side + 2;
    for (uint32_t t = threadIdx.x; t < ts * ts * np; t += blockDim.x) {
        const uint32_t j = t / (ts * ts), tt = t % (ts * ts);
        const uint32_t p = p0 + j;
        const uint32_t tx = tt % ts, ty = tt / ts;
        uint32_t ix, iy;
        tile_source(g.side, tx, ty, ix, iy);
        const f3 nt = texel_dir(g.side, ix, iy);
        const float4* lp = ls + 2 * R * j;
        float sr = 0.0f, sgr = 0.0f, sb = 0.0f, tw = 0.0f, d0 = 0.0f, d1 = 0.0f, tdw = 0.0f;
        for (uint32_t i = 0; i < R; ++i) {
            const float4 a = lp[2 * i], b = lp[2 * i + 1];
            float rd = g.max_dist;
            if (b.w != 0.0f) rd = fminf(rd, b.z);
            const float w = fmaxf(0.0f, dot(nt, mk(a.x, a.y, a.z)));
            const float dw = depth_weight(w, g.sharpness);
            if (w > 1e-6f) {
                sr = sr + w * a.w;
                sgr = sgr + w * b.x;
                sb = sb + w * b.y;
                tw = tw + w;
            }
            if (dw > 1e-6f) {
                d0 = d0 + dw * rd;
                d1 = d1 + dw * (rd * rd);
                tdw = tdw + dw;
            }
        }
        float nr = 0.0f, ng = 0.0f, nb = 0.0f, nd0 = 0.0f, nd1 = 0.0f;
        if (tw > 0.0f) { nr = sr / tw; ng = sgr / tw; nb = sb / tw; }
        if (tdw > 0.0f) { nd0 = d0 / tdw; nd1 = d1 / tdw; }
        const size_t q = tile_texel(g, p, tx, ty);
        if (!first) {
            const float h = g.hysteresis, h1 = 1.0f - g.hysteresis;
            const float4 pi = atl[2 * q], pd = atl[2 * q + 1];
            nr = h * pi.x + h1 * nr;
            ng = h * pi.y + h1 * ng;
            nb = h * pi.z + h1 * nb;
            nd0 = h * pd.x + h1 * nd0;
            nd1 = h * pd.y + h1 * nd1;
        }
        atl[2 * q] = make_float4(nr, ng, nb, 0.0f);
        atl[2 * q + 1] = make_float4(nd0, nd1, 0.0f, 0.0f);
    }
}

// texel records from the planar atlases (rt_ddgi_set_atlas)
__global__ void __launch_bounds__(DG_BLOCK) k_ddgi_pack(uint32_t n, const float* irr, const float* dep, float4* atl) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    atl[2 * (size_t)i] = make_float4(irr[4 * (size_t)i], irr[4 * (size_t)i + 1], irr[4 * (size_t)i + 2], 0.0f);
    atl[2 * (size_t)i + 1] = make_float4(dep[2 * (size_t)i], dep[2 * (size_t)i + 1], 0.0f, 0.0f);
}

// planar atlases from the texel records (rt_ddgi_get_atlas); either output may be NULL
__global__ void __launch_bounds__(DG_BLOCK) k_ddgi_unpack(uint32_t n, const float4* atl, float* irr, float* dep) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 a = atl[2 * (size_t)i], b = atl[2 * (size_t)i + 1];
    if (irr) {
        irr[4 * (size_t)i] = a.x; irr[4 * (size_t)i + 1] = a.y; irr[4 * (size_t)i + 2] = a.z;
        irr[4 * (size_t)i + 3] = 0.0f;
    }
    if (dep) {
        dep[2 * (size_t)i] = b.x; dep[2 * (size_t)i + 1] = b.y;
    }
}

__global__ void __launch_bounds__(DG_BLOCK) k_ddgi_eval(DdgiGrid g, const float4* atl, const float* f, uint32_t n,
                                                       float* out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* q = f + 9 * (size_t)i;
    const f3 e = ddgi_lookup(g, atl, mk(q[0], q[1], q[2]), mk(q[3], q[4], q[5]), mk(q[6], q[7], q[8]));
    out[3 * (size_t)i] = e.x; out[3 * (size_t)i + 1] = e.y; out[3 * (size_t)i + 2] = e.z;
}

// acc[3 idx[i] + ch] += w[3 i + ch] * lookup_i[ch]
// n_dev (may be NULL): the query count is read from device memory (at most n_max)
__global__ void __launch_bounds__(DG_BLOCK) k_ddgi_scatter(DdgiGrid g, const float4* atl, const float* f,
                                                          const float* w, const uint32_t* idx, uint32_t n_max,
                                                          const uint32_t* n_dev, float* acc) {
    const uint32_t n = n_dev ? min(*n_dev, n_max) : n_max;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float* q = f + 9 * (size_t)i;
        const f3 e = ddgi_lookup(g, atl, mk(q[0], q[1], q[2]), mk(q[3], q[4], q[5]), mk(q[6], q[7], q[8]));
        float* a = acc + 3 * (size_t)idx[i];
        atomicAdd(a, w[3 * (size_t)i] * e.x);
        atomicAdd(a + 1, w[3 * (size_t)i + 1] * e.y);
        atomicAdd(a + 2, w[3 * (size_t)i + 2] * e.z);
    }
}

inline uint32_t blocks(uint64_t n) { return (uint32_t)((n + DG_BLOCK - 1) / DG_BLOCK); }

// Uniform random rotation from three uniforms (Shoemake, "Uniform random
// rotations", Graphics Gems III, 1992): unit quaternion, then its matrix.
void random_rotation(uint64_t seed, uint64_t update, float m[9]) {
    const u4 r = philox4x32_10(u4{(uint32_t)update, (uint32_t)(update >> 32), 0u, 0xDD610000u}, (uint32_t)seed,
                               (uint32_t)(seed >> 32));
    const double u1 = (double)(r.x >> 8) / 16777216.0, u2 = (double)(r.y >> 8) / 16777216.0,
                 u3 = (double)(r.z >> 8) / 16777216.0;
    const double a = sqrt(1.0 - u1), b = sqrt(u1), t2 = 2.0 * M_PI * u2, t3 = 2.0 * M_PI * u3;
    const double x = a * sin(t2), y = a * cos(t2), z = b * sin(t3), w = b * cos(t3);
    const double q[9] = {1 - 2 * (y * y + z * z), 2 * (x * y - z * w),     2 * (x * z + y * w),
                         2 * (x * y + z * w),     1 - 2 * (x * x + z * z), 2 * (y * z - x * w),
                         2 * (x * z - y * w),     2 * (y * z + x * w),     1 - 2 * (x * x + y * y)};
    for (int k = 0; k < 9; ++k) m[k] = (float)q[k];
}

int launch_update(rt_ddgi* g) {
    rt_ctx* c = g->ctx;
    // probes per workgroup: as many as its threads have texels for and its
    // LDS has surfels for
    const uint32_t texels = (g->cfg.side + 2) * (g->cfg.side + 2);
    uint32_t ppb = DG_BLOCK / texels;
    if (ppb > MAX_RAYS / g->R) ppb = MAX_RAYS / g->R;
    if (ppb < 1) ppb = 1;
    hipLaunchKernelGGL(k_ddgi_update, dim3((g->P + ppb - 1) / ppb), dim3(DG_BLOCK), 0, c->stream, g->grid,
                       (const float*)g->surf.ptr, (float4*)g->atl.ptr, g->updates == 0 ? 1u : 0u, ppb, g->P);
    RT_HIP(c, hipGetLastError());
    g->updates += 1;
    return RT_OK;
}

int stage_in(rt_ddgi* g, DevBuf& b, const void* src, size_t bytes, int ptr_kind, const void** out) {
    if (ptr_kind == RT_PTR_DEVICE) {
        *out = src;
        return RT_OK;
    }
    int rc = ensure(g->ctx, b, bytes);
    if (rc) return rc;
    RT_HIP(g->ctx, hipMemcpyAsync(b.ptr, src, bytes, hipMemcpyHostToDevice, g->ctx->stream));
    *out = b.ptr;
    return RT_OK;
}

int check_kind(rt_ddgi* g, int ptr_kind, const char* fn) {
    if (ptr_kind != RT_PTR_HOST && ptr_kind != RT_PTR_DEVICE) return dg_fail(g, RT_E_INVALID, std::string(fn) + ": bad ptr_kind");
    return RT_OK;
}

}  // namespace

namespace rt {

int ddgi_query_scatter(rt_ddgi* g, const float* d_feats, const float* d_w, const uint32_t* d_idx, uint32_t n,
                       float* d_acc, const uint32_t* d_n) {
    if (n == 0) return RT_OK;
    rt_ctx* c = g->ctx;
    // with a device count, a grid of a few waves per CU strides over the queries
    uint32_t grid = blocks(n);
    if (d_n && grid > 8u * (uint32_t)c->num_cus) grid = 8u * (uint32_t)c->num_cus;
    hipLaunchKernelGGL(k_ddgi_scatter, dim3(grid), dim3(DG_BLOCK), 0, c->stream, g->grid,
                       (const float4*)g->atl.ptr, d_feats, d_w, d_idx, n, d_n, d_acc);
    RT_HIP(c, hipGetLastError());
    return RT_OK;
}

rt_ctx* ddgi_ctx(const rt_ddgi* g) { return g->ctx; }

void ddgi_detach(rt_ddgi* g) {
    rt_ctx* c = g->ctx;
    if (!c) return;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    DevBuf* all[] = {&g->atl, &g->surf, &g->rays, &g->hits, &g->sh_ray, &g->sh_w, &g->sh_occ, &g->io_in, &g->io_out};
    for (DevBuf* b : all) {
        if (b->ptr) (void)hipFree(b->ptr);
        b->ptr = nullptr;
        b->bytes = 0;
    }
    for (size_t i = 0; i < c->probe_fields.size(); ++i)
        if (c->probe_fields[i] == g) {
            c->probe_fields.erase(c->probe_fields.begin() + (long)i);
            break;
        }
    g->ctx = nullptr;
}

}  // namespace rt

extern "C" {

int rt_ddgi_create(rt_ctx* c, const rt_ddgi_config* cfg, rt_ddgi** out) {
    if (!c) return RT_E_INVALID;
    if (!out || !cfg) return fail(c, RT_E_INVALID, "rt_ddgi_create: NULL argument");
    *out = nullptr;
    rt_ddgi_config k = *cfg;
    if (k.energy_preservation == 0.0f) k.energy_preservation = 1.0f;
    if (k.shadow_count == 0) k.shadow_count = 1;
    uint64_t P = 1;
    for (int a = 0; a < 3; ++a) {
        if (k.probe_count[a] == 0) return fail(c, RT_E_INVALID, "rt_ddgi_create: probe_count must be >= 1 on every axis");
        if (!(k.probe_spacing[a] > 0.0f)) return fail(c, RT_E_INVALID, "rt_ddgi_create: probe_spacing must be > 0");
        P *= k.probe_count[a];
    }
    if (k.rays_per_probe == 0 || k.rays_per_probe > MAX_RAYS)
        return fail(c, RT_E_INVALID, "rt_ddgi_create: rays_per_probe must be 1..1024");
    if (k.side < 2 || k.side > 32) return fail(c, RT_E_INVALID, "rt_ddgi_create: side must be 2..32");
    if (!(k.max_ray_distance > 0.0f)) return fail(c, RT_E_INVALID, "rt_ddgi_create: max_ray_distance must be > 0");
    if (!(k.hysteresis >= 0.0f && k.hysteresis <= 1.0f)) return fail(c, RT_E_INVALID, "rt_ddgi_create: hysteresis must be in [0, 1]");
    if (k.shadow_count > 64) return fail(c, RT_E_INVALID, "rt_ddgi_create: shadow_count must be 1..64");
    if (!(k.depth_sharpness >= 0.0f) || !(k.normal_bias >= 0.0f) || !(k.energy_preservation > 0.0f))
        return fail(c, RT_E_INVALID, "rt_ddgi_create: depth_sharpness, normal_bias must be >= 0, energy_preservation > 0");
    const uint64_t tiles_x = (uint64_t)k.probe_count[0] * k.probe_count[1];
    const uint64_t W = tiles_x * (k.side + 2), H = (uint64_t)k.probe_count[2] * (k.side + 2);
    if (W > (1u << 20) || H > (1u << 20) || W * H > (1ull << 28))
        return fail(c, RT_E_UNSUPPORTED, "rt_ddgi_create: atlas above 2^28 texels");
    if (P * k.rays_per_probe * k.shadow_count >= (1ull << 31))
        return fail(c, RT_E_UNSUPPORTED, "rt_ddgi_create: probes * rays * shadow_count must be below 2^31");
    rt_ddgi* g = new (std::nothrow) rt_ddgi();
    if (!g) return fail(c, RT_E_NOMEM, "rt_ddgi_create: host allocation failed");
    g->ctx = c;
    g->cfg = k;
    g->P = (uint32_t)P;
    g->R = k.rays_per_probe;
    g->W = (uint32_t)W;
    g->H = (uint32_t)H;
    DdgiGrid& d = g->grid;
    for (int a = 0; a < 3; ++a) {
        d.count[a] = k.probe_count[a];
        d.start[a] = k.probe_start[a];
        d.spacing[a] = k.probe_spacing[a];
    }
    d.side = k.side;
    d.tiles_x = (uint32_t)tiles_x;
    d.atlas_w = (uint32_t)W;
    d.rays = k.rays_per_probe;
    d.max_dist = k.max_ray_distance;
    d.sharpness = k.depth_sharpness;
    d.normal_bias = k.normal_bias;
    d.energy = k.energy_preservation;
    d.hysteresis = k.hysteresis;
    d.flags = k.flags;
    (void)hipSetDevice(c->device);
    c->probe_fields.push_back(g);  // from here on rt_ddgi_destroy releases whatever was allocated
    const size_t texels = (size_t)W * H;
    int rc;
    if ((rc = ensure(c, g->atl, 32 * texels)) ||
        (rc = ensure(c, g->surf, 32 * (size_t)P * g->R))) {
        rt_ddgi_destroy(g);
        return rc;
    }
    hipError_t e = hipMemsetAsync(g->atl.ptr, 0, 32 * texels, c->stream);
    if (e == hipSuccess) e = hipMemsetAsync(g->surf.ptr, 0, 32 * (size_t)P * g->R, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        rt_ddgi_destroy(g);
        return hip_fail(c, e, "rt_ddgi_create");
    }
    *out = g;
    return RT_OK;
}

void rt_ddgi_destroy(rt_ddgi* g) {
    if (!g) return;
    ddgi_detach(g);
    delete g;
}

int rt_ddgi_atlas_size(const rt_ddgi* g, uint32_t* width, uint32_t* height) {
    if (!g || !width || !height) return RT_E_INVALID;
    *width = g->W;
    *height = g->H;
    return RT_OK;
}

int rt_ddgi_update(rt_ddgi* g, const rt_settings* s) {
    if (!g) return RT_E_INVALID;
    if (!g->ctx) return RT_E_STATE;
    rt_ctx* c = g->ctx;
    if (!s) return dg_fail(g, RT_E_INVALID, "rt_ddgi_update: NULL settings");
    if (!c->built) return dg_fail(g, RT_E_STATE, "rt_ddgi_update: call rt_build first");
    (void)hipSetDevice(c->device);
    const uint32_t n = g->P * g->R;
    const uint32_t nc = n * g->cfg.shadow_count;
    int rc;
    if ((rc = ensure(c, g->rays, sizeof(rt_ray) * (size_t)n)) || (rc = ensure(c, g->hits, sizeof(rt_hit) * (size_t)n)) ||
        (rc = ensure(c, g->sh_ray, sizeof(rt_ray) * (size_t)nc)) || (rc = ensure(c, g->sh_w, 12 * (size_t)nc)) ||
        (rc = ensure(c, g->sh_occ, nc)))
        return rc;
    if (g->cfg.flags & RT_DDGI_ROTATE) random_rotation(s->seed, g->updates, g->rot);
    Rot rot;
    for (int k = 0; k < 9; ++k) rot.m[k] = g->rot[k];
    RT_HIP(c, hipEventRecord(c->ev[0], c->stream));
    hipLaunchKernelGGL(k_ddgi_rays, dim3(blocks(n)), dim3(DG_BLOCK), 0, c->stream, g->grid, rot, n,
                       (rt_ray*)g->rays.ptr, (float*)g->surf.ptr);
    RT_HIP(c, hipGetLastError());
    if ((rc = launch_trace(c, (const rt_ray*)g->rays.ptr, n, (rt_hit*)g->hits.ptr, false))) return rc;
    DdgiShade sh;
    sh.grid = g->grid;
    sh.rays = (const rt_ray*)g->rays.ptr;
    sh.hits = (const rt_hit*)g->hits.ptr;
    sh.surf = (float*)g->surf.ptr;
    sh.atl = (const float4*)g->atl.ptr;
    sh.multi = (g->cfg.flags & RT_DDGI_MULTI_BOUNCE) && g->updates > 0 ? 1u : 0u;
    sh.update = (uint32_t)g->updates;
    sh.count = g->cfg.shadow_count;
    sh.sh_ray = (rt_ray*)g->sh_ray.ptr;
    sh.sh_w = (float*)g->sh_w.ptr;
    sh.n = n;
    if ((rc = launch_ddgi_shade(c, s, sh))) return rc;
    if ((rc = launch_trace_any(c, (const rt_ray*)g->sh_ray.ptr, nc, (uint8_t*)g->sh_occ.ptr))) return rc;
    hipLaunchKernelGGL(k_ddgi_resolve, dim3(blocks(n)), dim3(DG_BLOCK), 0, c->stream, n, g->cfg.shadow_count,
                       (const float*)g->sh_w.ptr, (const uint8_t*)g->sh_occ.ptr, (float*)g->surf.ptr);
    RT_HIP(c, hipGetLastError());
    if ((rc = launch_update(g))) return rc;
    RT_HIP(c, hipEventRecord(c->ev[1], c->stream));
    uint32_t err = 0;
    if ((rc = trace_error(c, &err))) return rc;
    if (err) return dg_fail(g, RT_E_DEVICE, "rt_ddgi_update: traversal exceeded its stack or step bound");
    float ms = 0.0f;
    RT_HIP(c, hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
    c->timing.last_call_ms = ms;
    c->timing.launches = 7;
    c->last_op = 4;
    return RT_OK;
}

int rt_ddgi_update_surfels(rt_ddgi* g, const float* surfels, int ptr_kind) {
    if (!g) return RT_E_INVALID;
    if (!g->ctx) return RT_E_STATE;
    if (!surfels) return dg_fail(g, RT_E_INVALID, "rt_ddgi_update_surfels: NULL surfels");
    int rc;
    if ((rc = check_kind(g, ptr_kind, "rt_ddgi_update_surfels"))) return rc;
    rt_ctx* c = g->ctx;
    (void)hipSetDevice(c->device);
    const size_t bytes = 32 * (size_t)g->P * g->R;
    const hipMemcpyKind kind = ptr_kind == RT_PTR_HOST ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice;
    RT_HIP(c, hipMemcpyAsync(g->surf.ptr, surfels, bytes, kind, c->stream));
    if ((rc = launch_update(g))) return rc;
    RT_HIP(c, hipStreamSynchronize(c->stream));
    return RT_OK;
}

int rt_ddgi_surfels(rt_ddgi* g, float* surfels, float* rotation, int ptr_kind) {
    if (!g) return RT_E_INVALID;
    if (!g->ctx) return RT_E_STATE;
    int rc;
    if ((rc = check_kind(g, ptr_kind, "rt_ddgi_surfels"))) return rc;
    if (rotation)
        for (int k = 0; k < 9; ++k) rotation[k] = g->rot[k];
    if (!surfels) return RT_OK;
    rt_ctx* c = g->ctx;
    (void)hipSetDevice(c->device);
    const hipMemcpyKind kind = ptr_kind == RT_PTR_HOST ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
    RT_HIP(c, hipMemcpyAsync(surfels, g->surf.ptr, 32 * (size_t)g->P * g->R, kind, c->stream));
    RT_HIP(c, hipStreamSynchronize(c->stream));
    return RT_OK;
}

int rt_ddgi_get_atlas(rt_ddgi* g, float* irradiance, float* depth, int ptr_kind) {
    if (!g) return RT_E_INVALID;
    if (!g->ctx) return RT_E_STATE;
    int rc;
    if ((rc = check_kind(g, ptr_kind, "rt_ddgi_get_atlas"))) return rc;
    rt_ctx* c = g->ctx;
    (void)hipSetDevice(c->device);
    const size_t texels = (size_t)g->W * g->H;
    float* di = irradiance;
    float* dd = depth;
    if (ptr_kind == RT_PTR_HOST) {  // unpack into device staging, then copy out
        if ((rc = ensure(c, g->io_out, 24 * texels))) return rc;
        di = irradiance ? (float*)g->io_out.ptr : nullptr;
        dd = depth ? (float*)g->io_out.ptr + 4 * texels : nullptr;
    }
    hipLaunchKernelGGL(k_ddgi_unpack, dim3(blocks(texels)), dim3(DG_BLOCK), 0, c->stream, (uint32_t)texels,
                       (const float4*)g->atl.ptr, di, dd);
    RT_HIP(c, hipGetLastError());
    if (ptr_kind == RT_PTR_HOST) {
        if (irradiance) RT_HIP(c, hipMemcpyAsync(irradiance, di, 16 * texels, hipMemcpyDeviceToHost, c->stream));
        if (depth) RT_HIP(c, hipMemcpyAsync(depth, dd, 8 * texels, hipMemcpyDeviceToHost, c->stream));
    }
    RT_HIP(c, hipStreamSynchronize(c->stream));
    return RT_OK;
}

int rt_ddgi_set_atlas(rt_ddgi* g, const float* irradiance, const float* depth, int ptr_kind) {
    if (!g) return RT_E_INVALID;
    if (!g->ctx) return RT_E_STATE;
    if (!irradiance || !depth) return dg_fail(g, RT_E_INVALID, "rt_ddgi_set_atlas: NULL atlas");
    int rc;
    if ((rc = check_kind(g, ptr_kind, "rt_ddgi_set_atlas"))) return rc;
    rt_ctx* c = g->ctx;
    (void)hipSetDevice(c->device);
    const size_t texels = (size_t)g->W * g->H;
    const float* si = irradiance;
    const float* sd = depth;
    if (ptr_kind == RT_PTR_HOST) {  // stage on the device, then pack
        if ((rc = ensure(c, g->io_in, 24 * texels))) return rc;
        RT_HIP(c, hipMemcpyAsync(g->io_in.ptr, irradiance, 16 * texels, hipMemcpyHostToDevice, c->stream));
        RT_HIP(c, hipMemcpyAsync((float*)g->io_in.ptr + 4 * texels, depth, 8 * texels, hipMemcpyHostToDevice,
                                 c->stream));
        si = (const float*)g->io_in.ptr;
        sd = (const float*)g->io_in.ptr + 4 * texels;
    }
    hipLaunchKernelGGL(k_ddgi_pack, dim3(blocks(texels)), dim3(DG_BLOCK), 0, c->stream, (uint32_t)texels, si, sd,
                       (float4*)g->atl.ptr);
    RT_HIP(c, hipGetLastError());
    RT_HIP(c, hipStreamSynchronize(c->stream));
    if (g->updates == 0) g->updates = 1;  // later updates blend into the set field
    return RT_OK;
}

int rt_ddgi_eval(rt_ddgi* g, const float* samples, uint32_t n, float* out, int ptr_kind) {
    if (!g) return RT_E_INVALID;
    if (!g->ctx) return RT_E_STATE;
    int rc;
    if ((rc = check_kind(g, ptr_kind, "rt_ddgi_eval"))) return rc;
    if (n == 0) return RT_OK;
    if (!samples || !out) return dg_fail(g, RT_E_INVALID, "rt_ddgi_eval: NULL argument");
    rt_ctx* c = g->ctx;
    (void)hipSetDevice(c->device);
    const void* df;
    if ((rc = stage_in(g, g->io_in, samples, 36 * (size_t)n, ptr_kind, &df))) return rc;
    float* dout = out;
    if (ptr_kind == RT_PTR_HOST) {
        if ((rc = ensure(c, g->io_out, 12 * (size_t)n))) return rc;
        dout = (float*)g->io_out.ptr;
    }
    hipLaunchKernelGGL(k_ddgi_eval, dim3(blocks(n)), dim3(DG_BLOCK), 0, c->stream, g->grid, (const float4*)g->atl.ptr,
                       (const float*)df, n, dout);
    RT_HIP(c, hipGetLastError());
    if (ptr_kind == RT_PTR_HOST) {
        RT_HIP(c, hipMemcpyAsync(out, dout, 12 * (size_t)n, hipMemcpyDeviceToHost, c->stream));
        RT_HIP(c, hipStreamSynchronize(c->stream));
    }
    return RT_OK;
}

}  // extern "C"
