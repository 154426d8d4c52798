// rt_api.hip — the C ABI (include/rt.h): context, scene upload, build,
// trace and render dispatch.  Every entry point validates its arguments on the
// host before anything is launched, returns a status code and never throws.
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <new>
#include <string>

#include "rt_internal.hpp"
#include "rt_core.hpp"

namespace rt {

static thread_local std::string g_create_err;

int fail(rt_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg; else g_create_err = msg;
    return code;
}

int hip_fail(rt_ctx* c, hipError_t e, const char* what) {
    std::string m = std::string(what) + ": " + hipGetErrorString(e);
    int code = (e == hipErrorOutOfMemory) ? RT_E_NOMEM : RT_E_DEVICE;
    return fail(c, code, m);
}

int ensure(rt_ctx* c, DevBuf& b, size_t bytes) {
    if (bytes == 0) bytes = 4;
    if (b.bytes >= bytes) return RT_OK;
    if (b.ptr) {
        hipError_t e = hipFree(b.ptr);  // waits for work that may still use it
        b.ptr = nullptr;
        b.bytes = 0;
        if (e != hipSuccess) return hip_fail(c, e, "hipFree");
    }
    hipError_t e = hipMalloc(&b.ptr, bytes);
    if (e != hipSuccess) {
        b.ptr = nullptr;
        return fail(c, RT_E_NOMEM, "device allocation of " + std::to_string(bytes) + " bytes failed: " +
                                       hipGetErrorString(e));
    }
    b.bytes = bytes;
    return RT_OK;
}

int ensure_snap(rt_ctx* c) {
    if (c->snap) return RT_OK;
    RT_HIP(c, hipHostMalloc((void**)&c->snap, 2 * 16 * sizeof(uint32_t), hipHostMallocDefault));
    for (int k = 0; k < 2; ++k) RT_HIP(c, hipEventCreateWithFlags(&c->snap_ev[k], hipEventDisableTiming));
    return RT_OK;
}

static void release(rt_ctx* c, DevBuf& b) {
    if (b.ptr) (void)hipFree(b.ptr);
    b.ptr = nullptr;
    b.bytes = 0;
}

namespace {

__global__ void k_shade_records(uint32_t T, const float* __restrict__ pos, const float* __restrict__ nrm,
                                const uint32_t* __restrict__ idx, const uint32_t* __restrict__ tri_mat,
                                ShadeGeom* __restrict__ g, ShadeNrm* __restrict__ nr) {
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= T) return;
    uint32_t a = idx[3 * t], b = idx[3 * t + 1], c = idx[3 * t + 2];
    ShadeGeom s;
    ShadeNrm q;
    for (int k = 0; k < 3; ++k) {
        float p0 = pos[3 * a + k];
        s.v0[k] = p0;
        s.e0[k] = pos[3 * b + k] - p0;
        s.e1[k] = pos[3 * c + k] - p0;
        q.n0[k] = nrm[3 * a + k];
        q.n1[k] = nrm[3 * b + k];
        q.n2[k] = nrm[3 * c + k];
    }
    // Triangle constructor, triangle.cpp:10-13
    f3 e0 = mk(s.e0[0], s.e0[1], s.e0[2]), e1 = mk(s.e1[0], s.e1[1], s.e1[2]);
    s.d00 = dot(e0, e0);
    s.d01 = dot(e0, e1);
    s.d11 = dot(e1, e1);
    s.denom = s.d00 * s.d11 - s.d01 * s.d01;
    s.mat = tri_mat[t];
    s.pad[0] = s.pad[1] = 0.0f;
    q.pad[0] = q.pad[1] = q.pad[2] = 0.0f;
    g[t] = s;
    nr[t] = q;
}

struct DeviceGuard {
    explicit DeviceGuard(int d) { (void)hipSetDevice(d); }
};

}  // namespace
}  // namespace rt

using namespace rt;

extern "C" {

int rt_create(rt_ctx** out, int device) {
    if (!out) return fail(nullptr, RT_E_INVALID, "rt_create: out is NULL");
    *out = nullptr;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0)
        return fail(nullptr, RT_E_DEVICE, std::string("rt_create: no HIP device (") + hipGetErrorString(e) + ")");
    if (device < 0 || device >= n)
        return fail(nullptr, RT_E_INVALID, "rt_create: device " + std::to_string(device) + " out of range");
    hipDeviceProp_t prop;
    e = hipGetDeviceProperties(&prop, device);
    if (e != hipSuccess) return fail(nullptr, RT_E_DEVICE, std::string("hipGetDeviceProperties: ") + hipGetErrorString(e));
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(nullptr, RT_E_DEVICE, std::string("libmtrt is built for gfx950; device reports ") + prop.gcnArchName);
    e = hipSetDevice(device);
    if (e != hipSuccess) return fail(nullptr, RT_E_DEVICE, std::string("hipSetDevice: ") + hipGetErrorString(e));
    rt_ctx* c = new (std::nothrow) rt_ctx();
    if (!c) return fail(nullptr, RT_E_NOMEM, "rt_create: host allocation failed");
    c->device = device;
    c->num_cus = prop.multiProcessorCount;
    if (const char* r = getenv("MTRT_TRACE_REFILL")) {  // tuning knob, 1..64
        long v = strtol(r, nullptr, 10);
        c->trace_refill_min = (uint32_t)(v < 1 ? 1 : (v > 64 ? 64 : v));
    }
    if (const char* r = getenv("MTRT_TRACE_REFILL_LARGE")) {  // tuning knob, 1..64
        long v = strtol(r, nullptr, 10);
        c->trace_refill_large = (uint32_t)(v < 1 ? 1 : (v > 64 ? 64 : v));
    }
    if (const char* r = getenv("MTRT_BVH_ALPHA")) {  // tuning knob, >= 0
        const float v = strtof(r, nullptr);
        c->bvh_leaf_alpha = v >= 0.0f ? v : 0.0f;
    }
    if (const char* r = getenv("MTRT_TRACE_MAX_STEPS")) {  // test hook: a smaller per-ray step bound
        const unsigned long v = strtoul(r, nullptr, 10);
        c->trace_max_steps = (uint32_t)(v > 0xFFFFFFFFul ? 0xFFFFFFFFul : v);
    }
    if (const char* r = getenv("MTRT_TRACE_UNIFIED_MAX")) {  // tuning knob: rays (0: the default rule)
        const unsigned long v = strtoul(r, nullptr, 10);
        c->trace_unified_max = (uint32_t)(v > 0xFFFFFFFFul ? 0xFFFFFFFFul : v);
    }
    if (const char* r = getenv("MTRT_TRACE_COOP")) {  // tuning knob, 0..6: smallest group width exponent (0: off)
        long v = strtol(r, nullptr, 10);
        c->trace_coop = (uint32_t)(v < 0 ? 0 : (v > 6 ? 6 : v));
    }
    if (const char* r = getenv("MTRT_TRACE_COOP_NODES")) {  // tuning knob, 1..4096
        long v = strtol(r, nullptr, 10);
        c->trace_coop_nodes = (uint32_t)(v < 1 ? 1 : (v > 4096 ? 4096 : v));
    }
    if (const char* r = getenv("MTRT_TRACE_COOP_GMAX")) {  // tuning knob, 1..6
        long v = strtol(r, nullptr, 10);
        c->trace_coop_gmax = (uint32_t)(v < 1 ? 1 : (v > 6 ? 6 : v));
    }
    if (const char* r = getenv("MTRT_PT_SLOTS_LOG2")) {  // tuning knob, 16..24: path slots in flight
        long v = strtol(r, nullptr, 10);
        c->pt_slots_log2 = (uint32_t)(v < 16 ? 16 : (v > 24 ? 24 : v));
    }
    if (const char* r = getenv("MTRT_TRACE_SLICE")) {  // tuning knob, rounded to 128..2^24
        long v = strtol(r, nullptr, 10);
        v = v < 128 ? 128 : (v > (1l << 24) ? (1l << 24) : v);
        c->trace_slice = (uint32_t)(v & ~127l);
    }
    // a blocking stream: work here is ordered after earlier work on the null
    // stream (where e.g. torch's default stream writes the rays)
    e = hipStreamCreateWithFlags(&c->own_stream, hipStreamDefault);
    if (e != hipSuccess) {
        delete c;
        return fail(nullptr, RT_E_DEVICE, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    }
    c->stream = c->own_stream;
    if (noalloc_scratch_init(c) != RT_OK) {
        const std::string msg = c->err;
        rt_destroy(c);
        return fail(nullptr, RT_E_DEVICE, "rt_create: " + msg);
    }
    // the ev[] events only time work (host reads of device data wait on
    // snap_ev or a stream synchronisation), so they skip the system-scope
    // cache fence that an event record carries by default
    for (int k = 0; k < rt_ctx::NUM_EV; ++k) {
        e = hipEventCreateWithFlags(&c->ev[k], hipEventDisableSystemFence);
        if (e != hipSuccess) {
            rt_destroy(c);
            return fail(nullptr, RT_E_DEVICE, std::string("hipEventCreate: ") + hipGetErrorString(e));
        }
    }
    *out = c;
    return RT_OK;
}

void rt_destroy(rt_ctx* c) {
    if (!c) return;
    DeviceGuard g(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    // caches outliving their context keep only their host part
    std::vector<rt_nrc*> caches = c->caches;
    for (rt_nrc* n : caches) nrc_detach(n);
    std::vector<rt_ddgi*> fields = c->probe_fields;
    for (rt_ddgi* g : fields) ddgi_detach(g);
    DevBuf* all[] = {&c->pos, &c->idx, &c->mat, &c->shade_geom, &c->shade_nrm, &c->vnrm, &c->nodes, &c->qnodes, &c->tris,
                     &c->keys[0], &c->keys[1], &c->vals[0], &c->vals[1], &c->hist, &c->inner, &c->leafbox,
                     &c->innerbox, &c->boxtab, &c->queue[0], &c->queue[1], &c->counters, &c->rays, &c->hits,
                     &c->spill, &c->work, &c->stats, &c->pt_state, &c->pt_queue[0], &c->pt_queue[1],
                     &c->pt_image, &c->pt_img8, &c->pt_dimg, &c->pt_dcolor, &c->pt_ctr, &c->pt_partial,
                     &c->pt_cams, &c->pt_nrc_query, &c->pt_nrc_train, &c->pt_nrc_sorted, &c->pt_shadow, &c->lights,
                     &c->light_objs, &c->light_obj_tris, &c->bd_vert, &c->bd_state, &c->bd_queue[0],
                     &c->bd_queue[1], &c->bd_seg, &c->bd_occ, &c->bd_val};
    for (DevBuf* b : all) release(c, *b);
    for (int k = 0; k < rt_ctx::NUM_EV; ++k)
        if (c->ev[k]) (void)hipEventDestroy(c->ev[k]);
    for (int k = 0; k < 2; ++k)
        if (c->snap_ev[k]) (void)hipEventDestroy(c->snap_ev[k]);
    if (c->snap) (void)hipHostFree(c->snap);
    if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
    delete c;
}

const char* rt_last_error(const rt_ctx* c) {
    return c ? c->err.c_str() : g_create_err.c_str();
}

int rt_set_stream(rt_ctx* c, void* s) {
    if (!c) return RT_E_INVALID;
    // work on the old stream finishes first: consecutive launches share
    // device-side counters that each launch prepares for the next
    DeviceGuard g(c->device);
    RT_HIP(c, hipStreamSynchronize(c->stream));
    c->stream = s ? (hipStream_t)s : c->own_stream;
    return RT_OK;
}

int rt_get_timing(const rt_ctx* cc, rt_timing* out) {
    rt_ctx* c = const_cast<rt_ctx*>(cc);
    if (!c || !out) return RT_E_INVALID;
    DeviceGuard g(c->device);
    if (c->last_op == 1 || c->last_op == 5) {  // calls that bracket their kernel with ev[2] / ev[3]
        RT_HIP(c, hipEventSynchronize(c->ev[3]));
        float ms = 0.0f;
        RT_HIP(c, hipEventElapsedTime(&ms, c->ev[2], c->ev[3]));
        c->timing.last_kernel_ms = ms;
        if (c->last_op == 1) c->timing.last_call_ms = ms;
    } else if (c->last_op != 0) {
        c->timing.last_kernel_ms = 0.0f;  // renders and probe updates time the whole call only
    }
    *out = c->timing;
    return RT_OK;
}

int rt_scene_upload(rt_ctx* c, const rt_scene_desc* s) {
    if (!c) return RT_E_INVALID;
    if (!s || !s->pos || !s->nrm || !s->idx || !s->tri_mat || !s->mat)
        return fail(c, RT_E_INVALID, "rt_scene_upload: NULL array in scene description");
    if (s->num_vertices == 0 || s->num_triangles == 0 || s->num_materials == 0)
        return fail(c, RT_E_INVALID, "rt_scene_upload: empty scene");
    if (s->num_triangles > RT_LEAF_FIRST_MASK)
        return fail(c, RT_E_UNSUPPORTED, "rt_scene_upload: more than 2^27-1 triangles");
    const uint64_t T = s->num_triangles, V = s->num_vertices, M = s->num_materials;
    for (uint64_t i = 0; i < 3 * T; ++i)
        if (s->idx[i] >= V)
            return fail(c, RT_E_INVALID, "rt_scene_upload: vertex index out of range at triangle " + std::to_string(i / 3));
    for (uint64_t i = 0; i < T; ++i)
        if (s->tri_mat[i] >= M)
            return fail(c, RT_E_INVALID, "rt_scene_upload: material index out of range at triangle " + std::to_string(i));
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint64_t v = 0; v < V; ++v)
        for (int k = 0; k < 3; ++k) {
            float x = s->pos[3 * v + k];
            if (!(fabsf(x) < INFINITY)) return fail(c, RT_E_INVALID, "rt_scene_upload: non-finite vertex position");
            lo[k] = fminf(lo[k], x);
            hi[k] = fmaxf(hi[k], x);
        }
    // emissive triangles and their cumulative area (light sampling of the
    // cached render); float arithmetic in triangle order
    std::vector<LightTri> lights;
    float area_sum = 0.0f;
    for (uint64_t t = 0; t < T; ++t) {
        if (!s->mat[s->tri_mat[t]].light_source) continue;
        const float* p0 = s->pos + 3 * s->idx[3 * t];
        const float* p1 = s->pos + 3 * s->idx[3 * t + 1];
        const float* p2 = s->pos + 3 * s->idx[3 * t + 2];
        const float area = tri_area(mk(p0[0], p0[1], p0[2]), mk(p1[0], p1[1], p1[2]), mk(p2[0], p2[1], p2[2]));
        if (!(area > 0.0f)) continue;
        area_sum = area_sum + area;
        lights.push_back(LightTri{(uint32_t)t, area_sum});
    }
    for (LightTri& l : lights) l.cdf = l.cdf / area_sum;
    if (!lights.empty()) lights.back().cdf = 1.0f;
    // light objects of the bidirectional tracer: every emissive material with
    // its triangles in triangle order (the reference's lightSources list, one
    // GraphicsObject per material, graphic/graphics_engine.cpp:20-30)
    std::vector<uint32_t> lobj_tris;
    std::vector<uint32_t> lobj;  // pairs {first, count}
    {
        std::vector<std::vector<uint32_t>> per(M);
        for (uint64_t t = 0; t < T; ++t)
            if (s->mat[s->tri_mat[t]].light_source) per[s->tri_mat[t]].push_back((uint32_t)t);
        for (uint64_t m = 0; m < M; ++m)
            if (s->mat[m].light_source && !per[m].empty()) {
                lobj.push_back((uint32_t)lobj_tris.size());
                lobj.push_back((uint32_t)per[m].size());
                lobj_tris.insert(lobj_tris.end(), per[m].begin(), per[m].end());
            }
    }
    DeviceGuard g(c->device);
    // from here on a failure leaves the context without a scene: buffers of
    // the previous one may already have been reallocated
    c->V = c->T = c->M = 0;
    c->built = false;
    c->num_lights = 0;
    c->num_light_objs = 0;
    c->light_area = 0.0f;
    int rc;
    if (!lights.empty()) {
        if ((rc = ensure(c, c->lights, sizeof(LightTri) * lights.size()))) return rc;
        RT_HIP(c, hipMemcpyAsync(c->lights.ptr, lights.data(), sizeof(LightTri) * lights.size(), hipMemcpyHostToDevice,
                                 c->stream));
    }
    if (!lobj.empty()) {
        if ((rc = ensure(c, c->light_objs, 4 * lobj.size())) || (rc = ensure(c, c->light_obj_tris, 4 * lobj_tris.size())))
            return rc;
        RT_HIP(c, hipMemcpyAsync(c->light_objs.ptr, lobj.data(), 4 * lobj.size(), hipMemcpyHostToDevice, c->stream));
        RT_HIP(c, hipMemcpyAsync(c->light_obj_tris.ptr, lobj_tris.data(), 4 * lobj_tris.size(), hipMemcpyHostToDevice,
                                 c->stream));
    }
    if ((rc = ensure(c, c->pos, 12 * V)) || (rc = ensure(c, c->idx, 12 * T)) ||
        (rc = ensure(c, c->mat, sizeof(rt_material) * M)) || (rc = ensure(c, c->shade_geom, sizeof(ShadeGeom) * T)) ||
        (rc = ensure(c, c->shade_nrm, sizeof(ShadeNrm) * T)))
        return rc;
    struct TmpBuf {  // freed on every return
        DevBuf b;
        ~TmpBuf() { if (b.ptr) (void)hipFree(b.ptr); }
    } tmp;
    DevBuf& tmat = tmp.b;
    if ((rc = ensure(c, c->vnrm, 12 * V)) || (rc = ensure(c, tmat, 4 * T))) return rc;
    RT_HIP(c, hipMemcpyAsync(c->pos.ptr, s->pos, 12 * V, hipMemcpyHostToDevice, c->stream));
    RT_HIP(c, hipMemcpyAsync(c->idx.ptr, s->idx, 12 * T, hipMemcpyHostToDevice, c->stream));
    RT_HIP(c, hipMemcpyAsync(c->mat.ptr, s->mat, sizeof(rt_material) * M, hipMemcpyHostToDevice, c->stream));
    RT_HIP(c, hipMemcpyAsync(c->vnrm.ptr, s->nrm, 12 * V, hipMemcpyHostToDevice, c->stream));
    RT_HIP(c, hipMemcpyAsync(tmat.ptr, s->tri_mat, 4 * T, hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_shade_records, dim3((uint32_t)((T + 255) / 256)), dim3(256), 0, c->stream, (uint32_t)T,
                       (const float*)c->pos.ptr, (const float*)c->vnrm.ptr, (const uint32_t*)c->idx.ptr,
                       (const uint32_t*)tmat.ptr, (ShadeGeom*)c->shade_geom.ptr, (ShadeNrm*)c->shade_nrm.ptr);
    RT_HIP(c, hipGetLastError());
    RT_HIP(c, hipStreamSynchronize(c->stream));
    c->V = (uint32_t)V;
    c->T = (uint32_t)T;
    c->M = (uint32_t)M;
    for (int k = 0; k < 3; ++k) { c->scene_lo[k] = lo[k]; c->scene_hi[k] = hi[k]; }
    c->num_lights = (uint32_t)lights.size();
    c->light_area = area_sum;
    c->num_light_objs = (uint32_t)(lobj.size() / 2);
    c->built = false;
    return RT_OK;
}

int rt_set_material_colors(rt_ctx* c, const float* rgb, int ptr_kind) {
    if (!c) return RT_E_INVALID;
    if (!rgb) return fail(c, RT_E_INVALID, "rt_set_material_colors: NULL colours");
    if (c->M == 0) return fail(c, RT_E_STATE, "rt_set_material_colors: no scene uploaded");
    DeviceGuard g(c->device);
    // colour is the first 12 bytes of each 32-byte rt_material record
    RT_HIP(c, hipMemcpy2DAsync(c->mat.ptr, sizeof(rt_material), rgb, 12, 12, c->M,
                               ptr_kind == RT_PTR_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, c->stream));
    if (ptr_kind != RT_PTR_DEVICE) RT_HIP(c, hipStreamSynchronize(c->stream));
    return RT_OK;
}

int rt_build(rt_ctx* c, const rt_build_opts* o, rt_bvh_info* info) {
    if (!c) return RT_E_INVALID;
    if (c->T == 0) return fail(c, RT_E_STATE, "rt_build: no scene uploaded");
    uint32_t leaf = (o && o->leaf_size) ? o->leaf_size : 4u;
    if (leaf > RT_MAX_LEAF_SIZE) return fail(c, RT_E_INVALID, "rt_build: leaf_size above 16");
    const uint32_t flags = o ? o->flags : 0u;
    if (flags & ~(uint32_t)RT_BUILD_CLUSTER) return fail(c, RT_E_INVALID, "rt_build: unknown flags");
    DeviceGuard g(c->device);
    int rc = build_lbvh(c, leaf, flags);
    if (rc) return rc;
    if (info) *info = c->info;
    return RT_OK;
}

int rt_bvh_export(rt_ctx* c, rt_bvh4_node* nodes, uint32_t* num_nodes, rt_tri_rec* tris, uint32_t* num_tris) {
    if (!c || !num_nodes || !num_tris) return RT_E_INVALID;
    if (!c->built) return fail(c, RT_E_STATE, "rt_bvh_export: no BVH built");
    uint32_t cap_n = *num_nodes, cap_t = *num_tris;
    *num_nodes = c->num_nodes;
    *num_tris = c->T;
    if (!nodes && !tris) return RT_OK;
    if ((nodes && cap_n < c->num_nodes) || (tris && cap_t < c->T))
        return fail(c, RT_E_INVALID, "rt_bvh_export: output capacity too small");
    DeviceGuard g(c->device);
    if (nodes)
        RT_HIP(c, hipMemcpyAsync(nodes, c->nodes.ptr, sizeof(rt_bvh4_node) * c->num_nodes, hipMemcpyDeviceToHost,
                                 c->stream));
    if (tris) RT_HIP(c, hipMemcpyAsync(tris, c->tris.ptr, sizeof(rt_tri_rec) * c->T, hipMemcpyDeviceToHost, c->stream));
    RT_HIP(c, hipStreamSynchronize(c->stream));
    return RT_OK;
}

int rt_bvh_export_traversal(rt_ctx* c, rt_bvh4q_node* nodes, uint32_t* num_nodes) {
    if (!c || !num_nodes) return RT_E_INVALID;
    if (!c->built) return fail(c, RT_E_STATE, "rt_bvh_export_traversal: no BVH built");
    const uint32_t cap = *num_nodes;
    *num_nodes = c->num_nodes;
    if (!nodes) return RT_OK;
    if (cap < c->num_nodes) return fail(c, RT_E_INVALID, "rt_bvh_export_traversal: output capacity too small");
    DeviceGuard g(c->device);
    RT_HIP(c, hipMemcpyAsync(nodes, c->qnodes.ptr, sizeof(rt_bvh4q_node) * c->num_nodes, hipMemcpyDeviceToHost,
                             c->stream));
    RT_HIP(c, hipStreamSynchronize(c->stream));
    return RT_OK;
}

int rt_trace_closest(rt_ctx* c, const rt_ray* rays, uint32_t n, rt_hit* hits, int ptr_kind,
                     rt_trace_counters* counters) {
    if (!c) return RT_E_INVALID;
    if (n > 0 && (!rays || !hits)) return fail(c, RT_E_INVALID, "rt_trace_closest: NULL rays or hits");
    if (n > RT_MAX_RAYS) return fail(c, RT_E_UNSUPPORTED, "rt_trace_closest: more than RT_MAX_RAYS rays in one call");
    if (ptr_kind != RT_PTR_HOST && ptr_kind != RT_PTR_DEVICE)
        return fail(c, RT_E_INVALID, "rt_trace_closest: bad ptr_kind");
    if (!c->built) return fail(c, RT_E_STATE, "rt_trace_closest: call rt_build first");
    if (counters) memset(counters, 0, sizeof(*counters));
    if (n == 0) return RT_OK;
    DeviceGuard g(c->device);
    int rc;
    const rt_ray* d_rays = rays;
    rt_hit* d_hits = hits;
    if (ptr_kind == RT_PTR_HOST) {
        if ((rc = ensure(c, c->rays, sizeof(rt_ray) * (size_t)n)) || (rc = ensure(c, c->hits, sizeof(rt_hit) * (size_t)n)))
            return rc;
        RT_HIP(c, hipMemcpyAsync(c->rays.ptr, rays, sizeof(rt_ray) * (size_t)n, hipMemcpyHostToDevice, c->stream));
        d_rays = (const rt_ray*)c->rays.ptr;
        d_hits = (rt_hit*)c->hits.ptr;
    }
    if ((rc = launch_trace(c, d_rays, n, d_hits, counters != nullptr))) return rc;
    c->last_op = 1;
    if (ptr_kind == RT_PTR_HOST)
        RT_HIP(c, hipMemcpyAsync(hits, d_hits, sizeof(rt_hit) * (size_t)n, hipMemcpyDeviceToHost, c->stream));
    if (ptr_kind == RT_PTR_HOST || counters) {
        unsigned long long st[4];
        if (counters) RT_HIP(c, hipMemcpyAsync(st, c->stats.ptr, 32, hipMemcpyDeviceToHost, c->stream));
        uint32_t err = 0;
        if ((rc = trace_error(c, &err))) return rc;
        if (err)
            return fail(c, RT_E_DEVICE, err == 1 ? "rt_trace_closest: traversal stack exceeded its bound"
                                                 : "rt_trace_closest: traversal exceeded its step bound");
        if (counters) {
            counters->rays = st[0];
            counters->node_visits = st[1];
            counters->tri_tests = st[2];
            counters->hits = st[3];
        }
    }
    return RT_OK;
}

int rt_trace_closest_noalloc(rt_ctx* c, const rt_ray* rays, uint32_t n, rt_hit* hits, uint32_t* status) {
    if (!c) return RT_E_INVALID;
    if (n > 0 && (!rays || !hits)) return fail(c, RT_E_INVALID, "rt_trace_closest_noalloc: NULL rays or hits");
    if (n > RT_MAX_RAYS)
        return fail(c, RT_E_UNSUPPORTED, "rt_trace_closest_noalloc: more than RT_MAX_RAYS rays in one call");
    if (!c->built) return fail(c, RT_E_STATE, "rt_trace_closest_noalloc: call rt_build first");
    if (n == 0) return RT_OK;
    DeviceGuard g(c->device);
    // no timing events are recorded around this call (rt_get_timing keeps
    // the last timed call's figures)
    c->last_op = 0;
    return trace_noalloc(c, rays, n, hits, status);
}

// any-hit queries: rays (32 B each) or segments (24 B each) -> one byte per query
static int occlusion_query(rt_ctx* c, const char* what, const void* in, size_t in_bytes, uint32_t n, uint8_t* out,
                           int ptr_kind, bool segments) {
    if (!c) return RT_E_INVALID;
    if (n > 0 && (!in || !out)) return fail(c, RT_E_INVALID, std::string(what) + ": NULL argument");
    if (n > RT_MAX_RAYS) return fail(c, RT_E_UNSUPPORTED, std::string(what) + ": more than RT_MAX_RAYS queries in one call");
    if (ptr_kind != RT_PTR_HOST && ptr_kind != RT_PTR_DEVICE) return fail(c, RT_E_INVALID, std::string(what) + ": bad ptr_kind");
    if (!c->built) return fail(c, RT_E_STATE, std::string(what) + ": call rt_build first");
    if (n == 0) return RT_OK;
    DeviceGuard g(c->device);
    int rc;
    const void* d_in = in;
    uint8_t* d_out = out;
    if (ptr_kind == RT_PTR_HOST) {
        if ((rc = ensure(c, c->rays, in_bytes * n)) || (rc = ensure(c, c->hits, n))) return rc;
        RT_HIP(c, hipMemcpyAsync(c->rays.ptr, in, in_bytes * n, hipMemcpyHostToDevice, c->stream));
        d_in = c->rays.ptr;
        d_out = (uint8_t*)c->hits.ptr;
    }
    if (segments) rc = launch_occluded(c, (const float*)d_in, n, d_out);
    else rc = launch_trace_any(c, (const rt_ray*)d_in, n, d_out);
    if (rc) return rc;
    c->last_op = 1;
    if (ptr_kind == RT_PTR_HOST) {
        RT_HIP(c, hipMemcpyAsync(out, d_out, n, hipMemcpyDeviceToHost, c->stream));
        uint32_t err = 0;
        if ((rc = trace_error(c, &err))) return rc;
        if (err) return fail(c, RT_E_DEVICE, std::string(what) + ": traversal exceeded its stack or step bound");
    }
    return RT_OK;
}

int rt_trace_any(rt_ctx* c, const rt_ray* rays, uint32_t n, uint8_t* occluded, int ptr_kind) {
    return occlusion_query(c, "rt_trace_any", rays, sizeof(rt_ray), n, occluded, ptr_kind, false);
}

int rt_occluded(rt_ctx* c, const float* segments, uint32_t n, uint8_t* occluded, int ptr_kind) {
    return occlusion_query(c, "rt_occluded", segments, 24, n, occluded, ptr_kind, true);
}

int rt_render(rt_ctx* c, const rt_camera* cam, const rt_settings* s, float* rgb, uint8_t* rgb8, int ptr_kind) {
    return rt_render_views(c, cam, 1, s, rgb, rgb8, ptr_kind);
}

int rt_render_views(rt_ctx* c, const rt_camera* cams, uint32_t num_views, const rt_settings* s, float* rgb,
                    uint8_t* rgb8, int ptr_kind) {
    if (!c) return RT_E_INVALID;
    if (!cams || !s) return fail(c, RT_E_INVALID, "rt_render: NULL camera or settings");
    if (num_views == 0) return fail(c, RT_E_INVALID, "rt_render_views: no views");
    if (!c->built) return fail(c, RT_E_STATE, "rt_render: call rt_build first");
    if (cams[0].width == 0 || cams[0].height == 0) return fail(c, RT_E_INVALID, "rt_render: empty image");
    for (uint32_t v = 1; v < num_views; ++v)
        if (cams[v].width != cams[0].width || cams[v].height != cams[0].height)
            return fail(c, RT_E_INVALID, "rt_render_views: views differ in size");
    if (s->spp == 0) return fail(c, RT_E_INVALID, "rt_render: spp must be >= 1");
    if (s->max_depth == 0 || s->max_depth > 64) return fail(c, RT_E_INVALID, "rt_render: max_depth must be 1..64");
    if (ptr_kind != RT_PTR_HOST && ptr_kind != RT_PTR_DEVICE) return fail(c, RT_E_INVALID, "rt_render: bad ptr_kind");
    DeviceGuard g(c->device);
    const size_t P = (size_t)cams[0].width * cams[0].height * num_views;
    float* d_rgb = rgb;
    uint8_t* d_rgb8 = rgb8;
    int rc;
    if (ptr_kind == RT_PTR_HOST) {
        if ((rc = ensure(c, c->pt_image, 12 * P)) || (rc = ensure(c, c->pt_img8, 3 * P))) return rc;
        d_rgb = (float*)c->pt_image.ptr;
        d_rgb8 = (uint8_t*)c->pt_img8.ptr;
    }
    if ((rc = launch_render(c, cams, num_views, s, d_rgb, d_rgb8, nullptr, nullptr))) return rc;
    if (ptr_kind == RT_PTR_HOST) {
        if (rgb) RT_HIP(c, hipMemcpyAsync(rgb, d_rgb, 12 * P, hipMemcpyDeviceToHost, c->stream));
        if (rgb8) RT_HIP(c, hipMemcpyAsync(rgb8, d_rgb8, 3 * P, hipMemcpyDeviceToHost, c->stream));
        RT_HIP(c, hipStreamSynchronize(c->stream));
    }
    return RT_OK;
}

int rt_render_nrc(rt_ctx* c, rt_nrc* nrc, const rt_camera* cams, uint32_t num_views, const rt_settings* s,
                  const rt_nrc_render* opts, float* rgb, uint8_t* rgb8, int ptr_kind, rt_nrc_stats* stats) {
    if (!c) return RT_E_INVALID;
    if (!nrc || !cams || !s) return fail(c, RT_E_INVALID, "rt_render_nrc: NULL cache, camera or settings");
    if (!nrc_ctx(nrc)) return fail(c, RT_E_STATE, "rt_render_nrc: the cache's context was destroyed");
    if (nrc_ctx(nrc) != c) return fail(c, RT_E_INVALID, "rt_render_nrc: the cache belongs to another context");
    if (num_views == 0) return fail(c, RT_E_INVALID, "rt_render_nrc: no views");
    if (!c->built) return fail(c, RT_E_STATE, "rt_render_nrc: call rt_build first");
    if (cams[0].width == 0 || cams[0].height == 0) return fail(c, RT_E_INVALID, "rt_render_nrc: empty image");
    for (uint32_t v = 1; v < num_views; ++v)
        if (cams[v].width != cams[0].width || cams[v].height != cams[0].height)
            return fail(c, RT_E_INVALID, "rt_render_nrc: views differ in size");
    if (s->spp == 0) return fail(c, RT_E_INVALID, "rt_render_nrc: spp must be >= 1");
    if (s->max_depth == 0 || s->max_depth > 64) return fail(c, RT_E_INVALID, "rt_render_nrc: max_depth must be 1..64");
    if (ptr_kind != RT_PTR_HOST && ptr_kind != RT_PTR_DEVICE) return fail(c, RT_E_INVALID, "rt_render_nrc: bad ptr_kind");
    NrcFrame f;
    f.cache = nrc;
    f.opt = opts ? *opts : rt_nrc_render{};
    if (f.opt.query_vertex == 0) f.opt.query_vertex = 1;
    if (f.opt.train_every == 0) f.opt.train_every = 16;
    if (f.opt.batch == 0) f.opt.batch = 16384;
    if (f.opt.max_steps == 0) f.opt.max_steps = 4;
    if (f.opt.query_vertex > 63) return fail(c, RT_E_INVALID, "rt_render_nrc: query_vertex must be 1..63");
    if (f.opt.batch % 64 != 0 || f.opt.batch > (1u << 22))
        return fail(c, RT_E_INVALID, "rt_render_nrc: batch must be a multiple of 64 up to 2^22");
    if (f.opt.max_steps > 64) return fail(c, RT_E_INVALID, "rt_render_nrc: max_steps must be <= 64");
    rt_nrc_stats local{};
    f.stats = stats ? stats : &local;
    *f.stats = rt_nrc_stats{};
    DeviceGuard g(c->device);
    const size_t P = (size_t)cams[0].width * cams[0].height * num_views;
    float* d_rgb = rgb;
    uint8_t* d_rgb8 = rgb8;
    int rc;
    if (ptr_kind == RT_PTR_HOST) {
        if ((rc = ensure(c, c->pt_image, 12 * P)) || (rc = ensure(c, c->pt_img8, 3 * P))) return rc;
        d_rgb = (float*)c->pt_image.ptr;
        d_rgb8 = (uint8_t*)c->pt_img8.ptr;
    }
    if ((rc = launch_render(c, cams, num_views, s, d_rgb, d_rgb8, nullptr, nullptr, &f))) return rc;
    if (ptr_kind == RT_PTR_HOST) {
        if (rgb) RT_HIP(c, hipMemcpyAsync(rgb, d_rgb, 12 * P, hipMemcpyDeviceToHost, c->stream));
        if (rgb8) RT_HIP(c, hipMemcpyAsync(rgb8, d_rgb8, 3 * P, hipMemcpyDeviceToHost, c->stream));
        RT_HIP(c, hipStreamSynchronize(c->stream));
    }
    return RT_OK;
}

int rt_render_ddgi(rt_ctx* c, rt_ddgi* g, const rt_camera* cams, uint32_t num_views, const rt_settings* s,
                   uint32_t query_vertex, float* rgb, uint8_t* rgb8, int ptr_kind, rt_nrc_stats* stats) {
    if (!c) return RT_E_INVALID;
    if (!g || !cams || !s) return fail(c, RT_E_INVALID, "rt_render_ddgi: NULL field, camera or settings");
    if (!ddgi_ctx(g)) return fail(c, RT_E_STATE, "rt_render_ddgi: the field's context was destroyed");
    if (ddgi_ctx(g) != c) return fail(c, RT_E_INVALID, "rt_render_ddgi: the field belongs to another context");
    if (num_views == 0) return fail(c, RT_E_INVALID, "rt_render_ddgi: no views");
    if (!c->built) return fail(c, RT_E_STATE, "rt_render_ddgi: call rt_build first");
    if (cams[0].width == 0 || cams[0].height == 0) return fail(c, RT_E_INVALID, "rt_render_ddgi: empty image");
    for (uint32_t v = 1; v < num_views; ++v)
        if (cams[v].width != cams[0].width || cams[v].height != cams[0].height)
            return fail(c, RT_E_INVALID, "rt_render_ddgi: views differ in size");
    if (s->spp == 0) return fail(c, RT_E_INVALID, "rt_render_ddgi: spp must be >= 1");
    if (s->max_depth == 0 || s->max_depth > 64) return fail(c, RT_E_INVALID, "rt_render_ddgi: max_depth must be 1..64");
    if (query_vertex > 63) return fail(c, RT_E_INVALID, "rt_render_ddgi: query_vertex must be 0..63");
    if (ptr_kind != RT_PTR_HOST && ptr_kind != RT_PTR_DEVICE) return fail(c, RT_E_INVALID, "rt_render_ddgi: bad ptr_kind");
    NrcFrame f;
    f.ddgi = g;
    f.opt.query_vertex = query_vertex;
    f.opt.flags = RT_NRC_NO_RECORDS | RT_NRC_NO_TRAIN;
    rt_nrc_stats local{};
    f.stats = stats ? stats : &local;
    *f.stats = rt_nrc_stats{};
    DeviceGuard dg(c->device);
    const size_t P = (size_t)cams[0].width * cams[0].height * num_views;
    float* d_rgb = rgb;
    uint8_t* d_rgb8 = rgb8;
    int rc;
    if (ptr_kind == RT_PTR_HOST) {
        if ((rc = ensure(c, c->pt_image, 12 * P)) || (rc = ensure(c, c->pt_img8, 3 * P))) return rc;
        d_rgb = (float*)c->pt_image.ptr;
        d_rgb8 = (uint8_t*)c->pt_img8.ptr;
    }
    if ((rc = launch_render(c, cams, num_views, s, d_rgb, d_rgb8, nullptr, nullptr, &f))) return rc;
    if (ptr_kind == RT_PTR_HOST) {
        if (rgb) RT_HIP(c, hipMemcpyAsync(rgb, d_rgb, 12 * P, hipMemcpyDeviceToHost, c->stream));
        if (rgb8) RT_HIP(c, hipMemcpyAsync(rgb8, d_rgb8, 3 * P, hipMemcpyDeviceToHost, c->stream));
        RT_HIP(c, hipStreamSynchronize(c->stream));
    }
    return RT_OK;
}

int rt_render_nrc_records(rt_ctx* c, float* feats, float* targets, uint32_t* count) {
    if (!c || !count) return RT_E_INVALID;
    const uint32_t cap = *count;
    *count = c->nrc_records;
    if (!feats && !targets) return RT_OK;
    if (cap < c->nrc_records) return fail(c, RT_E_INVALID, "rt_render_nrc_records: capacity below the record count");
    if (c->nrc_records == 0) return RT_OK;
    DeviceGuard g(c->device);
    const float* base = (const float*)c->pt_nrc_sorted.ptr;
    if (feats)
        RT_HIP(c, hipMemcpyAsync(feats, base, 36 * (size_t)c->nrc_records, hipMemcpyDeviceToHost, c->stream));
    if (targets)
        RT_HIP(c, hipMemcpyAsync(targets, base + 9 * (size_t)c->nrc_tcap, 12 * (size_t)c->nrc_records,
                                 hipMemcpyDeviceToHost, c->stream));
    RT_HIP(c, hipStreamSynchronize(c->stream));
    return RT_OK;
}

int rt_render_bdpt(rt_ctx* c, const rt_camera* cam, const rt_bdpt_settings* s, float* rgb, uint8_t* rgb8,
                   int ptr_kind) {
    if (!c) return RT_E_INVALID;
    if (!cam || !s) return fail(c, RT_E_INVALID, "rt_render_bdpt: NULL camera or settings");
    if (!c->built) return fail(c, RT_E_STATE, "rt_render_bdpt: call rt_build first");
    if (cam->width == 0 || cam->height == 0) return fail(c, RT_E_INVALID, "rt_render_bdpt: empty image");
    if ((uint64_t)cam->width * cam->height > (1ull << 31)) return fail(c, RT_E_UNSUPPORTED, "rt_render_bdpt: image above 2^31 pixels");
    if (s->spp == 0 || s->spp > (1u << 20)) return fail(c, RT_E_INVALID, "rt_render_bdpt: spp must be 1..2^20");
    if (s->vision_jumps == 0 || s->vision_jumps > 16 || s->light_jumps == 0 || s->light_jumps > 16)
        return fail(c, RT_E_INVALID, "rt_render_bdpt: vision_jumps and light_jumps must be 1..16");
    if (s->max_depth == 0 || s->max_depth > 32) return fail(c, RT_E_INVALID, "rt_render_bdpt: max_depth must be 1..32");
    if (ptr_kind != RT_PTR_HOST && ptr_kind != RT_PTR_DEVICE) return fail(c, RT_E_INVALID, "rt_render_bdpt: bad ptr_kind");
    DeviceGuard g(c->device);
    const size_t P = (size_t)cam->width * cam->height;
    float* d_rgb = rgb;
    uint8_t* d_rgb8 = rgb8;
    int rc;
    if (ptr_kind == RT_PTR_HOST) {
        if ((rc = ensure(c, c->pt_image, 12 * P)) || (rc = ensure(c, c->pt_img8, 3 * P))) return rc;
        d_rgb = (float*)c->pt_image.ptr;
        d_rgb8 = (uint8_t*)c->pt_img8.ptr;
    }
    if ((rc = launch_bdpt(c, cam, s, d_rgb, d_rgb8))) return rc;
    if (ptr_kind == RT_PTR_HOST) {
        if (rgb) RT_HIP(c, hipMemcpyAsync(rgb, d_rgb, 12 * P, hipMemcpyDeviceToHost, c->stream));
        if (rgb8) RT_HIP(c, hipMemcpyAsync(rgb8, d_rgb8, 3 * P, hipMemcpyDeviceToHost, c->stream));
        RT_HIP(c, hipStreamSynchronize(c->stream));
    }
    return RT_OK;
}

int rt_render_adjoint(rt_ctx* c, const rt_camera* cam, const rt_settings* s, const float* d_image, float* d_color,
                      int ptr_kind) {
    if (!c) return RT_E_INVALID;
    if (!cam || !s || !d_image || !d_color) return fail(c, RT_E_INVALID, "rt_render_adjoint: NULL argument");
    if (!c->built) return fail(c, RT_E_STATE, "rt_render_adjoint: call rt_build first");
    if (cam->width == 0 || cam->height == 0) return fail(c, RT_E_INVALID, "rt_render_adjoint: empty image");
    if (s->spp == 0) return fail(c, RT_E_INVALID, "rt_render_adjoint: spp must be >= 1");
    if (s->max_depth == 0 || s->max_depth > 64) return fail(c, RT_E_INVALID, "rt_render_adjoint: max_depth must be 1..64");
    if (ptr_kind != RT_PTR_HOST && ptr_kind != RT_PTR_DEVICE) return fail(c, RT_E_INVALID, "rt_render_adjoint: bad ptr_kind");
    DeviceGuard g(c->device);
    const size_t P = (size_t)cam->width * cam->height;
    const float* dimg = d_image;
    float* dcol = d_color;
    int rc;
    if (ptr_kind == RT_PTR_HOST) {
        if ((rc = ensure(c, c->pt_dimg, 12 * P)) || (rc = ensure(c, c->pt_dcolor, 12 * (size_t)c->M))) return rc;
        RT_HIP(c, hipMemcpyAsync(c->pt_dimg.ptr, d_image, 12 * P, hipMemcpyHostToDevice, c->stream));
        dimg = (const float*)c->pt_dimg.ptr;
        dcol = (float*)c->pt_dcolor.ptr;
    }
    if ((rc = launch_render(c, cam, 1, s, nullptr, nullptr, dimg, dcol))) return rc;
    if (ptr_kind == RT_PTR_HOST) {
        RT_HIP(c, hipMemcpyAsync(d_color, dcol, 12 * (size_t)c->M, hipMemcpyDeviceToHost, c->stream));
        RT_HIP(c, hipStreamSynchronize(c->stream));
    }
    return RT_OK;
}

}  // extern "C"
