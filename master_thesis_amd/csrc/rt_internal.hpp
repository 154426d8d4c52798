// rt_internal.hpp — context object and launcher declarations shared by the
// HIP translation units of libmtrt.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/rt.h"
#include "ddgi_field.hpp"

namespace rt {

// Per-triangle shading record in the caller's triangle order: what
// Triangle::getBarycentricCoords (triangle.cpp:47-55) and Scene::traceRay's
// normal interpolation (scene.cpp:43-53) read.  64 + 48 bytes.
struct ShadeGeom {
    float v0[3], e0[3], e1[3];
    float d00, d01, d11, denom;
    uint32_t mat;
    float pad[2];
};
struct ShadeNrm {
    float n0[3], n1[3], n2[3];
    float pad[3];
};

// Emissive triangle for light sampling: caller's triangle index and the
// cumulative area fraction up to and including it (last = 1).
struct LightTri {
    uint32_t tri;
    float cdf;
};

// Device buffer that grows on demand and is freed with the context.
struct DevBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
};

}  // namespace rt

struct rt_ctx {
    int device = 0;
    int num_cus = 0;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    std::string err;

    // scene
    uint32_t V = 0, T = 0, M = 0;
    float scene_lo[3] = {0, 0, 0}, scene_hi[3] = {0, 0, 0};
    rt::DevBuf pos, idx, mat, shade_geom, shade_nrm, vnrm;  // vnrm: per-vertex world normals
    rt::DevBuf lights;  // LightTri per emissive triangle of positive area
    uint32_t num_lights = 0;
    float light_area = 0.0f;
    // light objects (one per emissive material, in material order) for the
    // bidirectional tracer: uint2 {first, count} into a triangle-id list
    // grouped by material in triangle order
    rt::DevBuf light_objs, light_obj_tris;
    uint32_t num_light_objs = 0;
    // BVH
    bool built = false;
    rt::DevBuf nodes, tris, qnodes;  // build nodes (float boxes), triangle records, traversal nodes (64 B)
    uint32_t num_nodes = 0, bvh_depth = 0;
    rt_bvh_info info{};
    // cooperative traversal of small launches (groups of 2^g lanes per ray):
    // nodes with index < coop_split[g] are walked by every lane of a group,
    // the subtrees below them are dealt out; coop_gmax = the largest g the
    // tree has a split for (0: not used).  The collapse numbers nodes level by
    // level, so a split is the first index of a level.
    uint32_t coop_split[7] = {0, 0, 0, 0, 0, 0, 0};
    uint32_t coop_gmax = 0;
    uint32_t trace_coop = 1;         // launches with >= 2^trace_coop lanes per ray walk cooperatively (0: never; MTRT_TRACE_COOP)
    uint32_t trace_coop_nodes = 256;  // nodes per lane of a group at its split level (MTRT_TRACE_COOP_NODES)
    uint32_t trace_coop_gmax = 6;    // widest group: 2^trace_coop_gmax lanes (MTRT_TRACE_COOP_GMAX, 1..6)
    // build scratch
    rt::DevBuf keys[2], vals[2], hist, inner, leafbox, innerbox, boxtab, queue[2], counters;
    // trace scratch
    rt::DevBuf rays, hits, spill, work, stats;
    uint32_t work_half = 0;  // traversal: which half of `work` the next launch counts rays in
    int trace_per_cu[4] = {0, 0, 0, 0};  // traversal: resident workgroups per CU (closest, counting, any, mixed)
    uint32_t trace_max_steps = 0;  // traversal: per-ray step bound (0: from the tree size; MTRT_TRACE_MAX_STEPS)
    uint32_t trace_refill_min = 8;  // traversal: idle lanes that trigger a ray refill (MTRT_TRACE_REFILL)
    uint32_t trace_refill_large = 24;  // the same in launches too large for the unified loads (MTRT_TRACE_REFILL_LARGE)
    uint32_t pt_slots_log2 = 24;     // path tracers: at most 2^this paths in flight (MTRT_PT_SLOTS_LOG2)
    uint32_t trace_slice = 16384;   // traversal: rays per XCD slice, a multiple of 128 (MTRT_TRACE_SLICE)
    uint32_t trace_unified_max = 0;  // traversal: launches of at most this many rays load each step in one round (0: 4 per resident lane; MTRT_TRACE_UNIFIED_MAX)
    // build: node-step / triangle-step cost of the leaf test (MTRT_BVH_ALPHA);
    // 0.5 measured best on the 1M-triangle soup ray-cast (3% over 1.0), within
    // 1% of the best on the terrain
    float bvh_leaf_alpha = 0.5f;
    uint32_t* noalloc_scratch = nullptr;  // this device's copy of the allocation-free trace's scratch array
    // path tracer state
    rt::DevBuf pt_state, pt_queue[2], pt_image, pt_img8, pt_dimg, pt_dcolor, pt_ctr, pt_partial, pt_cams;
    rt::DevBuf pt_nrc_query, pt_nrc_train;  // cache queries (features, weights, pixel) and training records
    rt::DevBuf pt_nrc_sorted;               // training records in sample order, and the sort's key / index buffers
    rt::DevBuf pt_shadow;                   // shadow rays of light sampling (rays, weights, indices, flags)
    rt::DevBuf bd_vert, bd_state, bd_queue[2], bd_seg, bd_occ, bd_val;  // bidirectional tracer
    uint32_t nrc_records = 0, nrc_tcap = 0; // training records of the last cached render, record capacity
    std::vector<rt_nrc*> caches;            // live radiance caches; rt_destroy detaches them
    std::vector<rt_ddgi*> probe_fields;     // live irradiance-probe fields; rt_destroy detaches them
    uint32_t render_iters = 0;
    // pinned host copies of device counters (2 x 16 words) and their copies'
    // completion events: the path tracer's and the BVH build's loops read
    // them one step behind instead of draining the stream
    uint32_t* snap = nullptr;
    hipEvent_t snap_ev[2] = {nullptr, nullptr};

    // ev[0..1] bracket a whole call, ev[2..3] the last traversal kernel,
    // ev[4] the end of a cached render's path phase
    static constexpr int NUM_EV = 5;
    hipEvent_t ev[NUM_EV] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    rt_timing timing{};
    int last_op = 0;  // 0 = untimed, 1 = trace, 2 = render, 4 = probe update, 5 = post-filter
};

namespace rt {

// error helpers (rt_api.hip)
int fail(rt_ctx* c, int code, const std::string& msg);
int hip_fail(rt_ctx* c, hipError_t e, const char* what);
int ensure(rt_ctx* c, DevBuf& b, size_t bytes);
// allocates the pinned counter copies and their events once per context
int ensure_snap(rt_ctx* c);

#define RT_HIP(c, expr)                                      \
    do {                                                     \
        hipError_t _e = (expr);                              \
        if (_e != hipSuccess) return rt::hip_fail(c, _e, #expr); \
    } while (0)

// lbvh.hip
int build_lbvh(rt_ctx* c, uint32_t leaf_size, uint32_t flags);
// Stable LSD radix sort of n (key, value) pairs by the low 8 * passes bits of
// the keys, ping-ponging between keys[0]/vals[0] (input) and keys[1]/vals[1];
// *cur is the index of the buffers holding the result.  Uses c->hist.
int radix_sort_pairs(rt_ctx* c, uint32_t* keys[2], uint32_t* vals[2], uint32_t n, uint32_t passes, int* cur);

// trace.hip
constexpr int TRACE_BLOCK = 256;
// 16 LDS entries per lane (16 KiB per 256-lane block) leave room for six
// waves per SIMD; deeper entries go to the global spill area.
constexpr int STACK_LDS = 16;
// n: ray count, or an upper bound on it when n_dev (device-resident count) is
// set.  zero0 / zero1: device words the launch sets to 0 before any later
// launch on the stream runs (queue counters of a wavefront loop, so that the
// loop needs no separate memset).
int launch_trace(rt_ctx* c, const rt_ray* d_rays, uint32_t n, rt_hit* d_hits, bool stats,
                 const uint32_t* n_dev = nullptr, uint32_t* zero0 = nullptr, uint32_t* zero1 = nullptr);
// occlusion: d_occluded[i] = 1 when ray i hits any triangle in (tmin, tmax)
int launch_trace_any(rt_ctx* c, const rt_ray* d_rays, uint32_t n, uint8_t* d_occluded,
                     const uint32_t* n_dev = nullptr, uint32_t* zero0 = nullptr);
// Both in one launch: closest hit for the *n_dev rays of d_rays into d_hits
// and any hit for the *n2_dev rays of d_rays2 into d_occluded; capacity
// bounds the two counts together.
int launch_trace_mixed(rt_ctx* c, const rt_ray* d_rays, const uint32_t* n_dev, rt_hit* d_hits, const rt_ray* d_rays2,
                       const uint32_t* n2_dev, uint8_t* d_occluded, uint32_t capacity, uint32_t* zero0 = nullptr,
                       uint32_t* zero1 = nullptr);
// Sets the context's cooperative-traversal splits from the node count of each
// BVH level (level_start[l] = index of level l's first node, the last entry
// the node count).
void set_coop_splits(rt_ctx* c, const std::vector<uint32_t>& level_start);
// Traversal error of the launches since the last call (0 none, 1 stack
// overflow, 2 step bound; both mean a corrupt tree): read, then cleared.
// Synchronises the stream.
int trace_error(rt_ctx* c, uint32_t* err);
// Scene::isOccluded for n segments (start, end: 6 floats each)
int launch_occluded(rt_ctx* c, const float* d_segs, uint32_t n, uint8_t* d_occluded);
// Looks up the device address of the allocation-free trace's scratch array on
// the context's device (once per device; the device must be current).
int noalloc_scratch_init(rt_ctx* c);
// rt_trace_closest_noalloc's launch: zeroes the scratch's counters and
// *d_status (may be NULL) with a fill kernel, then traces with the scratch
int trace_noalloc(rt_ctx* c, const rt_ray* d_rays, uint32_t n, rt_hit* d_hits, uint32_t* d_status);

// nrc.hip: cache inference / training on device-resident samples
int nrc_infer_device(rt_nrc* s, const float* d_feats, uint32_t n, float* d_out);
// acc[3 * idx[i] + ch] += w[3 * i + ch] * max(prediction_i[ch], 0) (atomically)
int nrc_infer_scatter(rt_nrc* s, const float* d_feats, const float* d_w, const uint32_t* d_idx, uint32_t n,
                      float* d_acc, const uint32_t* d_n = nullptr);
// up to max_steps Adam steps over `count` records (feats 9 + targets 3
// floats each, in two arrays); batch b of `batch` records takes records
// b + stride * j with stride = count / batch.  Returns the steps taken.
int nrc_train_records(rt_nrc* s, const float* d_feats, const float* d_targets, uint32_t count, uint32_t batch,
                      uint32_t max_steps, uint32_t* steps, float* last_loss);
rt_ctx* nrc_ctx(const rt_nrc* s);
// frees a cache's device memory and cuts it from its context (rt_destroy)
void nrc_detach(rt_nrc* s);

// pathtrace.hip
// Radiance-cache options of one frame (rt_render_nrc)
struct NrcFrame {
    rt_nrc* cache = nullptr;
    rt_nrc_render opt{};
    rt_nrc_stats* stats = nullptr;
    rt_ddgi* ddgi = nullptr;  // set: the probe field answers the queries (rt_render_ddgi)
};
// ddgi.hip: probe-field lookups of queued cache queries, added to their
// pixels: acc[3 * idx[i] + ch] += w[3 * i + ch] * lookup_i[ch]
int ddgi_query_scatter(rt_ddgi* g, const float* d_feats, const float* d_w, const uint32_t* d_idx, uint32_t n,
                       float* d_acc, const uint32_t* d_n = nullptr);
rt_ctx* ddgi_ctx(const rt_ddgi* g);
// frees a field's device memory and cuts it from its context (rt_destroy)
void ddgi_detach(rt_ddgi* g);
// pathtrace.hip: radiance of a probe update's rays (surfel floats 3..7) and
// their light samples (count per ray; unused slots get a zero direction)
struct DdgiShade {
    DdgiGrid grid;
    const rt_ray* rays;
    const rt_hit* hits;
    float* surf;              // n * 8
    const float4* atl;        // the field before this update (multi-bounce term), texel records
    uint32_t multi;           // add the field's diffuse term at surface hits
    uint32_t update;          // update index: Philox counter word
    uint32_t count;           // light samples per ray
    rt_ray* sh_ray;           // n * count
    float* sh_w;              // n * count * 3
    uint32_t n;               // probes * rays
};
int launch_ddgi_shade(rt_ctx* c, const rt_settings* s, const DdgiShade& d);

// bdpt.hip: one frame of the bidirectional tracer into device rgb / rgb8
int launch_bdpt(rt_ctx* c, const rt_camera* cam, const rt_bdpt_settings* s, float* d_rgb, uint8_t* d_rgb8);

// cams: V host cameras of one size; d_rgb / d_rgb8 / d_dimage view-major
int launch_render(rt_ctx* c, const rt_camera* cams, uint32_t V, const rt_settings* s, float* d_rgb,
                  uint8_t* d_rgb8, const float* d_dimage, float* d_dcolor, const NrcFrame* nrc = nullptr);

}  // namespace rt
