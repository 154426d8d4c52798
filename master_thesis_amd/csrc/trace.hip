// trace.hip — closest-hit traversal of the 4-wide BVH.
//
// Replaces Scene::traceRay's geometric part (software_renderer/graphic/
// scene.cpp:22-41 -> graphics_object.cpp:64-87 -> triangle.cpp:21-45 and the
// candidate enumeration of math/bounding_volume_hierachy.cpp:81-101).
//
// Result contract (identical to the CPU restatement in oracle/): among all
// triangles whose Moller-Trumbore test accepts the ray (t > 1e-7, t > tmin,
// t < tmax), the winner has the smallest squared distance from the ray origin
// to its hit point o + t*d; exact ties go to the lowest triangle index.  The
// reference enumerates every box-hit leaf and keeps the first strict minimum
// in its own tree order; with the tie rule made explicit the answer does not
// depend on the tree, so any BVH gives the same bits.
//
// Culling is conservative: a child box is skipped only when its entry
// distance exceeds t_best * (1 + 2^-10) + sigma, where sigma = 2^-18 *
// max|o| / max|d| covers the cancellation in |(o + t d) - o|^2 for hits close
// to a distant origin; boxes are padded outward at build time.
//
// Execution: persistent waves pull 64-ray batches from a global counter
// (Aila & Laine, "Understanding the efficiency of ray traversal on GPUs",
// HPG 2009).  Each lane keeps a 32-entry ring stack in LDS; deeper entries
// spill to a per-lane global area sized from the BVH depth.
#include "rt_internal.hpp"
#include "rt_core.hpp"

namespace rt {
namespace {

struct TraceArgs {
    const rt_ray* rays;
    rt_hit* hits;
    uint32_t n;
    const uint32_t* n_dev;  // when set, the ray count is read from device memory
    const rt_bvh4_node* nodes;
    const rt_tri_rec* tris;
    uint32_t* work;        // [0] = next ray batch, [1] = stack overflow flag
    uint32_t* spill;       // [spill_cap][grid threads]
    uint32_t spill_cap;
    uint32_t spill_stride;
    unsigned long long* stats;  // rays, nodes, tris, hits
};

template <bool STATS>
__device__ __forceinline__ rt_hit trace_one(const TraceArgs& a, const rt_ray& r, uint32_t* stk, uint32_t tid,
                                            uint32_t gtid, uint64_t& c_nodes, uint64_t& c_tris) {
    rt_hit out;
    out.tri = RT_MISS;
    out.t = 0.0f; out.u = 0.0f; out.v = 0.0f;
    const f3 o = mk(r.ox, r.oy, r.oz), d = mk(r.dx, r.dy, r.dz);
    const float dmax = fmaxf(fabsf(d.x), fmaxf(fabsf(d.y), fabsf(d.z)));
    const float omax = fmaxf(fabsf(o.x), fmaxf(fabsf(o.y), fabsf(o.z)));
    if (!(dmax > 0.0f) || !(dmax < INFINITY) || !(omax < INFINITY)) return out;
    const float tmin = r.tmin, tmax = r.tmax;
    const float sigma = 3.814697265625e-06f * omax / dmax;  // 2^-18 * max|o| / max|d|
    const float ix = safe_rcp(d.x), iy = safe_rcp(d.y), iz = safe_rcp(d.z);
    const float nox = -o.x * ix, noy = -o.y * iy, noz = -o.z * iz;
    const float tlo = fmaxf(tmin, 0.0f);

    float best_d2 = INFINITY, best_t = 0.0f, best_u = 0.0f, best_v = 0.0f;
    uint32_t best_id = RT_MISS;
    float tcull = tmax;

    uint32_t sp = 0;
    uint32_t cur = 0;  // root: inner node 0
    while (true) {
        if (!(cur & RT_LEAF_BIT)) {
            const rt_bvh4_node* nd = a.nodes + cur;
            const float4 lx = *(const float4*)nd->lo_x, hx = *(const float4*)nd->hi_x;
            const float4 ly = *(const float4*)nd->lo_y, hy = *(const float4*)nd->hi_y;
            const float4 lz = *(const float4*)nd->lo_z, hz = *(const float4*)nd->hi_z;
            const uint4 ch = *(const uint4*)nd->child;
            if (STATS) ++c_nodes;
            float key[4];
            uint32_t ref[4] = {ch.x, ch.y, ch.z, ch.w};
            const float lxa[4] = {lx.x, lx.y, lx.z, lx.w}, hxa[4] = {hx.x, hx.y, hx.z, hx.w};
            const float lya[4] = {ly.x, ly.y, ly.z, ly.w}, hya[4] = {hy.x, hy.y, hy.z, hy.w};
            const float lza[4] = {lz.x, lz.y, lz.z, lz.w}, hza[4] = {hz.x, hz.y, hz.z, hz.w};
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                float x0 = fmaf(lxa[c], ix, nox), x1 = fmaf(hxa[c], ix, nox);
                float y0 = fmaf(lya[c], iy, noy), y1 = fmaf(hya[c], iy, noy);
                float z0 = fmaf(lza[c], iz, noz), z1 = fmaf(hza[c], iz, noz);
                float tn = fmaxf(fmaxf(fminf(x0, x1), fminf(y0, y1)), fmaxf(fminf(z0, z1), tlo));
                float tf = fminf(fminf(fmaxf(x0, x1), fmaxf(y0, y1)), fminf(fmaxf(z0, z1), tcull));
                bool hit = (tn <= tf) && (ref[c] != RT_CHILD_EMPTY);
                key[c] = hit ? tn : INFINITY;
            }
            // sort the four (key, ref) pairs ascending: 5-comparator network
#define RT_CSWAP(i, j)                                                        \
    if (key[j] < key[i]) {                                                    \
        float tk = key[i]; key[i] = key[j]; key[j] = tk;                      \
        uint32_t tr = ref[i]; ref[i] = ref[j]; ref[j] = tr;                   \
    }
            RT_CSWAP(0, 1) RT_CSWAP(2, 3) RT_CSWAP(0, 2) RT_CSWAP(1, 3) RT_CSWAP(1, 2)
#undef RT_CSWAP
            if (key[0] < INFINITY) {
#pragma unroll
                for (int c = 3; c >= 1; --c) {
                    if (key[c] < INFINITY) {
                        const uint32_t slot = (sp & (STACK_LDS - 1)) * TRACE_BLOCK + tid;
                        if (sp >= (uint32_t)STACK_LDS) {
                            uint32_t deep = sp - STACK_LDS;
                            if (deep < a.spill_cap) a.spill[(size_t)deep * a.spill_stride + gtid] = stk[slot];
                            else a.work[1] = 1u;  // cannot happen for a stack sized from the depth
                        }
                        stk[slot] = ref[c];
                        ++sp;
                    }
                }
                cur = ref[0];
                continue;
            }
        } else {
            const uint32_t first = cur & RT_LEAF_FIRST_MASK;
            const uint32_t cnt = ((cur >> RT_LEAF_COUNT_SHIFT) & 15u) + 1u;
            for (uint32_t j = 0; j < cnt; ++j) {
                const float4* rp = (const float4*)(a.tris + first + j);
                const float4 p0 = rp[0], p1 = rp[1], p2 = rp[2];
                if (STATS) ++c_tris;
                float t, u, v;
                if (intersect_mt(o, d, mk(p0.x, p0.y, p0.z), mk(p1.x, p1.y, p1.z), mk(p2.x, p2.y, p2.z), t, u, v) &&
                    t > tmin && t < tmax) {
                    const uint32_t id = __float_as_uint(p0.w);
                    const float dd = dist2(o, hit_point(o, d, t));
                    if (closer(dd, id, best_d2, best_id)) {
                        best_d2 = dd; best_id = id; best_t = t; best_u = u; best_v = v;
                        tcull = fminf(tmax, fmaf(best_t, 1.0009765625f, sigma));
                    }
                }
            }
        }
        if (sp == 0) break;
        --sp;
        const uint32_t slot = (sp & (STACK_LDS - 1)) * TRACE_BLOCK + tid;
        cur = stk[slot];
        if (sp >= (uint32_t)STACK_LDS) {
            uint32_t deep = sp - STACK_LDS;
            if (deep < a.spill_cap) stk[slot] = a.spill[(size_t)deep * a.spill_stride + gtid];
        }
    }
    if (best_id != RT_MISS) {
        out.tri = best_id; out.t = best_t; out.u = best_u; out.v = best_v;
    }
    return out;
}

template <bool STATS>
__global__ void __launch_bounds__(TRACE_BLOCK) k_trace(TraceArgs a) {
    __shared__ uint32_t stk[STACK_LDS * TRACE_BLOCK];
    const uint32_t tid = threadIdx.x;
    const uint32_t gtid = blockIdx.x * TRACE_BLOCK + tid;
    const uint32_t lane = tid & 63u;
    uint64_t c_nodes = 0, c_tris = 0, c_hits = 0, c_rays = 0;
    const uint32_t n = a.n_dev ? *a.n_dev : a.n;
    while (true) {
        uint32_t base = 0;
        if (lane == 0) base = atomicAdd(a.work, 64u);
        base = __shfl(base, 0);
        if (base >= n) break;
        const uint32_t ri = base + lane;
        if (ri < n) {
            const rt_ray r = a.rays[ri];
            rt_hit h = trace_one<STATS>(a, r, stk, tid, gtid, c_nodes, c_tris);
            a.hits[ri] = h;
            if (STATS) { ++c_rays; c_hits += (h.tri != RT_MISS); }
        }
    }
    if (STATS) {
        atomicAdd(a.stats + 0, (unsigned long long)c_rays);
        atomicAdd(a.stats + 1, (unsigned long long)c_nodes);
        atomicAdd(a.stats + 2, (unsigned long long)c_tris);
        atomicAdd(a.stats + 3, (unsigned long long)c_hits);
    }
}

}  // namespace

int launch_trace(rt_ctx* c, const rt_ray* d_rays, uint32_t n, rt_hit* d_hits, bool stats, const uint32_t* n_dev) {
    if (n == 0) return RT_OK;
    int per_cu = 0;
    if (stats) RT_HIP(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_trace<true>, TRACE_BLOCK, 0));
    else RT_HIP(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_trace<false>, TRACE_BLOCK, 0));
    if (per_cu < 1) per_cu = 1;
    uint32_t grid = (uint32_t)(c->num_cus * per_cu);
    uint32_t need_blocks = (n + TRACE_BLOCK - 1) / TRACE_BLOCK;
    if (grid > need_blocks) grid = need_blocks;
    const uint32_t threads = grid * TRACE_BLOCK;
    const uint32_t max_stack = 3u * c->bvh_depth + 1u;
    const uint32_t spill_cap = max_stack > (uint32_t)STACK_LDS ? max_stack - STACK_LDS : 1u;
    int rc;
    if ((rc = ensure(c, c->spill, 4ull * spill_cap * threads))) return rc;
    if ((rc = ensure(c, c->work, 64))) return rc;
    if ((rc = ensure(c, c->stats, 64))) return rc;
    RT_HIP(c, hipMemsetAsync(c->work.ptr, 0, 8, c->stream));
    if (stats) RT_HIP(c, hipMemsetAsync(c->stats.ptr, 0, 32, c->stream));
    TraceArgs a;
    a.rays = d_rays;
    a.hits = d_hits;
    a.n = n;
    a.n_dev = n_dev;
    a.nodes = (const rt_bvh4_node*)c->nodes.ptr;
    a.tris = (const rt_tri_rec*)c->tris.ptr;
    a.work = (uint32_t*)c->work.ptr;
    a.spill = (uint32_t*)c->spill.ptr;
    a.spill_cap = spill_cap;
    a.spill_stride = threads;
    a.stats = (unsigned long long*)c->stats.ptr;
    RT_HIP(c, hipEventRecord(c->ev[2], c->stream));
    if (stats) hipLaunchKernelGGL(k_trace<true>, dim3(grid), dim3(TRACE_BLOCK), 0, c->stream, a);
    else hipLaunchKernelGGL(k_trace<false>, dim3(grid), dim3(TRACE_BLOCK), 0, c->stream, a);
    RT_HIP(c, hipGetLastError());
    RT_HIP(c, hipEventRecord(c->ev[3], c->stream));
    c->timing.launches = 1;
    return RT_OK;
}

}  // namespace rt
