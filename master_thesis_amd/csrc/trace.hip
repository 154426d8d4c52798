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
// Execution: persistent waves; each lane fetches a new ray as soon as its ray
// finishes (dynamic fetch, Aila & Laine, "Understanding the efficiency of ray
// traversal on GPUs", HPG 2009) from a wave-local pool of consecutive rays.
// One loop iteration is one traversal step.  Each lane keeps a 16-entry ring
// stack in LDS; deeper entries spill to a per-lane global area sized from the
// BVH depth.  Nodes are the 64-byte quantised form (rt_bvh4q_node, two per
// cache line): the slab test decodes each child plane with one FMA from two
// per-node, per-axis terms, taking the near and far bytes by the sign of the
// direction; the triangle test rejects most candidates before its IEEE
// division (intersect_mt_fast).
//
// Any-hit mode (occlusion, k_trace<.., true>): the same walk, ending a ray at
// the first triangle the test accepts; the output is one flag per ray.  As
// the box tests never reject a triangle the triangle test accepts, a ray is
// occluded exactly when the closest-hit walk finds a hit.  Segment queries
// (Scene::isOccluded, graphic/scene.cpp:58-83) build the ray from (start,
// end) as the reference does and accept a triangle when its hit point's
// squared distance from start is <= |end - start|^2.
#include <mutex>

#include "rt_internal.hpp"
#include "rt_core.hpp"

namespace rt {
namespace {

struct TraceArgs {
    const rt_ray* rays;
    rt_hit* hits;
    uint8_t* occluded;     // any-hit mode output
    const float* segs;     // any-hit mode: n * 6 floats (start, end) in place of rays

    uint32_t n;
    const uint32_t* n_dev;  // when set, the ray count is read from device memory
    const rt_ray* rays2;    // mixed launches: any-hit rays after the n closest-hit ones,
    const uint32_t* n2_dev; // their count (device), their flags into `occluded`
    const rt_bvh4q_node* nodes;  // the 64-byte quantised traversal nodes
    const rt_tri_rec* tris;
    uint32_t* work;        // [32 (1 + x)] = next ray of XCD x's slices (zero at launch)
    uint32_t* work_next;   // the other half of the buffer: zeroed here for the next launch
    uint32_t* err;         // sticky: 1 = stack overflow, 2 = step bound
    uint32_t* zero0;       // optional words zeroed for the launches after this one
    uint32_t* zero1;
    uint32_t* spill;       // [spill_cap][grid threads]
    uint32_t spill_cap;
    uint32_t spill_stride;
    uint32_t max_steps;    // bound on steps per ray (each node and triangle is visited at most once)
    uint32_t refill_min;   // idle lanes that trigger a refill (1..64) in launches of the unified form
    uint32_t refill_min_large;  // the same in larger launches
    uint32_t unified_max;  // launches of at most this many rays load each step's record in one round
    uint32_t slice;        // rays per XCD slice (a multiple of POOL_CHUNK)
    unsigned long long* stats;  // rays, nodes, tris, hits
    // cooperative walk (k_trace_coop): launches of at most coop_max rays go to
    // it (k_trace returns at once), with 2^g lanes per ray, g <= coop_gmax,
    // and the nodes below index coop_split[g] walked by every lane of a group
    uint32_t coop_max;
    uint32_t coop_gmax;
    uint32_t coop_split[7];
    uint32_t coop_zero;    // the cooperative launch is the only one: it zeroes zero0 / zero1
};

constexpr uint32_t POOL_CHUNK = 128;  // rays a wave takes from a counter at a time
// The ray range is cut into slices of TraceArgs::slice rays, dealt round-robin
// to the 8 XCDs (slice k to XCD k mod 8), each with its own counter.  A
// persistent grid's workgroup b runs on XCD b mod 8 (workgroups are dispatched
// to the XCDs in turn), so the rays in flight on one XCD, and the nodes its
// own L2 holds for them, come from a few consecutive slices instead of
// interleaving with the other seven.  A wave whose XCD has no rays left takes
// them from the next XCD's counter.
constexpr uint32_t XCDS = 8;
constexpr uint32_t XCD_COUNTER_STRIDE = 32;  // counters 128 B apart

// Per-lane traversal state of one ray.
struct Lane {
    f3 o, d;
    float ix, iy, iz, nox, noy, noz;
    float tmin, tmax, tlo, sigma, tcull;
    float best_d2, best_t, best_u, best_v;
    float seg_d2;  // segment query: accept hits with squared distance <= seg_d2
    uint32_t best_id;
    uint32_t negx, negy, negz;  // per axis: 1 when the direction is negative
    uint32_t cur, sp, steps;
    bool any;  // mixed launches: the ray is an any-hit one
};

__device__ __forceinline__ bool lane_start(Lane& L, const rt_ray& r) {
    L.o = mk(r.ox, r.oy, r.oz);
    L.d = mk(r.dx, r.dy, r.dz);
    const float dmax = fmaxf(fabsf(L.d.x), fmaxf(fabsf(L.d.y), fabsf(L.d.z)));
    const float omax = fmaxf(fabsf(L.o.x), fmaxf(fabsf(L.o.y), fabsf(L.o.z)));
    L.tmin = r.tmin;
    L.tmax = r.tmax;
    L.tlo = fmaxf(r.tmin, 0.0f);
    L.sigma = 3.814697265625e-06f * omax / dmax;  // 2^-18 * max|o| / max|d|
    L.ix = safe_rcp(L.d.x);
    L.iy = safe_rcp(L.d.y);
    L.iz = safe_rcp(L.d.z);
    L.nox = -L.o.x * L.ix;
    L.noy = -L.o.y * L.iy;
    L.noz = -L.o.z * L.iz;
    // for a negative direction the near plane is the high bound
    L.negx = __float_as_uint(L.ix) >> 31;
    L.negy = __float_as_uint(L.iy) >> 31;
    L.negz = __float_as_uint(L.iz) >> 31;
    L.best_d2 = INFINITY;
    L.best_t = L.best_u = L.best_v = 0.0f;
    L.best_id = RT_MISS;
    L.tcull = r.tmax;
    L.cur = 0;
    L.sp = 0;
    L.steps = 0;
    return (dmax > 0.0f) && (dmax < INFINITY) && (omax < INFINITY);
}

// One traversal step (one inner node, or one leaf's triangles, then a pop).
// Returns true when the ray is finished.
// the child reference a sort key names (its two low bits)
__device__ __forceinline__ uint32_t child_of(const uint32_t (&ref)[4], uint32_t key) {
    const uint32_t lo = (key & 1u) ? ref[1] : ref[0], hi = (key & 1u) ? ref[3] : ref[2];
    return (key & 2u) ? hi : lo;
}

// Where a step loads its record.  Split (launches with many rays): lanes at
// nodes load the node inside the node block and lanes at triangles the
// triangle inside the triangle block, so a wave's loads go out in two
// rounds; with the GPU full of rays this spreads the requests of the
// incoherent, cache-missing walks best.  Unified (launches with few rays,
// which are bound by how long their slowest walks take): every lane loads
// its record, a node's 64 bytes or a triangle's 48, before either block, so
// a step waits for one round of loads.  The choice is per launch, from the
// ray count (TraceArgs::unified_max), and in larger launches per wave: a
// wave whose rays all point into one octant (coherent rays, whose records
// the caches mostly hold) takes the unified form too.
// Cooperative walks (COOP): the group's lane `sub` of 2^g pushes a child of
// a node above the split (index < split) only when the child is above the
// split too or is dealt to it (coop_owner); below the split it pushes every
// child as usual.
__device__ __forceinline__ uint32_t coop_owner(uint32_t ref, uint32_t g) {
    return g ? (ref * 0x9E3779B1u) >> (32u - g) : 0u;
}

template <bool STATS, bool ANY, bool MIXED, bool COOP = false>
__device__ __forceinline__ bool lane_step(const TraceArgs& a, Lane& L, uint32_t* stk, uint32_t tid, uint32_t gtid,
                                          uint64_t& c_nodes, uint64_t& c_tris, bool unified, uint32_t split = 0,
                                          uint32_t g = 0, uint32_t sub = 0) {
    const bool at_leaf = (L.cur & RT_LEAF_BIT) != 0u;
    uint4 w0, w1, w2, ch;
    if (unified) {
        const uint4* rec =
            at_leaf ? (const uint4*)(a.tris + (L.cur & RT_LEAF_FIRST_MASK)) : (const uint4*)(a.nodes + L.cur);
        w0 = rec[0]; w1 = rec[1]; w2 = rec[2]; ch = rec[at_leaf ? 0 : 3];
        // (an empty asm that reads all four keeps the compiler from moving
        // parts of these loads into the blocks below)
        typedef uint32_t v4u __attribute__((ext_vector_type(4)));
        __asm__ __volatile__("" ::"v"(__builtin_bit_cast(v4u, w0)), "v"(__builtin_bit_cast(v4u, w1)),
                             "v"(__builtin_bit_cast(v4u, w2)), "v"(__builtin_bit_cast(v4u, ch)));
    }
    if (!at_leaf) {
        // the 64-byte node as four 16-byte loads: origin + exponents, the x
        // and y bytes, the z bytes, the child references.  The plane of byte
        // q is origin + q * 2^e, so its entry distance is
        // (origin - o) / d + q * (2^e / d): one FMA per plane from two
        // per-axis terms, and the near plane's bytes are the low ones when
        // the direction is positive.
        if (!unified) {
            const uint4* qp = (const uint4*)(a.nodes + L.cur);
            w0 = qp[0]; w1 = qp[1]; w2 = qp[2]; ch = qp[3];
        }
        if (STATS) ++c_nodes;
        const int ex = (int)(int8_t)(w0.w & 255u), ey = (int)(int8_t)((w0.w >> 8) & 255u),
                  ez = (int)(int8_t)((w0.w >> 16) & 255u);
        const float ax = fmaf(__uint_as_float(w0.x), L.ix, L.nox), bx = ldexpf(L.ix, ex);
        const float ay = fmaf(__uint_as_float(w0.y), L.iy, L.noy), by = ldexpf(L.iy, ey);
        const float az = fmaf(__uint_as_float(w0.z), L.iz, L.noz), bz = ldexpf(L.iz, ez);
        const uint32_t nxw = L.negx ? w1.y : w1.x, fxw = L.negx ? w1.x : w1.y;
        const uint32_t nyw = L.negy ? w1.w : w1.z, fyw = L.negy ? w1.z : w1.w;
        const uint32_t nzw = L.negz ? w2.y : w2.x, fzw = L.negz ? w2.x : w2.y;
        const uint32_t ref[4] = {ch.x, ch.y, ch.z, ch.w};
        // sort key of child c: its entry distance's bits (non-negative, sign
        // cleared so that -0 sorts as 0) with the two low mantissa bits
        // replaced by c, or 0x7F800000 | c (above every finite distance) when
        // the child is missed or empty; the keys then sort as integers and
        // name their children
        uint32_t pk[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const float tn = fmaxf(fmaxf(fmaf((float)((nxw >> (8 * c)) & 255u), bx, ax),
                                         fmaf((float)((nyw >> (8 * c)) & 255u), by, ay)),
                                   fmaxf(fmaf((float)((nzw >> (8 * c)) & 255u), bz, az), L.tlo));
            const float tf = fminf(fminf(fmaf((float)((fxw >> (8 * c)) & 255u), bx, ax),
                                         fmaf((float)((fyw >> (8 * c)) & 255u), by, ay)),
                                   fminf(fmaf((float)((fzw >> (8 * c)) & 255u), bz, az), L.tcull));
            pk[c] = ((tn <= tf && ref[c] != RT_CHILD_EMPTY) ? (__float_as_uint(tn) & 0x7FFFFFFCu) : 0x7F800000u) |
                    (uint32_t)c;
        }
        if (COOP && L.cur < split) {
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const bool above = (ref[c] & RT_LEAF_BIT) == 0u && ref[c] < split;
                if (!above && coop_owner(ref[c], g) != sub) pk[c] = 0x7F800000u | (uint32_t)c;
            }
        }
        // ascending: 5-comparator network on integer min / max
#define RT_CSWAP(i, j)                                  \
    {                                                   \
        const uint32_t lo = min(pk[i], pk[j]), hi = max(pk[i], pk[j]); \
        pk[i] = lo;                                     \
        pk[j] = hi;                                     \
    }
        RT_CSWAP(0, 1) RT_CSWAP(2, 3) RT_CSWAP(0, 2) RT_CSWAP(1, 3) RT_CSWAP(1, 2)
#undef RT_CSWAP
        if (pk[0] < 0x7F800000u) {
#pragma unroll
            for (int c = 3; c >= 1; --c) {
                if (pk[c] < 0x7F800000u) {
                    const uint32_t slot = (L.sp & (STACK_LDS - 1)) * TRACE_BLOCK + tid;
                    if (L.sp >= (uint32_t)STACK_LDS) {
                        const uint32_t deep = L.sp - STACK_LDS;
                        if (deep < a.spill_cap) a.spill[(size_t)deep * a.spill_stride + gtid] = stk[slot];
                        else *a.err = 1u;  // cannot happen for a stack sized from the depth
                    }
                    stk[slot] = child_of(ref, pk[c]);
                    ++L.sp;
                }
            }
            L.cur = child_of(ref, pk[0]);
            return false;
        }
    } else {
        // one triangle per step: a lane in a leaf costs one triangle test per
        // iteration, not the whole leaf, so lanes at inner nodes wait less;
        // the leaf reference shrinks in place to the remaining triangles
        const uint32_t first = L.cur & RT_LEAF_FIRST_MASK;
        const uint32_t rest = (L.cur >> RT_LEAF_COUNT_SHIFT) & 15u;  // triangles after this one
        if (!unified) {
            const uint4* rp = (const uint4*)(a.tris + first);
            w0 = rp[0]; w1 = rp[1]; w2 = rp[2];
        }
        const float4 p0 = __builtin_bit_cast(float4, w0), p1 = __builtin_bit_cast(float4, w1),
                     p2 = __builtin_bit_cast(float4, w2);
        if (STATS) ++c_tris;
        float t, u, v;
        if (intersect_mt_fast(L.o, L.d, mk(p0.x, p0.y, p0.z), mk(p1.x, p1.y, p1.z), mk(p2.x, p2.y, p2.z), L.tmin,
                              L.tmax, L.tcull, t, u, v) &&
            t > L.tmin && t < L.tmax) {
            const uint32_t id = __float_as_uint(p0.w);
            if (ANY || (MIXED && L.any)) {
                if (!a.segs || dist2(L.o, hit_point(L.o, L.d, t)) <= L.seg_d2) {
                    L.best_id = id;
                    return true;
                }
            } else {
            const float dd = dist2(L.o, hit_point(L.o, L.d, t));
            if (closer(dd, id, L.best_d2, L.best_id)) {
                L.best_d2 = dd; L.best_id = id; L.best_t = t; L.best_u = u; L.best_v = v;
                L.tcull = fminf(L.tmax, fmaf(L.best_t, 1.0009765625f, L.sigma));
            }
            }
        }
        if (rest) {
            L.cur = RT_LEAF_BIT | ((rest - 1u) << RT_LEAF_COUNT_SHIFT) | (first + 1u);
            return false;
        }
    }
    if (L.sp == 0) return true;
    if (++L.steps > a.max_steps) {  // cannot happen for a tree; ends the ray instead of looping
        *a.err = 2u;
        return true;
    }
    --L.sp;
    const uint32_t slot = (L.sp & (STACK_LDS - 1)) * TRACE_BLOCK + tid;
    L.cur = stk[slot];
    if (L.sp >= (uint32_t)STACK_LDS) {
        const uint32_t deep = L.sp - STACK_LDS;
        if (deep < a.spill_cap) stk[slot] = a.spill[(size_t)deep * a.spill_stride + gtid];
    }
    return false;
}

// Persistent waves; every lane takes a new ray as soon as its ray finishes
// (dynamic fetch, Aila & Laine 2009), from a wave-local pool of POOL_CHUNK
// consecutive rays refilled with one global atomic, so that a wave is not
// held up by its slowest ray.
// Waves per SIMD: six for the closest-hit and any-hit walks (the walk is
// latency bound, so the extra waves hide the node and triangle loads; seven
// is slower), five for the counting variant, whose counters would spill at six.
// MIXED: one launch for two batches, closest hit for the n rays of `rays`
// (hits into `hits`) and any hit for the n2 rays of `rays2` (flags into
// `occluded`), so that a path tracer's shadow rays of one step travel with
// the next step's rays instead of in a launch of their own.
template <bool STATS, bool ANY, bool MIXED = false>
__global__ void __launch_bounds__(TRACE_BLOCK)
__attribute__((amdgpu_waves_per_eu(STATS ? 5 : 6, STATS ? 5 : 6))) k_trace(TraceArgs a) {
    __shared__ uint32_t stk[STACK_LDS * TRACE_BLOCK];
    const uint32_t tid = threadIdx.x;
    if (blockIdx.x == 0) {
        // the counters of the next launch (they last served the launch before
        // this one, which has finished), and the caller's words
        for (uint32_t i = tid; i < XCD_COUNTER_STRIDE * (1u + XCDS); i += TRACE_BLOCK) a.work_next[i] = 0u;
        if (tid == 0 && a.zero0) *a.zero0 = 0u;
        if (tid == 0 && a.zero1) *a.zero1 = 0u;
    }
    const uint32_t gtid = blockIdx.x * TRACE_BLOCK + tid;
    const uint32_t lane = tid & 63u;
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    uint64_t c_nodes = 0, c_tris = 0, c_hits = 0, c_rays = 0;
    const uint32_t n1 = a.n_dev ? *a.n_dev : a.n;  // closest-hit rays (any-hit ones in an any-hit launch)
    const uint32_t n = MIXED ? n1 + *a.n2_dev : n1;  // rays in all
    if (n <= a.coop_max) return;                      // the cooperative walk takes this launch
    const bool unified = n <= a.unified_max;          // launch-uniform
    // large launches refill later: rays that start together stay in step
    // longer, and their waves more often take the unified loads
    const uint32_t refill_min = unified ? a.refill_min : a.refill_min_large;
    bool coherent = false;                            // wave-uniform: every ray of the wave in one octant
    uint32_t pool_next = 0, pool_end = 0;  // wave-uniform
    bool exhausted = false;                // wave-uniform
    uint32_t part = blockIdx.x % XCDS;     // wave-uniform: the XCD whose slices this wave takes
    uint32_t drained = 0;                  // wave-uniform: XCD counters seen past the end
    bool have = false, retired = false;
    uint32_t ri = 0;
    Lane L;
    while (true) {
        const uint64_t need = __ballot(!have && !retired);
        // refill once enough lanes are idle (or none is busy), so lanes that
        // start together stay in step
        if (need && ((uint32_t)__popcll(need) >= refill_min || __ballot(have) == 0ull)) {
            if (pool_next >= pool_end && !exhausted) {
                const int leader = __ffsll((unsigned long long)need) - 1;
                while (true) {
                    uint32_t v = 0;
                    if ((int)lane == leader) v = atomicAdd(a.work + XCD_COUNTER_STRIDE * (1u + part), POOL_CHUNK);
                    v = __shfl(v, leader);
                    // XCD part's v-th ray: ray v mod slice of its (v / slice)-th slice;
                    // increasing in v, so the first v past n ends this counter
                    const uint64_t g = (uint64_t)(v / a.slice) * (XCDS * (uint64_t)a.slice) +
                                       (uint64_t)part * a.slice + v % a.slice;
                    if (g < n) {
                        pool_next = (uint32_t)g;
                        pool_end = min(pool_next + POOL_CHUNK, n);
                        break;
                    }
                    if (++drained == XCDS) {
                        exhausted = true;
                        break;
                    }
                    part = (part + 1u) % XCDS;
                }
            }
            if (pool_next < pool_end) {
                if (!have && !retired) {
                    const uint32_t idx = pool_next + (uint32_t)__popcll(need & lt);
                    if (idx < pool_end) {
                        ri = idx;
                        have = true;
                        bool ok;
                        if (ANY && a.segs) {
                            // Scene::isOccluded: direction = end - start, dist2, normalize
                            const float* sg = a.segs + 6 * (size_t)ri;
                            f3 dir = mk(sg[3] - sg[0], sg[4] - sg[1], sg[5] - sg[2]);
                            const float d2 = dot(dir, dir);
                            dir = normalize_ref(dir);
                            rt_ray r;
                            r.ox = sg[0]; r.oy = sg[1]; r.oz = sg[2]; r.tmin = 0.0f;
                            r.dx = dir.x; r.dy = dir.y; r.dz = dir.z; r.tmax = INFINITY;
                            ok = lane_start(L, r);
                            L.seg_d2 = d2;
                            // conservative: a hit within the segment has t <= |end - start| (1 + 2^-10) + sigma
                            L.tcull = fmaf(sqrtf(d2), 1.0009765625f, L.sigma);
                        } else if (MIXED) {
                            L.any = ri >= n1;
                            ok = lane_start(L, L.any ? a.rays2[ri - n1] : a.rays[ri]);
                        } else {
                            ok = lane_start(L, a.rays[ri]);
                        }
                        if (!ok) {
                            if (ANY || (MIXED && L.any)) {
                                a.occluded[MIXED ? ri - n1 : ri] = 0;
                            } else {
                                rt_hit h;
                                h.tri = RT_MISS; h.t = 0.0f; h.u = 0.0f; h.v = 0.0f;
                                a.hits[ri] = h;
                            }
                            have = false;
                            if (STATS) ++c_rays;
                        }
                    }
                }
                pool_next = min(pool_end, pool_next + (uint32_t)__popcll(need));
                if (!unified) {
                    const uint64_t act = __ballot(have), bx = __ballot(have && L.negx),
                                   by = __ballot(have && L.negy), bz = __ballot(have && L.negz);
                    coherent = act != 0ull && (bx == 0ull || bx == act) && (by == 0ull || by == act) &&
                               (bz == 0ull || bz == act);
                }
            } else if (exhausted && !have) {
                retired = true;
            }
        }
        if (__ballot(!retired) == 0ull) break;
        if (have && lane_step<STATS, ANY, MIXED>(a, L, stk, tid, gtid, c_nodes, c_tris, unified || coherent)) {
            const bool hit = L.best_id != RT_MISS;
            if (ANY || (MIXED && L.any)) {
                a.occluded[MIXED ? ri - n1 : ri] = hit ? 1 : 0;
            } else {
                rt_hit h;
                h.tri = L.best_id;
                h.t = hit ? L.best_t : 0.0f;
                h.u = hit ? L.best_u : 0.0f;
                h.v = hit ? L.best_v : 0.0f;
                a.hits[ri] = h;
            }
            have = false;
            if (STATS) { ++c_rays; c_hits += hit; }
        }
    }
    if (STATS) {
        atomicAdd(a.stats + 0, (unsigned long long)c_rays);
        atomicAdd(a.stats + 1, (unsigned long long)c_nodes);
        atomicAdd(a.stats + 2, (unsigned long long)c_tris);
        atomicAdd(a.stats + 3, (unsigned long long)c_hits);
    }
}

// Few rays: a group of 2^g lanes (g <= coop_gmax, as large as the grid
// allows for n rays) walks each ray.  Every lane of a group walks the nodes
// above the split (index < coop_split[g]: the top levels); each child of
// such a node that is not itself above the split is dealt to one lane by a
// hash of its reference, and that lane walks its whole subtree.  So every
// subtree below the split is walked by exactly one lane, and the union of
// the group's walks covers the tree as one walk does.  After every step the
// lanes of a group take the smallest culling bound among them (a box is
// still skipped only when it cannot hold a hit closer than one already
// found), and at the end the group's hit is the closer() minimum of its
// lanes' hits: the same answer as one lane walking alone.  A small launch
// lasts as long as its longest walk (a few hundred dependent steps for rays
// skimming a terrain), which a group cuts by up to its width.  Any hit: a hit
// found by one lane ends the group's walk.
template <bool ANY, bool MIXED>
__global__ void __launch_bounds__(TRACE_BLOCK) __attribute__((amdgpu_waves_per_eu(6, 6))) k_trace_coop(TraceArgs a) {
    __shared__ uint32_t stk[STACK_LDS * TRACE_BLOCK];
    const uint32_t tid = threadIdx.x;
    if (a.coop_zero && blockIdx.x == 0 && tid == 0) {
        if (a.zero0) *a.zero0 = 0u;
        if (a.zero1) *a.zero1 = 0u;
    }
    const uint32_t n1 = a.n_dev ? *a.n_dev : a.n;
    const uint32_t n = MIXED ? n1 + *a.n2_dev : n1;
    if (n == 0 || n > a.coop_max) return;  // empty, or the plain walk takes this launch
    const uint32_t lanes = gridDim.x * TRACE_BLOCK;
    uint32_t g = 0;  // launch-uniform: the widest group with n groups in the grid
    while (g < a.coop_gmax && (lanes >> (g + 1u)) >= n) ++g;
    const uint32_t gtid = blockIdx.x * TRACE_BLOCK + tid;
    const uint32_t ri = gtid >> g, sub = gtid & ((1u << g) - 1u);
    if (ri >= n) return;  // whole groups: a group's lanes are consecutive in one wave
    const uint32_t split = a.coop_split[g];
    const uint32_t lane = tid & 63u;
    // this lane's group in the wave's ballot masks
    const uint64_t gmask = (g == 6u ? ~0ull : ((1ull << (1u << g)) - 1ull)) << (lane & ~((1u << g) - 1u));
    Lane L;
    bool ok;
    if (MIXED) {
        L.any = ri >= n1;
        ok = lane_start(L, L.any ? a.rays2[ri - n1] : a.rays[ri]);
    } else {
        ok = lane_start(L, a.rays[ri]);
    }
    const bool any_ray = ANY || (MIXED && L.any);  // group-uniform
    bool done = !ok;
    uint64_t c_nodes = 0, c_tris = 0;
    while (true) {
        if (!done) done = lane_step<false, ANY, MIXED, true>(a, L, stk, tid, gtid, c_nodes, c_tris, true, split, g, sub);
        if (ANY || MIXED) {
            if (__ballot(any_ray && L.best_id != RT_MISS) & gmask) done = true;
        }
        if (__ballot(!done) == 0ull) break;
        // the group's culling bound: the smallest of its lanes'
        float tc = L.tcull;
        for (uint32_t s = 1; s < (1u << g); s <<= 1) tc = fminf(tc, __shfl_xor(tc, (int)s));
        L.tcull = tc;
    }
    if (any_ray) {
        // (the ballot outside the sub == 0 test: it sees only the lanes that run it)
        const bool hit = (__ballot(L.best_id != RT_MISS) & gmask) != 0ull;
        if (sub == 0) a.occluded[MIXED ? ri - n1 : ri] = hit ? 1 : 0;
        return;
    }
    float d2 = L.best_d2, t = L.best_t, u = L.best_u, v = L.best_v;
    uint32_t id = L.best_id;
    for (uint32_t s = 1; s < (1u << g); s <<= 1) {
        const float d2o = __shfl_xor(d2, (int)s), to = __shfl_xor(t, (int)s), uo = __shfl_xor(u, (int)s),
                    vo = __shfl_xor(v, (int)s);
        const uint32_t ido = __shfl_xor(id, (int)s);
        if (closer(d2o, ido, d2, id)) {
            d2 = d2o; id = ido; t = to; u = uo; v = vo;
        }
    }
    if (sub == 0) {
        const bool hit = id != RT_MISS;
        rt_hit h;
        h.tri = id;
        h.t = hit ? t : 0.0f;
        h.u = hit ? u : 0.0f;
        h.v = hit ? v : 0.0f;
        a.hits[ri] = h;
    }
}

}  // namespace

namespace {
constexpr uint32_t WORK_HALF = XCD_COUNTER_STRIDE * (1 + XCDS);  // words per half of `work`

// Launches up to this many rays take the unified loads: four rays per
// resident lane (measured on the 1M-triangle soup: unified loads are
// 2-13 % faster up to 2^20 incoherent rays, 7-9 % slower at 2^22-2^23, and
// never slower for coherent rays); MTRT_TRACE_UNIFIED_MAX overrides it.
uint32_t unified_max(const rt_ctx* c, uint32_t threads) {
    return c->trace_unified_max ? c->trace_unified_max : 4u * threads;
}
constexpr uint32_t WORK_ERR = 2 * WORK_HALF;                      // the sticky error word

// d_rays2 set: a mixed launch, closest hit for d_rays (count *n_dev) and any
// hit for d_rays2 (count *n2_dev, flags into d_occl); n is then the two
// batches' capacity together.
int launch(rt_ctx* c, const rt_ray* d_rays, uint32_t n, rt_hit* d_hits, uint8_t* d_occl, bool stats,
           const uint32_t* n_dev, const float* d_segs = nullptr, uint32_t* zero0 = nullptr,
           uint32_t* zero1 = nullptr, const rt_ray* d_rays2 = nullptr, const uint32_t* n2_dev = nullptr) {
    // the kernel's 32-bit ray chunks (pool_next + POOL_CHUNK) stay below 2^32
    if (n > RT_MAX_RAYS) return fail(c, RT_E_UNSUPPORTED, "trace: more than RT_MAX_RAYS rays in one launch");
    if (n == 0) {
        if (zero0) RT_HIP(c, hipMemsetAsync(zero0, 0, 4, c->stream));
        if (zero1) RT_HIP(c, hipMemsetAsync(zero1, 0, 4, c->stream));
        return RT_OK;
    }
    const bool mixed = d_rays2 != nullptr;
    const bool any = d_occl != nullptr && !mixed;
    // resident workgroups per CU of each variant, asked once per context
    int& per_cu = c->trace_per_cu[mixed ? 3 : (any ? 2 : (stats ? 1 : 0))];
    if (per_cu == 0) {
        if (mixed)
            RT_HIP(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_trace<false, false, true>, TRACE_BLOCK, 0));
        else if (any) RT_HIP(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_trace<false, true>, TRACE_BLOCK, 0));
        else if (stats) RT_HIP(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_trace<true, false>, TRACE_BLOCK, 0));
        else RT_HIP(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_trace<false, false>, TRACE_BLOCK, 0));
        if (per_cu < 1) per_cu = 1;
    }
    // cooperative walks for small launches (not for the counting variant or
    // segment queries): the grid then leaves room for 64 lanes per ray
    const bool coop = c->trace_coop > 0 && c->coop_gmax > 0 && !stats && !d_segs;
    const uint64_t want = coop ? 64ull * n : (uint64_t)n;
    uint32_t grid = (uint32_t)(c->num_cus * per_cu);
    const uint64_t need_blocks = (want + TRACE_BLOCK - 1) / TRACE_BLOCK;
    if (grid > need_blocks) grid = (uint32_t)need_blocks;
    const uint32_t threads = grid * TRACE_BLOCK;
    // launches of at most coop_max rays have at least 2^trace_coop lanes per ray
    const uint32_t coop_max = coop ? threads >> c->trace_coop : 0u;
    const uint32_t max_stack = 3u * c->bvh_depth + 1u;
    const uint32_t spill_cap = max_stack > (uint32_t)STACK_LDS ? max_stack - STACK_LDS : 1u;
    int rc;
    if ((rc = ensure(c, c->spill, 4ull * spill_cap * threads))) return rc;
    // two halves of ray counters, each zeroed by the launch before the one
    // that uses it, and the error word: cleared once, at allocation
    if (!c->work.ptr) {
        const size_t work_bytes = 4ull * (WORK_ERR + 32);
        if ((rc = ensure(c, c->work, work_bytes))) return rc;
        RT_HIP(c, hipMemsetAsync(c->work.ptr, 0, work_bytes, c->stream));
        c->work_half = 0;
    }
    if ((rc = ensure(c, c->stats, 64))) return rc;
    if (stats) RT_HIP(c, hipMemsetAsync(c->stats.ptr, 0, 32, c->stream));
    TraceArgs a;
    a.rays = d_rays;
    a.hits = d_hits;
    a.occluded = d_occl;
    a.segs = d_segs;
    a.n = n;
    a.n_dev = n_dev;
    a.rays2 = d_rays2;
    a.n2_dev = n2_dev;
    a.nodes = (const rt_bvh4q_node*)c->qnodes.ptr;
    a.tris = (const rt_tri_rec*)c->tris.ptr;
    a.work = (uint32_t*)c->work.ptr + WORK_HALF * c->work_half;
    a.work_next = (uint32_t*)c->work.ptr + WORK_HALF * (c->work_half ^ 1u);
    a.err = (uint32_t*)c->work.ptr + WORK_ERR;
    a.zero0 = zero0;
    a.zero1 = zero1;
    a.spill = (uint32_t*)c->spill.ptr;
    a.spill_cap = spill_cap;
    a.spill_stride = threads;
    a.max_steps = c->trace_max_steps ? c->trace_max_steps : 2u * c->num_nodes + 2u * c->T + 16u;
    a.refill_min = c->trace_refill_min;
    a.refill_min_large = c->trace_refill_large;
    a.unified_max = unified_max(c, threads);
    a.slice = c->trace_slice;
    a.stats = (unsigned long long*)c->stats.ptr;
    a.coop_max = coop_max;
    a.coop_gmax = c->coop_gmax;
    for (int g = 0; g < 7; ++g) a.coop_split[g] = c->coop_split[g];
    a.coop_zero = 0;
    // A count on the device: both kernels, each returning at once when the
    // launch is the other's (the plain one first: it also zeroes the next
    // launch's counters).  A host count: the one that takes it.
    const bool run_plain = !coop || n_dev || n > coop_max;
    const bool run_coop = coop && (n_dev || n <= coop_max);
    RT_HIP(c, hipEventRecord(c->ev[2], c->stream));
    if (run_plain) {
        if (mixed) hipLaunchKernelGGL((k_trace<false, false, true>), dim3(grid), dim3(TRACE_BLOCK), 0, c->stream, a);
        else if (any) hipLaunchKernelGGL((k_trace<false, true>), dim3(grid), dim3(TRACE_BLOCK), 0, c->stream, a);
        else if (stats) hipLaunchKernelGGL((k_trace<true, false>), dim3(grid), dim3(TRACE_BLOCK), 0, c->stream, a);
        else hipLaunchKernelGGL((k_trace<false, false>), dim3(grid), dim3(TRACE_BLOCK), 0, c->stream, a);
        RT_HIP(c, hipGetLastError());
        c->work_half ^= 1u;  // (a launch of the cooperative walk alone leaves the counters untouched)
    }
    if (run_coop) {
        a.coop_zero = run_plain ? 0u : 1u;
        if (mixed) hipLaunchKernelGGL((k_trace_coop<false, true>), dim3(grid), dim3(TRACE_BLOCK), 0, c->stream, a);
        else if (any) hipLaunchKernelGGL((k_trace_coop<true, false>), dim3(grid), dim3(TRACE_BLOCK), 0, c->stream, a);
        else hipLaunchKernelGGL((k_trace_coop<false, false>), dim3(grid), dim3(TRACE_BLOCK), 0, c->stream, a);
        RT_HIP(c, hipGetLastError());
    }
    RT_HIP(c, hipEventRecord(c->ev[3], c->stream));
    c->timing.launches = 1;
    return RT_OK;
}
}  // namespace

int launch_trace(rt_ctx* c, const rt_ray* d_rays, uint32_t n, rt_hit* d_hits, bool stats, const uint32_t* n_dev,
                 uint32_t* zero0, uint32_t* zero1) {
    return launch(c, d_rays, n, d_hits, nullptr, stats, n_dev, nullptr, zero0, zero1);
}
int launch_trace_any(rt_ctx* c, const rt_ray* d_rays, uint32_t n, uint8_t* d_occluded, const uint32_t* n_dev,
                     uint32_t* zero0) {
    return launch(c, d_rays, n, nullptr, d_occluded, false, n_dev, nullptr, zero0);
}
int launch_trace_mixed(rt_ctx* c, const rt_ray* d_rays, const uint32_t* n_dev, rt_hit* d_hits, const rt_ray* d_rays2,
                       const uint32_t* n2_dev, uint8_t* d_occluded, uint32_t capacity, uint32_t* zero0,
                       uint32_t* zero1) {
    return launch(c, d_rays, capacity, d_hits, d_occluded, false, n_dev, nullptr, zero0, zero1, d_rays2, n2_dev);
}
void set_coop_splits(rt_ctx* c, const std::vector<uint32_t>& level_start) {
    // group width 2^g: split at the first level (below the root) holding at
    // least trace_coop_nodes nodes per lane of the group
    c->coop_gmax = 0;
    for (int g = 0; g < 7; ++g) c->coop_split[g] = 0;
    const size_t levels = level_start.size() - 1;  // the last entry is the node count
    for (uint32_t g = 1; g <= c->trace_coop_gmax; ++g) {
        const uint64_t need = (uint64_t)c->trace_coop_nodes << g;
        size_t l = 1;
        while (l < levels && level_start[l + 1] - level_start[l] < need) ++l;
        if (l >= levels) break;
        c->coop_split[g] = level_start[l];
        c->coop_gmax = g;
    }
}
int trace_error(rt_ctx* c, uint32_t* err) {
    *err = 0;
    if (!c->work.ptr) return RT_OK;
    uint32_t* w = (uint32_t*)c->work.ptr + WORK_ERR;
    RT_HIP(c, hipMemcpyAsync(err, w, 4, hipMemcpyDeviceToHost, c->stream));
    RT_HIP(c, hipStreamSynchronize(c->stream));
    if (*err) RT_HIP(c, hipMemsetAsync(w, 0, 4, c->stream));
    return RT_OK;
}
int launch_occluded(rt_ctx* c, const float* d_segs, uint32_t n, uint8_t* d_occluded) {
    return launch(c, nullptr, n, nullptr, d_occluded, false, nullptr, d_segs);
}

// ---- allocation-free closest hit (rt_trace_closest_noalloc) ----------------
// Its scratch lives in a file-scope device array, one copy per GPU, kept
// until the process ends: NOALLOC_HEADER_WORDS of ray counters (two halves as
// in `work`) and the error word, then the per-lane spill stacks.  Host code
// never passes `scratch` itself (in host code the name is a stand-in, not a
// device address) but the address hipGetSymbolAddress returns for the current
// device, kept per device.
static __device__ __align__(256) uint32_t scratch[RT_NOALLOC_SCRATCH_BYTES / 4];

namespace {
constexpr uint32_t NOALLOC_HEADER_WORDS = 1024;
static_assert(2 * WORK_HALF + 1 <= NOALLOC_HEADER_WORDS, "counter header too small");

std::mutex g_scratch_mu;
uint32_t* g_scratch_dev[256] = {};  // per device ordinal

__global__ void k_zero_words(uint32_t* p, uint32_t n, uint32_t* status) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0u;
    if (status && blockIdx.x == 0 && threadIdx.x == 0) *status = 0u;
}
}  // namespace

int noalloc_scratch_init(rt_ctx* c) {
    if (c->device < 0 || c->device >= 256) return fail(c, RT_E_UNSUPPORTED, "device ordinal beyond 255");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[c->device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_fail(c, e, "hipGetSymbolAddress(scratch)");
        if (!p) return fail(c, RT_E_DEVICE, "hipGetSymbolAddress(scratch) returned NULL");
        g_scratch_dev[c->device] = (uint32_t*)p;
    }
    c->noalloc_scratch = g_scratch_dev[c->device];
    return RT_OK;
}

int trace_noalloc(rt_ctx* c, const rt_ray* d_rays, uint32_t n, rt_hit* d_hits, uint32_t* d_status) {
    uint32_t* base = c->noalloc_scratch;
    if (!base) return fail(c, RT_E_STATE, "rt_trace_closest_noalloc: no scratch address for this device");
    if (n > RT_MAX_RAYS) return fail(c, RT_E_UNSUPPORTED, "rt_trace_closest_noalloc: more than RT_MAX_RAYS rays");
    int& per_cu = c->trace_per_cu[0];
    if (per_cu == 0) {
        RT_HIP(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_trace<false, false>, TRACE_BLOCK, 0));
        if (per_cu < 1) per_cu = 1;
    }
    const uint32_t max_stack = 3u * c->bvh_depth + 1u;
    const uint32_t spill_cap = max_stack > (uint32_t)STACK_LDS ? max_stack - STACK_LDS : 1u;
    const size_t block_bytes = 4ull * spill_cap * TRACE_BLOCK;
    const size_t spill_bytes = (size_t)RT_NOALLOC_SCRATCH_BYTES - 4ull * NOALLOC_HEADER_WORDS;
    const size_t fit = spill_bytes / block_bytes;
    if (fit == 0)
        return fail(c, RT_E_UNSUPPORTED, "rt_trace_closest_noalloc: a BVH of depth " + std::to_string(c->bvh_depth) +
                                             " needs " + std::to_string(block_bytes) +
                                             " bytes of traversal stack per workgroup, beyond the " +
                                             std::to_string(spill_bytes) + "-byte scratch");
    uint32_t grid = (uint32_t)(c->num_cus * per_cu);
    const uint32_t need_blocks = (n + TRACE_BLOCK - 1) / TRACE_BLOCK;
    if (grid > need_blocks) grid = need_blocks;
    if (grid > fit) grid = (uint32_t)fit;
    // counters and the error word start at zero: a fill kernel, not a
    // memset, so that the call can be captured into a graph
    hipLaunchKernelGGL(k_zero_words, dim3(4), dim3(256), 0, c->stream, base, NOALLOC_HEADER_WORDS, d_status);
    RT_HIP(c, hipGetLastError());
    TraceArgs a;
    a.rays = d_rays;
    a.hits = d_hits;
    a.occluded = nullptr;
    a.segs = nullptr;
    a.n = n;
    a.n_dev = nullptr;
    a.rays2 = nullptr;
    a.n2_dev = nullptr;
    a.coop_max = 0;  // the plain walk only
    a.coop_gmax = 0;
    for (int g = 0; g < 7; ++g) a.coop_split[g] = 0;
    a.coop_zero = 0;
    a.nodes = (const rt_bvh4q_node*)c->qnodes.ptr;
    a.tris = (const rt_tri_rec*)c->tris.ptr;
    a.work = base;
    a.work_next = base + WORK_HALF;
    a.err = d_status ? d_status : base + 2 * WORK_HALF;
    a.zero0 = nullptr;
    a.zero1 = nullptr;
    a.spill = base + NOALLOC_HEADER_WORDS;
    a.spill_cap = spill_cap;
    a.spill_stride = grid * TRACE_BLOCK;
    a.max_steps = c->trace_max_steps ? c->trace_max_steps : 2u * c->num_nodes + 2u * c->T + 16u;
    a.refill_min = c->trace_refill_min;
    a.refill_min_large = c->trace_refill_large;
    a.unified_max = unified_max(c, grid * TRACE_BLOCK);
    a.slice = c->trace_slice;
    a.stats = nullptr;
    hipLaunchKernelGGL((k_trace<false, false>), dim3(grid), dim3(TRACE_BLOCK), 0, c->stream, a);
    RT_HIP(c, hipGetLastError());
    c->timing.launches = 2;
    return RT_OK;
}

}  // namespace rt
