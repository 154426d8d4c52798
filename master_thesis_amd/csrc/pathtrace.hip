// pathtrace.hip — wavefront diffuse path tracer with persistent path slots.
//
// Replaces GraphicsEngine::render's thread pool (software_renderer/graphic/
// graphics_engine.cpp:42-97) and PathTracer::renderPixel / traceSinglePath
// (graphic/path_tracer.cpp:18-97).
//
// Estimator (the reference's, restated): the primary ray of a pixel is fixed
// (no jitter) and shared by its N samples.  Per sample, throughput T = 1 and
// for depth i < max_depth: a miss ends the path; while backface culling is on
// and n.d > 0 the ray steps past the surface (o = p + 0.01 d) and the depth
// slot is consumed; otherwise culling switches off, xi0 picks diffuse /
// reflect / refract by the material thresholds, o = p + 0.01 d', a diffuse
// bounce multiplies T by colour * (n.d'), and a light hit from the front
// (n.d_prev <= 0) ends the path with contribution T.  The pixel is
// 2*pi/N * sum of contributions, summed in sample order, the scale applied
// in double.
//
// Scheduling: a slot owns one work item (pixel, chunk of samples) and runs
// its samples one after another, so each pixel's sum has the reference's
// order.  A pixel's N samples are split into K chunks (K <= N, chunk c holds
// samples [c N / K, (c + 1) N / K)) when there are fewer pixels than slots.  Every iteration is two kernels over the live slots:
//   extend  closest hit of every queued ray (trace.hip, device-resident count)
//   shade   the state machine above; a finished sample restarts the slot on
//           the next sample; a finished item takes the next item from a
//           global counter (the analogue of pixelCounter.fetch_add,
//           graphics_engine.cpp:74); surviving slots are compacted into the
//           next queue with one atomic per wave (ballot + popcount).
// The primary ray of a pixel is fixed and the culling steps after it draw no
// random numbers, so only sample 0 of an item traces that prefix: the slot
// caches its first accepted vertex (or the escape, or the depth running out)
// and later samples of the item start there (plain and adjoint modes).
// Random numbers: Philox-4x32-10 with counter (pixel, sample, depth, view)
// and the 64-bit seed as key, so a path does not depend on scheduling and the
// adjoint pass replays it exactly.  Several views (rt_render_views) share
// one slot pool: a work item is (chunk, view, pixel).
//
// Radiance cache (rt_render_nrc): at its query vertex (the nrc_q-th
// front-facing surface hit, unless that hit ends the path on a light) a
// sample either stops and enqueues (position, normal, direction, throughput,
// pixel) for the cache, or, if it is a training sample, records the features
// there, keeps a second throughput Tq from that vertex on, and at its end
// enqueues a training record with target Tq * (end term).  Every 8 iterations
// the host runs the cache over the queued queries, which add throughput *
// prediction to their pixels' partial sums atomically; after the frame the
// cache trains on the records.
//
// Adjoint (rt_render_adjoint): the same pass with dL/dI given.  A path's
// contribution is the product of its diffuse factors colour_k * cos_k (the
// light's own colour included), so d C / d colour_m sums, over the vertices k
// with material m, the product of all other factors times cos_k.  Paths do not
// depend on the colours (uniform hemisphere sampling, colour-independent
// thresholds), which makes this gradient exact for the rendered estimate.
#include <math.h>
#include <string.h>

#include "rt_internal.hpp"
#include "rt_core.hpp"

namespace rt {
namespace {

constexpr int PT_BLOCK = 256;
constexpr uint32_t ADJ_MAX_DEPTH = 16;
constexpr uint32_t ADJ_LDS_MAX = 3072;  // 3 * materials summed in LDS (12 KB)
constexpr uint32_t CULL_BIT = 0x10000u;
// cache mode: front-facing hits so far (bits 17-23), throughput-from-the-
// query-vertex tracked (bit 24), last bounce diffuse (bit 25)
constexpr uint32_t FH_SHIFT = 17;
constexpr uint32_t FH_MASK = 0x7Fu;
constexpr uint32_t REC_BIT = 1u << 24;
constexpr uint32_t DIFF_BIT = 1u << 25;
constexpr uint32_t NO_REC = 0xFFFFFFFFu;

// ctr slots
constexpr int C_ITEM = 0;   // next work item
constexpr int C_N0 = 1;     // live count, queue 0
constexpr int C_N1 = 2;     // live count, queue 1
constexpr int C_RAYS = 4;   // 64-bit count of traced rays (slots 4-5)
constexpr int C_QN = 6;     // cache queries queued since the last flush
constexpr int C_TN = 7;     // training records made
constexpr int C_QDROP = 8;  // queries dropped for want of room (stays 0: the host flushes in time)
constexpr int C_SN0 = 9;    // shadow rays queued by an iteration that read queue 0
constexpr int C_SN1 = 3;    // ... queue 1 (their trace runs with the next iteration's rays)
__host__ __device__ constexpr int sn_slot(int par) { return par ? C_SN1 : C_SN0; }
constexpr int C_SHRAYS = 10;  // 64-bit count of shadow rays (slots 10-11)
constexpr int C_QTOT = 12;    // 64-bit count of cache queries answered (slots 12-13)
constexpr int C_WORDS = 14;

struct PTArgs {
    const ShadeGeom* geom;
    const ShadeNrm* nrm;
    const rt_material* mat;
    const rt_camera* cams;  // V cameras in device memory
    uint32_t W, H, P;       // image size, pixels per view
    uint32_t V, VP;         // views, V * P
    uint32_t K, spp, items, S;  // chunks per pixel, samples per pixel, work items, slots
    uint32_t max_depth, sample_offset, flags;
    uint32_t seed_lo, seed_hi;
    float sky[3];
    float sun_dir[3], sun_cos, sun_rad[3];
    double scale;
    // slot state
    uint32_t* s_item;
    uint32_t* s_smp;
    uint32_t* s_state;   // depth | CULL_BIT
    float* s_T;          // 3 * S
    float* s_sum;        // 3 * S
    uint32_t* s_amat;    // adjoint: ADJ_MAX_DEPTH * S material ids
    float* s_acos;       // adjoint: ADJ_MAX_DEPTH * S cosines
    uint32_t* s_nv;      // adjoint: diffuse vertices so far
    float* s_fv;         // plain and adjoint: 11 * S first-vertex cache of the slot's item
    // queues
    uint32_t* q_slot[2];
    rt_ray* q_ray[2];
    rt_hit* hits;
    uint32_t* ctr;
    float* partial;      // 3 * VP * K
    // outputs
    float* rgb;
    uint8_t* rgb8;
    const float* dimg;
    float* dcolor;
    // adjoint with few materials: each workgroup sums its gradient in LDS and
    // adds it to its own column of adj_rows ([adj_cols][grid]); 0 = atomics
    // straight into dcolor
    float* adj_rows;
    uint32_t adj_cols;
    // radiance cache (nrc_q == 0: off)
    uint32_t nrc_q, nrc_every, nrc_query;
    float* q_feat;      // q_cap * 9
    float* q_w;         // q_cap * 3: throughput before the query vertex
    uint32_t* q_idx;    // q_cap: partial-sum index
    uint32_t q_cap;
    float* t_feat;      // t_cap * 9
    float* t_tgt;       // t_cap * 3
    uint32_t* t_key;    // t_cap: the record's sample, (view * P + pixel) * key_n + sample - sample_offset
    uint32_t t_cap, key_n;
    uint32_t* s_ridx;   // S: training record of the slot's sample (NO_REC: none)
    float* s_Tq;        // 3 * S: its throughput from the query vertex on
    // light sampling (cache mode): shadow rays to emissive triangles and the sun
    uint32_t nee, sun_on;
    const LightTri* lights;
    uint32_t num_lights;
    float light_area;
    rt_ray* sh_ray;     // S
    float* sh_w;        // 6 * S: pixel weight, record weight
    uint32_t* sh_idx;   // 2 * S: partial-sum index, record index
    uint8_t* sh_occ;    // S
    // probe field in place of the network (rt_render_ddgi): queries only at
    // purely diffuse vertices, which also take their light sample
    uint32_t ddgi_mode;
};

// Training-sample choice: lowbias32 (C. Wellons) over pixel, sample, view, seed.
__device__ __forceinline__ uint32_t train_hash(uint32_t pix, uint32_t smp, uint32_t view, uint32_t seed) {
    uint32_t h = pix * 0x9E3779B1u ^ (smp + 0x7F4A7C15u) * 0x85EBCA77u ^ (view + 0x632BE5ABu) * 0xC2B2AE3Du ^ seed;
    h ^= h >> 16;
    h *= 0x7FEB352Du;
    h ^= h >> 15;
    h *= 0x846CA68Bu;
    h ^= h >> 16;
    return h;
}

__device__ __forceinline__ rt_ray make_ray(f3 o, f3 d) {
    rt_ray r;
    r.ox = o.x; r.oy = o.y; r.oz = o.z; r.tmin = 0.0f;
    r.dx = d.x; r.dy = d.y; r.dz = d.z; r.tmax = INFINITY;
    return r;
}

// vp = view * P + pixel
// first sample of chunk c of a pixel (chunk c holds the samples up to the next chunk's first)
__device__ __forceinline__ uint32_t chunk_first(const PTArgs& a, uint32_t c) {
    return (uint32_t)(((uint64_t)c * a.spp) / a.K);
}

__device__ __forceinline__ void start_sample(const PTArgs& a, uint32_t slot, uint32_t vp, rt_ray& r, bool adj) {
    const rt_camera& c = a.cams[vp / a.P];
    const uint32_t pix = vp % a.P;
    f3 d = primary_dir(c.view_inv, c.proj_inv, pix % a.W, pix / a.W, a.W, a.H);
    r = make_ray(mk(c.origin[0], c.origin[1], c.origin[2]), d);
    a.s_state[slot] = CULL_BIT;
    a.s_T[3 * slot] = 1.0f; a.s_T[3 * slot + 1] = 1.0f; a.s_T[3 * slot + 2] = 1.0f;
    if (adj) a.s_nv[slot] = 0;
}

__device__ __forceinline__ void start_item(const PTArgs& a, uint32_t slot, uint32_t item, rt_ray& r, bool adj) {
    a.s_item[slot] = item;
    a.s_smp[slot] = 0;
    a.s_sum[3 * slot] = 0.0f; a.s_sum[3 * slot + 1] = 0.0f; a.s_sum[3 * slot + 2] = 0.0f;
    start_sample(a, slot, item % a.VP, r, adj);
}

// Radiance of an escaped ray (extension; the reference's miss is black):
// constant sky plus, when `sun` is set, the sun inside its cone.
__device__ __forceinline__ f3 miss_radiance(const PTArgs& a, f3 d, bool sun) {
    f3 L = mk(0.0f, 0.0f, 0.0f);
    if (a.flags & RT_RENDER_SKY) L = mk(a.sky[0], a.sky[1], a.sky[2]);
    if (sun && (a.flags & RT_RENDER_SUN) && dot(d, mk(a.sun_dir[0], a.sun_dir[1], a.sun_dir[2])) >= a.sun_cos)
        L = add(L, mk(a.sun_rad[0], a.sun_rad[1], a.sun_rad[2]));
    return L;
}

// Hit point, interpolated shading normal and material of a hit:
// Triangle::getBarycentricCoords (triangle.cpp:47-55) and Scene::traceRay's
// normal (scene.cpp:49-53).
__device__ __forceinline__ void surface(const PTArgs& a, const rt_hit& h, f3 o, f3 d, f3& p, f3& nrm, uint32_t& mat) {
    p = hit_point(o, d, h.t);
    const ShadeGeom g = a.geom[h.tri];
    const ShadeNrm nn = a.nrm[h.tri];
    const f3 v0 = mk(g.v0[0], g.v0[1], g.v0[2]);
    const f3 e0 = mk(g.e0[0], g.e0[1], g.e0[2]), e1 = mk(g.e1[0], g.e1[1], g.e1[2]);
    const f3 e2 = sub(p, v0);
    const float d20 = dot(e2, e0), d21 = dot(e2, e1);
    const float bv = (g.d11 * d20 - g.d01 * d21) / g.denom;
    const float bw = (g.d00 * d21 - g.d01 * d20) / g.denom;
    const float bu = 1.0f - bv - bw;
    nrm = mk(0.0f, 0.0f, 0.0f);
    nrm = add(nrm, scale(bu, mk(nn.n0[0], nn.n0[1], nn.n0[2])));
    nrm = add(nrm, scale(bv, mk(nn.n1[0], nn.n1[1], nn.n1[2])));
    nrm = add(nrm, scale(bw, mk(nn.n2[0], nn.n2[1], nn.n2[2])));
    nrm = normalize_ref(nrm);
    mat = g.mat;
}

template <bool ADJ>
__global__ void __launch_bounds__(PT_BLOCK) k_pt_init(PTArgs a) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        a.ctr[C_ITEM] = a.S;
        a.ctr[C_N0] = a.S;
        a.ctr[C_N1] = 0;
        *(unsigned long long*)&a.ctr[C_RAYS] = 0ull;
        *(unsigned long long*)&a.ctr[C_SHRAYS] = 0ull;
        *(unsigned long long*)&a.ctr[C_QTOT] = 0ull;
        a.ctr[C_SN0] = 0;
        a.ctr[C_SN1] = 0;
        a.ctr[C_QN] = 0;
        a.ctr[C_TN] = 0;
        a.ctr[C_QDROP] = 0;
    }
    if (i >= a.S) return;
    rt_ray r;
    start_item(a, i, i, r, ADJ);
    a.q_slot[0][i] = i;
    a.q_ray[0][i] = r;
}

// Scatter dL/dI * dC/dcolour over the diffuse vertices of a finished path.
// `term` is the factor after the last vertex (1 for a light hit, the sky
// radiance for an escaped path).
__device__ __forceinline__ void scatter_grad(const PTArgs& a, uint32_t slot, uint32_t pix, f3 term, float* acc) {
    const uint32_t nv = a.s_nv[slot];
    const float w[3] = {a.dimg[3 * pix] * (float)a.scale, a.dimg[3 * pix + 1] * (float)a.scale,
                        a.dimg[3 * pix + 2] * (float)a.scale};
    const float tm[3] = {term.x, term.y, term.z};
    uint32_t m[ADJ_MAX_DEPTH];
    float cs[ADJ_MAX_DEPTH];
    for (uint32_t k = 0; k < nv; ++k) {
        m[k] = a.s_amat[k * a.S + slot];
        cs[k] = a.s_acos[k * a.S + slot];
    }
    for (int ch = 0; ch < 3; ++ch) {
        if (w[ch] == 0.0f) continue;
        float f[ADJ_MAX_DEPTH];
        for (uint32_t k = 0; k < nv; ++k) f[k] = a.mat[m[k]].color[ch] * cs[k];
        // prefix products forward, suffix products backward
        float pre = 1.0f;
        float suf[ADJ_MAX_DEPTH + 1];
        suf[nv] = tm[ch];
        for (int k = (int)nv - 1; k >= 0; --k) suf[k] = f[k] * suf[k + 1];
        for (uint32_t k = 0; k < nv; ++k) {
            float g = pre * suf[k + 1] * cs[k];
            atomicAdd(a.adj_cols ? &acc[3 * m[k] + ch] : &a.dcolor[3 * m[k] + ch], w[ch] * g);
            pre = pre * f[k];
        }
    }
}

// First-vertex cache of a work item (plain and adjoint modes).  A pixel's
// primary ray is fixed and the culling steps before the first accepted hit
// draw no random numbers, so every sample of an item reaches the same first
// vertex along the same direction; only sample 0 traces that prefix and the
// slot keeps where it ended.  Later samples start at the cached vertex and
// draw their own numbers there, which gives the same paths, the same values
// and the same summation order as tracing the prefix again.
// s_fv layout, one array of S words each: 0-2 p, 3-5 shading normal,
// 6-8 direction, 9 material, 10 depth | kind << 16.
constexpr uint32_t FV_HIT = 1;   // accepted vertex at `depth`
constexpr uint32_t FV_MISS = 2;  // escaped before any accepted vertex
constexpr uint32_t FV_END = 3;   // depth ran out in culling steps

__device__ __forceinline__ void fv_put(const PTArgs& a, uint32_t slot, uint32_t kind, uint32_t depth, f3 p, f3 n,
                                       f3 d, uint32_t mid) {
    const size_t S = a.S;
    float* f = a.s_fv;
    if (kind == FV_HIT) {
        f[0 * S + slot] = p.x; f[1 * S + slot] = p.y; f[2 * S + slot] = p.z;
        f[3 * S + slot] = n.x; f[4 * S + slot] = n.y; f[5 * S + slot] = n.z;
        ((uint32_t*)f)[9 * S + slot] = mid;
    }
    f[6 * S + slot] = d.x; f[7 * S + slot] = d.y; f[8 * S + slot] = d.z;
    ((uint32_t*)f)[10 * S + slot] = depth | (kind << 16);
}

// The accepted-vertex step of the estimator at (p, nrm, material mid) reached
// along d: pick the branch, bounce, weight a diffuse bounce, end on a light
// seen from the front, then count the depth slot.
template <bool ADJ>
__device__ __forceinline__ void vertex_step(const PTArgs& a, uint32_t slot, uint32_t pix, uint32_t view,
                                            uint32_t gsmp, uint32_t& depth, f3 p, f3 nrm, uint32_t mid, f3& o,
                                            f3& d, f3& T, bool& end, bool& contrib) {
    const float ndotd = dot(nrm, d);
    const rt_material m = a.mat[mid];
    const bool light_end = m.light_source && ndotd <= 0.0f;
    u4 c4;
    c4.x = pix; c4.y = gsmp; c4.z = depth; c4.w = view;
    const u4 rnd = philox4x32_10(c4, a.seed_lo, a.seed_hi);
    const float xi0 = u01(rnd.x);
    const bool diffuse = xi0 <= m.diffuse_threshold;
    if (diffuse) d = hemisphere_dir(u01(rnd.y), u01(rnd.z), nrm);
    else if (xi0 <= m.reflect_threshold) d = reflect_ref(d, nrm);
    else if (xi0 <= m.transparent_threshold) d = custom_refract_ref(d, nrm, m.refraction_index);
    o = add(p, scale(0.01f, d));
    if (diffuse) {
        const float cs = dot(nrm, d);
        T = mulv(T, scale(cs, mk(m.color[0], m.color[1], m.color[2])));
        if (ADJ) {
            uint32_t nv = a.s_nv[slot];
            if (nv < ADJ_MAX_DEPTH) {
                a.s_amat[nv * a.S + slot] = mid;
                a.s_acos[nv * a.S + slot] = cs;
            }
            a.s_nv[slot] = nv + 1;
        }
    }
    if (light_end) {
        end = true;
        contrib = true;
    }
    ++depth;
    if (depth >= a.max_depth) end = true;
}

template <bool ADJ>
__global__ void __launch_bounds__(PT_BLOCK) k_pt_shade(PTArgs a, int cur) {
    __shared__ float adj_acc[ADJ ? ADJ_LDS_MAX : 1];
    if (ADJ && a.adj_cols) {
        for (uint32_t i = threadIdx.x; i < a.adj_cols; i += PT_BLOCK) adj_acc[i] = 0.0f;
        __syncthreads();
    }
    const uint32_t n = a.ctr[cur ? C_N1 : C_N0];
    uint32_t* n_out = &a.ctr[cur ? C_N0 : C_N1];
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k == 0) *(unsigned long long*)&a.ctr[C_RAYS] += n;  // rays the extend step just traced
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    bool alive = false, need_item = false;
    uint32_t slot = 0;
    rt_ray nr;
    if (k < n) {
        slot = a.q_slot[cur][k];
        const rt_ray r = a.q_ray[cur][k];
        const rt_hit h = a.hits[k];
        const uint32_t item = a.s_item[slot];
        const uint32_t vp = item % a.VP;   // view * P + pixel: index into images and partial sums
        const uint32_t chunk = item / a.VP;
        const uint32_t view = vp / a.P, pix = vp % a.P;
        const uint32_t c0 = chunk_first(a, chunk), cn = chunk_first(a, chunk + 1) - c0;
        uint32_t smp = a.s_smp[slot];
        uint32_t st = a.s_state[slot];
        uint32_t depth = st & 0xFFFFu;
        bool cull = (st & CULL_BIT) != 0;
        f3 T = mk(a.s_T[3 * slot], a.s_T[3 * slot + 1], a.s_T[3 * slot + 2]);
        f3 o = mk(r.ox, r.oy, r.oz), d = mk(r.dx, r.dy, r.dz);
        bool end = false, contrib = false, at_vertex = false;
        f3 term = mk(1.0f, 1.0f, 1.0f);
        f3 p = o, nrm = o;
        uint32_t mid = 0;
        if (h.tri == RT_MISS) {
            end = true;
            if (a.flags & (RT_RENDER_SKY | RT_RENDER_SUN)) {
                contrib = true;
                term = miss_radiance(a, d, true);
                T = mulv(T, term);
            }
            if (cull) fv_put(a, slot, FV_MISS, depth, o, o, d, 0);
        } else {
            surface(a, h, o, d, p, nrm, mid);
            if (cull && dot(nrm, d) > 0.0f) {
                o = add(p, scale(0.01f, d));
                ++depth;
                if (depth >= a.max_depth) {
                    end = true;
                    fv_put(a, slot, FV_END, depth, p, nrm, d, mid);
                }
            } else {
                if (cull) fv_put(a, slot, FV_HIT, depth, p, nrm, d, mid);
                cull = false;
                at_vertex = true;
            }
        }
        if (at_vertex)
            vertex_step<ADJ>(a, slot, pix, view, a.sample_offset + c0 + smp, depth, p, nrm, mid, o, d, T,
                             end, contrib);
        if (end) {
            // the sample ended: add it, then start the next one at the item's
            // cached first vertex (at most one such step per call; a sample
            // that ends there too is added and the next traces its primary ray)
            float sum[3] = {0.0f, 0.0f, 0.0f};
            if (!ADJ) {
                sum[0] = a.s_sum[3 * slot];
                sum[1] = a.s_sum[3 * slot + 1];
                sum[2] = a.s_sum[3 * slot + 2];
            }
            if (contrib) {
                if (ADJ) {
                    scatter_grad(a, slot, vp, term, adj_acc);
                } else {
                    sum[0] = sum[0] + T.x;
                    sum[1] = sum[1] + T.y;
                    sum[2] = sum[2] + T.z;
                }
            }
            ++smp;
            bool item_done = smp >= cn;
            if (!item_done) {
                const size_t S = a.S;
                const uint32_t w10 = ((const uint32_t*)a.s_fv)[10 * S + slot];
                const uint32_t kind = w10 >> 16;
                const bool sky = (a.flags & (RT_RENDER_SKY | RT_RENDER_SUN)) != 0;
                d = mk(a.s_fv[6 * S + slot], a.s_fv[7 * S + slot], a.s_fv[8 * S + slot]);
                if (kind == FV_END || (kind == FV_MISS && !sky)) {
                    item_done = true;  // every remaining sample contributes nothing
                } else if (kind == FV_MISS) {
                    // every remaining sample escapes along the primary ray: 1 times the miss radiance
                    const f3 t0 = miss_radiance(a, d, true);
                    if (!ADJ)
                        for (; smp < cn; ++smp) {
                            sum[0] = sum[0] + t0.x;
                            sum[1] = sum[1] + t0.y;
                            sum[2] = sum[2] + t0.z;
                        }
                    item_done = true;
                } else {
                    p = mk(a.s_fv[0 * S + slot], a.s_fv[1 * S + slot], a.s_fv[2 * S + slot]);
                    nrm = mk(a.s_fv[3 * S + slot], a.s_fv[4 * S + slot], a.s_fv[5 * S + slot]);
                    mid = ((const uint32_t*)a.s_fv)[9 * S + slot];
                    depth = w10 & 0xFFFFu;
                    T = mk(1.0f, 1.0f, 1.0f);
                    end = false;
                    contrib = false;
                    if (ADJ) a.s_nv[slot] = 0;
                    vertex_step<ADJ>(a, slot, pix, view, a.sample_offset + c0 + smp, depth, p, nrm, mid,
                                     o, d, T, end, contrib);
                    if (end) {
                        if (contrib) {
                            if (ADJ) {
                                scatter_grad(a, slot, vp, mk(1.0f, 1.0f, 1.0f), adj_acc);
                            } else {
                                sum[0] = sum[0] + T.x;
                                sum[1] = sum[1] + T.y;
                                sum[2] = sum[2] + T.z;
                            }
                        }
                        ++smp;
                        if (smp >= cn) {
                            item_done = true;
                        } else {
                            start_sample(a, slot, vp, nr, ADJ);
                            alive = true;
                        }
                    } else {
                        a.s_state[slot] = depth;
                        a.s_T[3 * slot] = T.x; a.s_T[3 * slot + 1] = T.y; a.s_T[3 * slot + 2] = T.z;
                        nr = make_ray(o, d);
                        alive = true;
                    }
                }
            }
            if (item_done) {
                if (!ADJ) {
                    const size_t q = (size_t)chunk * a.VP + vp;
                    a.partial[3 * q] = sum[0];
                    a.partial[3 * q + 1] = sum[1];
                    a.partial[3 * q + 2] = sum[2];
                }
                need_item = true;
            } else {
                if (!ADJ) {
                    a.s_sum[3 * slot] = sum[0];
                    a.s_sum[3 * slot + 1] = sum[1];
                    a.s_sum[3 * slot + 2] = sum[2];
                }
                a.s_smp[slot] = smp;
            }
        } else {
            a.s_state[slot] = depth | (cull ? CULL_BIT : 0u);
            a.s_T[3 * slot] = T.x; a.s_T[3 * slot + 1] = T.y; a.s_T[3 * slot + 2] = T.z;
            nr = make_ray(o, d);
            alive = true;
        }
    }
    // next work item for slots whose item finished, then the survivors
    // compacted into the next queue: one atomic per workgroup each
    __shared__ uint32_t lds_alloc[PT_BLOCK / 64 + 1];
    {
        const bool f[1] = {need_item};
        uint32_t* const c[1] = {&a.ctr[C_ITEM]};
        uint32_t item[1];
        block_alloc<PT_BLOCK, 1>(f, c, lds_alloc, lane, lt, item);
        if (need_item && item[0] < a.items) {
            start_item(a, slot, item[0], nr, ADJ);
            alive = true;
        }
    }
    const bool f[1] = {alive};
    uint32_t* const c[1] = {n_out};
    uint32_t pos[1];
    block_alloc<PT_BLOCK, 1>(f, c, lds_alloc, lane, lt, pos);  // also orders the LDS gradient sums
    if (alive) {
        a.q_slot[cur ^ 1][pos[0]] = slot;
        a.q_ray[cur ^ 1][pos[0]] = nr;
    }
    if (ADJ && a.adj_cols) {
        // the workgroup's gradient into its own column: launches on the
        // stream run one after another and the grid is the same every
        // iteration, so no other workgroup touches it
        for (uint32_t i = threadIdx.x; i < a.adj_cols; i += PT_BLOCK) {
            const float v = adj_acc[i];
            if (v != 0.0f) a.adj_rows[(size_t)i * gridDim.x + blockIdx.x] += v;
        }
    }
}

// dcolor[i] = sum over workgroups of adj_rows[i][*] (one block per entry)
__global__ void __launch_bounds__(256) k_adj_reduce(const float* __restrict__ rows, uint32_t nblocks,
                                                    float* __restrict__ dcolor) {
    __shared__ float part[4];
    const float* r = rows + (size_t)blockIdx.x * nblocks;
    float v = 0.0f;
    for (uint32_t j = threadIdx.x; j < nblocks; j += 256) v = v + r[j];
    for (int o = 32; o >= 1; o >>= 1) v = v + __shfl_xor(v, o);
    if ((threadIdx.x & 63u) == 0) part[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) dcolor[blockIdx.x] = (part[0] + part[1]) + (part[2] + part[3]);
}

// ---------------------------------------------------------------- cache mode
//
// Light sampling.  At a vertex x (shading normal n, material m with diffuse
// probability td and colour c) the reference estimator's diffuse branch
// draws w uniformly on the hemisphere (pdf 1/(2 pi)) and multiplies the
// throughput by c (n.w); a next hit y on a light from the front then ends the
// path with the light's own factor, whose expectation is
//   R_L = td_L * c_L / 2 + (1 - td_L)
// (a uniform hemisphere cosine averages 1/2; the other branches multiply by
// 1), and an escape inside the sun's cone adds the sun radiance.  The
// expected value of those terms is estimated here instead:
//   light: y uniform on the emissive area A (triangle by area, point by
//          sqrt-warped barycentrics), w = (y - x)/|y - x|
//          C = td c (n.w) R_L (-n_y.w) A / (2 pi |y - x|^2)
//   sun:   w uniform in the cone (solid angle 2 pi (1 - cos_sun))
//          C = td c (n.w) sun_radiance (1 - cos_sun)
// times the number of strategies present when one is picked at random, and
// an occlusion test from x + 0.01 w (the estimator's own offset).  The BSDF
// path then skips what was estimated: a front light hit after a diffuse
// bounce ends the path with no contribution, and an escape after a diffuse
// bounce gets the sky without the sun.  A light hit that is not estimated
// (seen directly, or after a mirror or refraction) contributes T * R_L, the
// expectation of the light's own random factor.  Both forms have the same
// expected pixel value, so the plain render stays the check on the cache's
// bias.
__device__ __forceinline__ f3 light_factor(const rt_material& lm) {
    const float tdl = lm.diffuse_threshold;
    return mk(tdl * lm.color[0] * 0.5f + (1.0f - tdl), tdl * lm.color[1] * 0.5f + (1.0f - tdl),
              tdl * lm.color[2] * 0.5f + (1.0f - tdl));
}

__device__ __forceinline__ bool nee_sample(const PTArgs& a, f3 p, f3 nrm, const rt_material& m, u4 rnd, rt_ray& sh,
                                           f3& C) {
    const uint32_t nls = (a.num_lights > 0 ? 1u : 0u) + (a.sun_on ? 1u : 0u);
    if (nls == 0 || !(m.diffuse_threshold > 0.0f)) return false;
    const bool use_light = a.num_lights > 0 && (!a.sun_on || u01(rnd.x) < 0.5f);
    const f3 col = mk(m.color[0], m.color[1], m.color[2]);
    if (use_light) {
        // smallest k with cdf[k] > u
        const float u = u01(rnd.y);
        uint32_t lo = 0, hi = a.num_lights - 1;
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (a.lights[mid].cdf > u) hi = mid;
            else lo = mid + 1;
        }
        const uint32_t t = a.lights[lo].tri;
        const ShadeGeom g = a.geom[t];
        const ShadeNrm nn = a.nrm[t];
        const float su = sqrtf(u01(rnd.z)), u2 = u01(rnd.w);
        const float b1 = su * (1.0f - u2), b2 = su * u2, b0 = 1.0f - su;
        const f3 y = add(add(mk(g.v0[0], g.v0[1], g.v0[2]), scale(b1, mk(g.e0[0], g.e0[1], g.e0[2]))),
                         scale(b2, mk(g.e1[0], g.e1[1], g.e1[2])));
        f3 ny = scale(b0, mk(nn.n0[0], nn.n0[1], nn.n0[2]));
        ny = add(ny, scale(b1, mk(nn.n1[0], nn.n1[1], nn.n1[2])));
        ny = add(ny, scale(b2, mk(nn.n2[0], nn.n2[1], nn.n2[2])));
        ny = normalize_ref(ny);
        const f3 wv = sub(y, p);
        const float d2 = dot(wv, wv);
        const float dist = sqrtf(d2);
        if (!(dist > 0.02f)) return false;
        const f3 w = scale(1.0f / dist, wv);
        const float cx = dot(nrm, w), cy = -dot(ny, w);
        if (!(cx > 0.0f) || !(cy > 0.0f)) return false;
        const f3 RL = light_factor(a.mat[g.mat]);
        const float f = m.diffuse_threshold * cx * cy * a.light_area * (float)nls / (6.28318530717958647f * d2);
        C = scale(f, mulv(col, RL));
        sh = make_ray(add(p, scale(0.01f, w)), w);
        sh.tmax = (dist - 0.01f) * 0.9990234375f;  // stop short of the light itself
    } else {
        const f3 s = mk(a.sun_dir[0], a.sun_dir[1], a.sun_dir[2]);
        const float ct = 1.0f - u01(rnd.y) * (1.0f - a.sun_cos);
        const float st = sqrtf(fmaxf(0.0f, 1.0f - ct * ct));
        float sp, cp;
        sincos_2pi(u01(rnd.z), sp, cp);
        // orthonormal basis around s (Duff et al., "Building an Orthonormal
        // Basis, Revisited", JCGT 2017)
        const float sg = copysignf(1.0f, s.z);
        const float ia = -1.0f / (sg + s.z);
        const float ib = s.x * s.y * ia;
        const f3 t1 = mk(1.0f + sg * s.x * s.x * ia, sg * ib, -sg * s.x);
        const f3 t2 = mk(ib, sg + s.y * s.y * ia, -s.y);
        const f3 w = add(scale(ct, s), scale(st, add(scale(cp, t1), scale(sp, t2))));
        const float cx = dot(nrm, w);
        if (!(cx > 0.0f)) return false;
        const float f = m.diffuse_threshold * cx * (1.0f - a.sun_cos) * (float)nls;
        C = scale(f, mulv(col, mk(a.sun_rad[0], a.sun_rad[1], a.sun_rad[2])));
        sh = make_ray(add(p, scale(0.01f, w)), w);
    }
    return true;
}

__device__ __forceinline__ void add3(float* dst, f3 v) {
    atomicAdd(dst, v.x);
    atomicAdd(dst + 1, v.y);
    atomicAdd(dst + 2, v.z);
}

// A training record's target has one writer per launch (the slot that owns
// the record: its path end in the shade step, or its one shadow ray in the
// resolve step), and the launches run in path order, so the target is a
// plain sum in the order of the path's vertices.
__device__ __forceinline__ void acc3(float* dst, f3 v) {
    dst[0] = dst[0] + v.x;
    dst[1] = dst[1] + v.y;
    dst[2] = dst[2] + v.z;
}

// Shade step of rt_render_nrc: the plain state machine plus the query
// vertex, training records and light sampling.
__global__ void __launch_bounds__(PT_BLOCK) k_pt_shade_nrc(PTArgs a, int cur) {
    const uint32_t n = a.ctr[cur ? C_N1 : C_N0];
    uint32_t* n_out = &a.ctr[cur ? C_N0 : C_N1];
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k == 0) *(unsigned long long*)&a.ctr[C_RAYS] += n;
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    bool alive = false, need_item = false;
    uint32_t slot = 0;
    rt_ray nr;
    // work this step enqueues
    bool push_q = false, want_rec = false, push_sh = false;
    float f9[9];
    f3 qw = mk(0.0f, 0.0f, 0.0f), shw_pix = mk(0.0f, 0.0f, 0.0f), shw_rec = mk(0.0f, 0.0f, 0.0f);
    uint32_t pidx = 0, ridx = NO_REC, rkey = 0;
    bool sh_to_rec = false;
    rt_ray sh;
    if (k < n) {
        slot = a.q_slot[cur][k];
        const rt_ray r = a.q_ray[cur][k];
        const rt_hit h = a.hits[k];
        const uint32_t item = a.s_item[slot];
        const uint32_t vp = item % a.VP;
        const uint32_t chunk = item / a.VP;
        const uint32_t view = vp / a.P, pix = vp % a.P;
        pidx = chunk * a.VP + vp;
        const uint32_t c0 = chunk_first(a, chunk), cn = chunk_first(a, chunk + 1) - c0;
        uint32_t smp = a.s_smp[slot];
        const uint32_t st = a.s_state[slot];
        uint32_t depth = st & 0xFFFFu;
        bool cull = (st & CULL_BIT) != 0;
        uint32_t fh = (st >> FH_SHIFT) & FH_MASK;
        bool rec = (st & REC_BIT) != 0;
        const bool prev_diff = (st & DIFF_BIT) != 0;
        bool diff_now = false;
        f3 T = mk(a.s_T[3 * slot], a.s_T[3 * slot + 1], a.s_T[3 * slot + 2]);
        f3 Tq = mk(1.0f, 1.0f, 1.0f);
        if (rec) {
            Tq = mk(a.s_Tq[3 * slot], a.s_Tq[3 * slot + 1], a.s_Tq[3 * slot + 2]);
            ridx = a.s_ridx[slot];
        }
        f3 o = mk(r.ox, r.oy, r.oz), d = mk(r.dx, r.dy, r.dz);
        bool end = false, contrib = false;
        if (h.tri == RT_MISS) {
            end = true;
            if (a.flags & (RT_RENDER_SKY | RT_RENDER_SUN)) {
                contrib = true;
                const f3 L = miss_radiance(a, d, !(a.nee && prev_diff));
                T = mulv(T, L);
                Tq = mulv(Tq, L);
            }
        } else {
            f3 p, nrm;
            uint32_t mid;
            surface(a, h, o, d, p, nrm, mid);
            const float ndotd = dot(nrm, d);
            if (cull && ndotd > 0.0f) {
                o = add(p, scale(0.01f, d));
            } else {
                cull = false;
                const rt_material m = a.mat[mid];
                const uint32_t gsmp = a.sample_offset + c0 + smp;
                const bool light_end = m.light_source && ndotd <= 0.0f;
                bool stop = false;
                if (light_end && a.nee) {
                    end = true;
                    stop = true;
                    if (!prev_diff) {  // not estimated by a light sample: the light's expected factor
                        const f3 RL = light_factor(m);
                        T = mulv(T, RL);
                        Tq = mulv(Tq, RL);
                        contrib = true;
                    }
                } else if (!light_end && fh == a.nrc_q && depth + 1 < a.max_depth &&
                           (!a.ddgi_mode || m.diffuse_threshold >= 1.0f)) {
                    const bool train = a.nrc_every != 0 && train_hash(pix, gsmp, view, a.seed_lo) % a.nrc_every == 0;
                    f9[0] = p.x; f9[1] = p.y; f9[2] = p.z;
                    f9[3] = nrm.x; f9[4] = nrm.y; f9[5] = nrm.z;
                    f9[6] = d.x; f9[7] = d.y; f9[8] = d.z;
                    if (train) {
                        want_rec = true;
                        rec = true;
                        Tq = mk(1.0f, 1.0f, 1.0f);
                        rkey = vp * a.key_n + (gsmp - a.sample_offset);
                    } else if (a.nrc_query) {
                        push_q = true;
                        // probe field: the diffuse bounce's expectation, albedo / 2 * lookup
                        qw = a.ddgi_mode ? mulv(T, scale(0.5f, mk(m.color[0], m.color[1], m.color[2]))) : T;
                        end = true;
                        stop = true;
                    }
                }
                ++fh;
                u4 c4;
                c4.x = pix; c4.y = gsmp; c4.z = depth; c4.w = view;
                // a probe-field query holds the indirect part only: its vertex
                // takes the light sample too
                if (a.nee && !light_end && depth + 1 < a.max_depth && (!stop || (push_q && a.ddgi_mode))) {
                    u4 c5 = c4;
                    c5.w = view | 0x80000000u;
                    f3 Cn;
                    if (nee_sample(a, p, nrm, m, philox4x32_10(c5, a.seed_lo, a.seed_hi), sh, Cn)) {
                        push_sh = true;
                        shw_pix = mulv(T, Cn);
                        sh_to_rec = rec;
                        shw_rec = mulv(Tq, Cn);
                    }
                }
                if (!stop) {
                    const u4 rnd = philox4x32_10(c4, a.seed_lo, a.seed_hi);
                    const float xi0 = u01(rnd.x);
                    const bool diffuse = xi0 <= m.diffuse_threshold;
                    if (diffuse) d = hemisphere_dir(u01(rnd.y), u01(rnd.z), nrm);
                    else if (xi0 <= m.reflect_threshold) d = reflect_ref(d, nrm);
                    else if (xi0 <= m.transparent_threshold) d = custom_refract_ref(d, nrm, m.refraction_index);
                    o = add(p, scale(0.01f, d));
                    if (diffuse) {
                        const f3 f = scale(dot(nrm, d), mk(m.color[0], m.color[1], m.color[2]));
                        T = mulv(T, f);
                        Tq = mulv(Tq, f);
                    }
                    diff_now = diffuse;
                    if (light_end) {
                        end = true;
                        contrib = true;
                    }
                }
            }
            ++depth;
            if (depth >= a.max_depth) end = true;
        }
        if (end) {
            if (contrib) {
                a.s_sum[3 * slot] = a.s_sum[3 * slot] + T.x;
                a.s_sum[3 * slot + 1] = a.s_sum[3 * slot + 1] + T.y;
                a.s_sum[3 * slot + 2] = a.s_sum[3 * slot + 2] + T.z;
                if (rec && !want_rec && ridx != NO_REC) acc3(&a.t_tgt[3 * (size_t)ridx], Tq);
            }
            ++smp;
            if (smp < cn) {
                a.s_smp[slot] = smp;
                start_sample(a, slot, vp, nr, false);
                alive = true;
            } else {
                // cache predictions and light samples land in the same sums
                add3(&a.partial[3 * (size_t)pidx], mk(a.s_sum[3 * slot], a.s_sum[3 * slot + 1], a.s_sum[3 * slot + 2]));
                need_item = true;
            }
        } else {
            a.s_state[slot] = depth | (cull ? CULL_BIT : 0u) | (fh << FH_SHIFT) | (rec ? REC_BIT : 0u) |
                              (diff_now ? DIFF_BIT : 0u);
            a.s_T[3 * slot] = T.x; a.s_T[3 * slot + 1] = T.y; a.s_T[3 * slot + 2] = T.z;
            if (rec) {
                a.s_Tq[3 * slot] = Tq.x; a.s_Tq[3 * slot + 1] = Tq.y; a.s_Tq[3 * slot + 2] = Tq.z;
            }
            nr = make_ray(o, d);
            alive = true;
        }
    }
    // cache queries, training records opened at the query vertex, shadow
    // rays and next work items: one atomic per counter and workgroup
    __shared__ uint32_t lds_alloc[4 * (PT_BLOCK / 64 + 1)];
    {
        const bool f[4] = {push_q, want_rec, push_sh, need_item};
        uint32_t* const c[4] = {&a.ctr[C_QN], &a.ctr[C_TN], &a.ctr[sn_slot(cur)], &a.ctr[C_ITEM]};
        uint32_t pos[4];
        block_alloc<PT_BLOCK, 4>(f, c, lds_alloc, lane, lt, pos);
        if (push_q) {
            if (pos[0] < a.q_cap) {
                for (int j = 0; j < 9; ++j) a.q_feat[9 * (size_t)pos[0] + j] = f9[j];
                a.q_w[3 * (size_t)pos[0]] = qw.x; a.q_w[3 * (size_t)pos[0] + 1] = qw.y;
                a.q_w[3 * (size_t)pos[0] + 2] = qw.z;
                a.q_idx[pos[0]] = pidx;
            } else {
                atomicAdd(&a.ctr[C_QDROP], 1u);
            }
        }
        // a training record: features now, target accumulated from here on
        // (records past the capacity are not kept)
        if (want_rec) {
            ridx = pos[1] < a.t_cap ? pos[1] : NO_REC;
            if (ridx != NO_REC) {
                for (int j = 0; j < 9; ++j) a.t_feat[9 * (size_t)ridx + j] = f9[j];
                a.t_tgt[3 * (size_t)ridx] = 0.0f; a.t_tgt[3 * (size_t)ridx + 1] = 0.0f;
                a.t_tgt[3 * (size_t)ridx + 2] = 0.0f;
                a.t_key[ridx] = rkey;
            }
            a.s_ridx[slot] = ridx;
        }
        if (push_sh) {
            const uint32_t q = pos[2];  // < S: one per live slot
            a.sh_ray[q] = sh;
            float* w = a.sh_w + 6 * (size_t)q;
            w[0] = shw_pix.x; w[1] = shw_pix.y; w[2] = shw_pix.z;
            w[3] = shw_rec.x; w[4] = shw_rec.y; w[5] = shw_rec.z;
            a.sh_idx[2 * (size_t)q] = pidx;
            a.sh_idx[2 * (size_t)q + 1] = sh_to_rec ? ridx : NO_REC;
        }
        if (need_item && pos[3] < a.items) {
            start_item(a, slot, pos[3], nr, false);
            alive = true;
        }
    }
    // compact the survivors into the next queue
    const bool f[1] = {alive};
    uint32_t* const c[1] = {n_out};
    uint32_t pos[1];
    block_alloc<PT_BLOCK, 1>(f, c, lds_alloc, lane, lt, pos);
    if (alive) {
        a.q_slot[cur ^ 1][pos[0]] = slot;
        a.q_ray[cur ^ 1][pos[0]] = nr;
    }
}

// Unoccluded light samples add their weights to the pixel sums and records
// (the shadow rays queued by the iteration of parity par).
__global__ void __launch_bounds__(PT_BLOCK) k_shadow_resolve(PTArgs a, int par) {
    const uint32_t n = a.ctr[sn_slot(par)];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) *(unsigned long long*)&a.ctr[C_SHRAYS] += n;
    if (i >= n || a.sh_occ[i]) return;
    const float* w = a.sh_w + 6 * (size_t)i;
    add3(&a.partial[3 * (size_t)a.sh_idx[2 * (size_t)i]], mk(w[0], w[1], w[2]));
    const uint32_t r = a.sh_idx[2 * (size_t)i + 1];
    if (r != NO_REC) acc3(&a.t_tgt[3 * (size_t)r], mk(w[3], w[4], w[5]));
}

__global__ void __launch_bounds__(PT_BLOCK) k_pt_finish(PTArgs a) {
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.VP) return;
    float s[3] = {0.0f, 0.0f, 0.0f};
    for (uint32_t c = 0; c < a.K; ++c)
        for (int ch = 0; ch < 3; ++ch) s[ch] = s[ch] + a.partial[3 * ((size_t)c * a.VP + p) + ch];
    for (int ch = 0; ch < 3; ++ch) {
        // finalColor *= 2.0f * M_PI * (1.0f / float(raysPerPixel))  (double scale)
        float v = (float)((double)s[ch] * a.scale);
        if (a.rgb) a.rgb[3 * (size_t)p + ch] = v;
        if (a.rgb8) {
            // clamp to [0, 1] then (char)(v * 255.0f) (graphics_engine.cpp:90-95)
            float cv = fminf(fmaxf(v, 0.0f), 1.0f);
            a.rgb8[3 * (size_t)p + ch] = (uint8_t)(int)(cv * 255.0f);
        }
    }
}

// after a cache flush: the queries answered join the total, the queue restarts
__global__ void k_flush_queries(uint32_t* ctr) {
    if (threadIdx.x == 0) {
        *(unsigned long long*)&ctr[C_QTOT] += ctr[C_QN];
        ctr[C_QN] = 0;
    }
}

__global__ void k_zero(float* p, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}

__global__ void k_iota(uint32_t* p, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = i;
}

// training records in sample order: out[i] = record order[i]
__global__ void k_gather_records(const float* __restrict__ feat, const float* __restrict__ tgt,
                                 const uint32_t* __restrict__ order, uint32_t n, float* __restrict__ ofeat,
                                 float* __restrict__ otgt) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t r = order[i];
    for (int j = 0; j < 9; ++j) ofeat[9 * (size_t)i + j] = feat[9 * (size_t)r + j];
    for (int j = 0; j < 3; ++j) otgt[3 * (size_t)i + j] = tgt[3 * (size_t)r + j];
}

// Radiance of the probe rays of a field update (rt_ddgi_update): what a path
// leaving a purely diffuse vertex along the ray adds beyond that vertex's own
// light sample.  A miss sees the sky without the sun; a light seen from the
// front, or any back face (the probe is inside geometry), adds 0; a surface
// adds the mean of `count` light samples (queued here, tested by the any-hit
// kernel, summed by k_ddgi_resolve) and, with multi-bounce, td * albedo / 2 *
// the previous field's lookup there.  The distance runs from the probe to the
// hit moved normal_bias along its normal (majercik2019_probe_raygen.rgen:40-47).
// Light sample k of ray j draws Philox (j, update, k, 0x50000000).
__global__ void __launch_bounds__(PT_BLOCK) k_ddgi_shade(PTArgs a, DdgiShade b) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= b.n) return;
    const rt_ray r = b.rays[j];
    const rt_hit h = b.hits[j];
    const f3 o = mk(r.ox, r.oy, r.oz), d = mk(r.dx, r.dy, r.dz);
    f3 rad = mk(0.0f, 0.0f, 0.0f);
    float dist = b.grid.max_dist, hit = 0.0f;
    bool sampled = false;
    if (h.tri == RT_MISS) {
        rad = miss_radiance(a, d, false);
    } else {
        f3 p, nrm;
        uint32_t mid;
        surface(a, h, o, d, p, nrm, mid);
        const f3 dv = sub(o, add(p, scale(b.grid.normal_bias, nrm)));
        dist = sqrtf(dot(dv, dv));
        hit = 1.0f;
        const rt_material m = a.mat[mid];
        if (!(dot(nrm, d) > 0.0f) && !m.light_source) {
            sampled = true;
            for (uint32_t k = 0; k < b.count; ++k) {
                const size_t q = (size_t)j * b.count + k;
                u4 c;
                c.x = j; c.y = b.update; c.z = k; c.w = 0x50000000u;
                rt_ray sr;
                f3 C = mk(0.0f, 0.0f, 0.0f);
                if (!(a.nee && nee_sample(a, p, nrm, m, philox4x32_10(c, a.seed_lo, a.seed_hi), sr, C))) {
                    sr = make_ray(p, mk(0.0f, 0.0f, 0.0f));  // no direction: never occluded, weight 0
                    C = mk(0.0f, 0.0f, 0.0f);
                }
                b.sh_ray[q] = sr;
                b.sh_w[3 * q] = C.x; b.sh_w[3 * q + 1] = C.y; b.sh_w[3 * q + 2] = C.z;
            }
            if (b.multi) {
                const f3 e = ddgi_lookup(b.grid, b.atl, p, nrm, d);
                rad = mulv(scale(0.5f * m.diffuse_threshold, mk(m.color[0], m.color[1], m.color[2])), e);
            }
        }
    }
    if (!sampled) {
        for (uint32_t k = 0; k < b.count; ++k) {
            const size_t q = (size_t)j * b.count + k;
            b.sh_ray[q] = make_ray(o, mk(0.0f, 0.0f, 0.0f));
            b.sh_w[3 * q] = 0.0f; b.sh_w[3 * q + 1] = 0.0f; b.sh_w[3 * q + 2] = 0.0f;
        }
    }
    float* sf = b.surf + 8 * (size_t)j;
    sf[3] = rad.x; sf[4] = rad.y; sf[5] = rad.z;
    sf[6] = dist;
    sf[7] = hit;
}

inline uint32_t div_up(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

}  // namespace

int launch_render(rt_ctx* c, const rt_camera* cams, uint32_t V, const rt_settings* s, float* d_rgb,
                  uint8_t* d_rgb8, const float* d_dimage, float* d_dcolor, const NrcFrame* nf) {
    const bool adj = d_dcolor != nullptr;
    const bool nrc = nf != nullptr;
    if (adj && nrc) return fail(c, RT_E_INVALID, "launch_render: adjoint and cache modes exclude each other");
    if (adj && s->max_depth > ADJ_MAX_DEPTH)
        return fail(c, RT_E_UNSUPPORTED, "rt_render_adjoint: max_depth above 16");
    const rt_camera* cam = &cams[0];
    const uint64_t P1 = (uint64_t)cam->width * cam->height;
    const uint64_t P = P1 * V;  // view-major (view, pixel) pairs
    if (P > (1ull << 31)) return fail(c, RT_E_UNSUPPORTED, "rt_render: views * pixels above 2^31");
    const uint32_t N = s->spp;
    // samples of a pixel are split into K <= N chunks when there are too
    // few pixels to fill the device; deterministic mode keeps one chunk so
    // every pixel's sum has the reference's sample order.  (The adjoint
    // scatters per path, so chunks change only its summation order.)
    const uint64_t target_slots = 1ull << c->pt_slots_log2;
    uint32_t K = 1;
    if (!(s->flags & RT_RENDER_DETERMINISTIC)) {
        const uint64_t want = (target_slots + P - 1) / P;
        K = (uint32_t)(want < N ? want : N);
    }
    const uint64_t items = P * K;
    if (items > 0xFFFFFFFFull) return fail(c, RT_E_UNSUPPORTED, "rt_render: pixels * chunks above 2^32");
    if (nrc && 3 * items > 0xFFFFFFFFull) return fail(c, RT_E_UNSUPPORTED, "rt_render_nrc: pixels * chunks above 2^32 / 3");
    const uint32_t S = (uint32_t)(items < target_slots ? items : target_slots);
    // adjoint: gradient sums in LDS per workgroup when 3 * materials fit
    const uint32_t adj_cols = adj && 3 * (uint64_t)c->M <= ADJ_LDS_MAX ? 3u * c->M : 0u;
    int rc;
    const size_t szS = S;
    // cache mode: queries flushed every 8 iterations (at most S per
    // iteration); training records: 4 frames' worth of steps, and at least
    // twice the expected number of training samples (a binomial count, so
    // the capacity is practically never reached and no record is dropped)
    const size_t q_cap = nrc ? 8 * szS : 0;
    size_t t_cap = nrc ? (size_t)4 * nf->opt.batch * nf->opt.max_steps : 0;
    const bool records = nrc && nf->opt.train_every && !(nf->opt.flags & RT_NRC_NO_RECORDS);
    if (records) {
        const size_t expect = (size_t)((P * N + nf->opt.train_every - 1) / nf->opt.train_every);
        if (2 * expect + 4096 > t_cap) t_cap = 2 * expect + 4096;
        // a record's sort key is its sample's index in the frame
        if (P * N > 0xFFFFFFFFull)
            return fail(c, RT_E_UNSUPPORTED, "rt_render_nrc: training records with pixels * samples above 2^32");
    }
    if (t_cap > 0x3FFFFFFFull) return fail(c, RT_E_UNSUPPORTED, "rt_render_nrc: training records above 2^30");
    // shadow rays (one per live slot and iteration): ray, 6 weights, 2 indices, flag
    const size_t sh_bytes = szS * (sizeof(rt_ray) + 24 + 8 + 1);
    if (nrc && ((rc = ensure(c, c->pt_nrc_query, q_cap * (36 + 12 + 4))) ||
                (rc = ensure(c, c->pt_nrc_train, (t_cap ? t_cap : 1) * (36 + 12 + 4))) ||
                (rc = ensure(c, c->pt_nrc_sorted, (t_cap ? t_cap : 1) * (36 + 12 + 12))) ||
                (rc = ensure(c, c->pt_shadow, sh_bytes))))
        return rc;
    if ((rc = ensure(c, c->pt_state, szS * (4 * 3 + 12 * 2) + (adj ? szS * (8 * ADJ_MAX_DEPTH + 4) : 0) +
                                         (nrc ? szS * 16 : szS * 44))) ||
        (rc = ensure(c, c->pt_queue[0], 4 * szS + 16 + sizeof(rt_ray) * szS)) ||
        (rc = ensure(c, c->pt_queue[1], 4 * szS + 16 + sizeof(rt_ray) * szS)) ||
        (rc = ensure(c, c->hits, szS * sizeof(rt_hit))) || (rc = ensure(c, c->pt_ctr, 64)) ||
        (rc = ensure(c, c->pt_partial, adj ? 4 * adj_cols * (size_t)div_up(S, PT_BLOCK) + 4 : 12 * items)) ||
        (rc = ensure(c, c->pt_cams, sizeof(rt_camera) * V)))
        return rc;
    RT_HIP(c, hipMemcpyAsync(c->pt_cams.ptr, cams, sizeof(rt_camera) * V, hipMemcpyHostToDevice, c->stream));
    PTArgs a;
    a.geom = (const ShadeGeom*)c->shade_geom.ptr;
    a.nrm = (const ShadeNrm*)c->shade_nrm.ptr;
    a.mat = (const rt_material*)c->mat.ptr;
    a.cams = (const rt_camera*)c->pt_cams.ptr;
    a.W = cam->width; a.H = cam->height; a.P = (uint32_t)P1;
    a.V = V; a.VP = (uint32_t)P;
    a.K = K; a.spp = N; a.items = (uint32_t)items; a.S = S;
    a.max_depth = s->max_depth; a.sample_offset = s->sample_offset; a.flags = s->flags;
    a.seed_lo = (uint32_t)s->seed; a.seed_hi = (uint32_t)(s->seed >> 32);
    for (int k = 0; k < 3; ++k) {
        a.sky[k] = s->sky[k];
        a.sun_dir[k] = s->sun_dir[k];
        a.sun_rad[k] = s->sun_radiance[k];
    }
    a.sun_cos = s->sun_cos;
    a.scale = 2.0 * M_PI * (double)(1.0f / (float)N);
    char* st = (char*)c->pt_state.ptr;
    a.s_item = (uint32_t*)st; st += 4 * szS;
    a.s_smp = (uint32_t*)st; st += 4 * szS;
    a.s_state = (uint32_t*)st; st += 4 * szS;
    a.s_T = (float*)st; st += 12 * szS;
    a.s_sum = (float*)st; st += 12 * szS;
    a.s_amat = adj ? (uint32_t*)st : nullptr; st += adj ? 4 * ADJ_MAX_DEPTH * szS : 0;
    a.s_acos = adj ? (float*)st : nullptr; st += adj ? 4 * ADJ_MAX_DEPTH * szS : 0;
    a.s_nv = adj ? (uint32_t*)st : nullptr; st += adj ? 4 * szS : 0;
    a.s_ridx = nrc ? (uint32_t*)st : nullptr; st += nrc ? 4 * szS : 0;
    a.s_Tq = nrc ? (float*)st : nullptr;
    a.s_fv = nrc ? nullptr : (float*)st;
    a.nee = nrc && !(nf->opt.flags & RT_NRC_NO_NEE) && (c->num_lights > 0 || ((s->flags & RT_RENDER_SUN) && s->sun_cos < 1.0f));
    a.sun_on = (s->flags & RT_RENDER_SUN) && s->sun_cos < 1.0f;
    a.lights = (const LightTri*)c->lights.ptr;
    a.num_lights = c->num_lights;
    a.light_area = c->light_area;
    {
        char* sp = (char*)c->pt_shadow.ptr;
        a.sh_ray = nrc ? (rt_ray*)sp : nullptr; sp += sizeof(rt_ray) * szS;
        a.sh_w = nrc ? (float*)sp : nullptr; sp += 24 * szS;
        a.sh_idx = nrc ? (uint32_t*)sp : nullptr; sp += 8 * szS;
        a.sh_occ = nrc ? (uint8_t*)sp : nullptr;
    }
    a.nrc_q = nrc ? nf->opt.query_vertex : 0;
    a.nrc_every = nrc && !(nf->opt.flags & RT_NRC_NO_RECORDS) ? nf->opt.train_every : 0;
    a.nrc_query = nrc && !(nf->opt.flags & RT_NRC_NO_QUERY);
    a.ddgi_mode = nrc && nf->ddgi != nullptr;
    a.q_feat = nrc ? (float*)c->pt_nrc_query.ptr : nullptr;
    a.q_w = nrc ? a.q_feat + 9 * q_cap : nullptr;
    a.q_idx = nrc ? (uint32_t*)(a.q_w + 3 * q_cap) : nullptr;
    a.q_cap = (uint32_t)q_cap;
    a.t_feat = nrc ? (float*)c->pt_nrc_train.ptr : nullptr;
    a.t_tgt = nrc ? a.t_feat + 9 * t_cap : nullptr;
    a.t_key = nrc ? (uint32_t*)(a.t_tgt + 3 * t_cap) : nullptr;
    a.t_cap = (uint32_t)t_cap;
    a.key_n = N;
    // queue buffers hold S slot ids then S rays (16-byte aligned)
    for (int q = 0; q < 2; ++q) {
        a.q_slot[q] = (uint32_t*)c->pt_queue[q].ptr;
        a.q_ray[q] = (rt_ray*)((char*)c->pt_queue[q].ptr + ((4 * szS + 15) & ~(size_t)15));
    }
    a.hits = (rt_hit*)c->hits.ptr;
    a.ctr = (uint32_t*)c->pt_ctr.ptr;
    a.partial = (float*)c->pt_partial.ptr;
    a.rgb = d_rgb;
    a.rgb8 = d_rgb8;
    a.dimg = d_dimage;
    a.dcolor = d_dcolor;
    a.adj_rows = adj_cols ? (float*)c->pt_partial.ptr : nullptr;
    a.adj_cols = adj_cols;

    hipStream_t stream = c->stream;
    RT_HIP(c, hipEventRecord(c->ev[0], stream));
    if (adj && !adj_cols)
        hipLaunchKernelGGL(k_zero, dim3(div_up(3ull * c->M, 256)), dim3(256), 0, stream, d_dcolor, 3u * c->M);
    if (adj_cols)
        hipLaunchKernelGGL(k_zero, dim3(div_up((uint64_t)adj_cols * div_up(S, PT_BLOCK), 256)), dim3(256), 0, stream,
                           a.adj_rows, adj_cols * div_up(S, PT_BLOCK));
    if (nrc) hipLaunchKernelGGL(k_zero, dim3(div_up(3ull * items, 256)), dim3(256), 0, stream, a.partial, (uint32_t)(3 * items));
    if (adj) hipLaunchKernelGGL(k_pt_init<true>, dim3(div_up(S, PT_BLOCK)), dim3(PT_BLOCK), 0, stream, a);
    else hipLaunchKernelGGL(k_pt_init<false>, dim3(div_up(S, PT_BLOCK)), dim3(PT_BLOCK), 0, stream, a);
    RT_HIP(c, hipGetLastError());
    // upper bound on iterations: an item is at most ceil(N / K) samples of at
    // most max_depth extends (L iterations), and a slot takes the next item as
    // soon as its own ends, so by iteration (ceil(items / S) - 1) L every slot
    // has finished that many items and every item has started: the last one
    // ends by ceil(items / S) L.  (Reaching the bound ends the frame: the
    // host then waits for the chunk it just queued instead of queueing more.)
    const uint64_t bound = (uint64_t)div_up(items, S) * div_up(N, K) * s->max_depth + 2;
    int cur = 0;
    uint64_t iter = 0;
    uint32_t launches = 1;
    // The host enqueues chunks of 8 extend/shade iterations and never drains
    // the stream inside the frame: after each chunk it copies the counters to
    // pinned memory behind an event, and before deciding on the next chunk it
    // waits only for the previous chunk's copy, so the device always has the
    // current chunk queued.  Iterations run after the last path ended are
    // no-ops (every kernel reads its count on the device).  Cache queries are
    // answered once per chunk, the inference reading the queue's count on the
    // device.
    if ((rc = ensure_snap(c))) return rc;
    int snap_cur[2] = {0, 0};
    uint64_t chunk = 0;
    bool done = false;
    while (!done) {
        // Shadow rays of a step are traced with the next step's rays, in one
        // mixed launch, and their weights added right after it, before that
        // step's shading; the last step of a chunk traces and resolves its own
        // at once, so every cache flush still follows all light samples of
        // the steps before it (the pixel sums take their terms in the same
        // order either way).
        bool pending = false;  // the previous step's shadow rays await their trace
        for (int rep = 0; rep < 8 && iter < bound; ++rep, ++iter) {
            // the trace clears the counters the shade step appends to (the
            // next queue's; this step's shadow rays')
            uint32_t* const z_sh = nrc && a.nee ? &a.ctr[sn_slot(cur)] : nullptr;
            if (pending) {
                if ((rc = launch_trace_mixed(c, a.q_ray[cur], &a.ctr[cur ? C_N1 : C_N0], a.hits, a.sh_ray,
                                             &a.ctr[sn_slot(cur ^ 1)], a.sh_occ, 2 * S, &a.ctr[cur ? C_N0 : C_N1],
                                             z_sh)))
                    return rc;
                hipLaunchKernelGGL(k_shadow_resolve, dim3(div_up(S, PT_BLOCK)), dim3(PT_BLOCK), 0, stream, a, cur ^ 1);
                RT_HIP(c, hipGetLastError());
                ++launches;
                pending = false;
            } else if ((rc = launch_trace(c, a.q_ray[cur], S, a.hits, false, &a.ctr[cur ? C_N1 : C_N0],
                                          &a.ctr[cur ? C_N0 : C_N1], z_sh))) {
                return rc;
            }
            if (nrc) {
                hipLaunchKernelGGL(k_pt_shade_nrc, dim3(div_up(S, PT_BLOCK)), dim3(PT_BLOCK), 0, stream, a, cur);
                RT_HIP(c, hipGetLastError());
                if (a.nee) {
                    if (rep < 7 && iter + 1 < bound) {
                        pending = true;
                    } else {
                        // light samples: occlusion of the queued shadow rays, then their weights
                        if ((rc = launch_trace_any(c, a.sh_ray, S, a.sh_occ, &a.ctr[sn_slot(cur)]))) return rc;
                        hipLaunchKernelGGL(k_shadow_resolve, dim3(div_up(S, PT_BLOCK)), dim3(PT_BLOCK), 0, stream,
                                           a, cur);
                        RT_HIP(c, hipGetLastError());
                        launches += 2;
                    }
                }
            } else if (adj) {
                hipLaunchKernelGGL(k_pt_shade<true>, dim3(div_up(S, PT_BLOCK)), dim3(PT_BLOCK), 0, stream, a, cur);
            } else {
                hipLaunchKernelGGL(k_pt_shade<false>, dim3(div_up(S, PT_BLOCK)), dim3(PT_BLOCK), 0, stream, a, cur);
            }
            RT_HIP(c, hipGetLastError());
            launches += 2;
            cur ^= 1;
        }
        if (nrc && a.nrc_query) {
            // the queued samples' cache predictions, added to their pixels
            // (at most q_cap: 8 iterations of at most S queries each)
            if (nf->ddgi) rc = ddgi_query_scatter(nf->ddgi, a.q_feat, a.q_w, a.q_idx, a.q_cap, a.partial, &a.ctr[C_QN]);
            else rc = nrc_infer_scatter(nf->cache, a.q_feat, a.q_w, a.q_idx, a.q_cap, a.partial, &a.ctr[C_QN]);
            if (rc) return rc;
            hipLaunchKernelGGL(k_flush_queries, dim3(1), dim3(64), 0, stream, a.ctr);
            RT_HIP(c, hipGetLastError());
            launches += 2;
        }
        const int slot = (int)(chunk & 1);
        RT_HIP(c, hipMemcpyAsync(c->snap + 16 * slot, a.ctr, C_WORDS * sizeof(uint32_t), hipMemcpyDeviceToHost,
                                 stream));
        RT_HIP(c, hipEventRecord(c->snap_ev[slot], stream));
        snap_cur[slot] = cur;
        ++chunk;
        // the previous chunk's counters (or, at the iteration bound, this one's)
        const bool at_bound = iter >= bound;
        if (chunk >= 2 || at_bound) {
            const int ps = at_bound ? slot : (int)((chunk - 2) & 1);
            RT_HIP(c, hipEventSynchronize(c->snap_ev[ps]));
            const uint32_t* hc = c->snap + 16 * ps;
            if (nrc && hc[C_QDROP] != 0) return fail(c, RT_E_DEVICE, "rt_render_nrc: cache query buffer overflow");
            const uint32_t live = hc[snap_cur[ps] ? C_N1 : C_N0];
            if (live == 0) done = true;
            else if (at_bound) return fail(c, RT_E_DEVICE, "rt_render: paths still live past the iteration bound");
        }
    }
    if (!adj) {
        hipLaunchKernelGGL(k_pt_finish, dim3(div_up(P, PT_BLOCK)), dim3(PT_BLOCK), 0, stream, a);
        RT_HIP(c, hipGetLastError());
        ++launches;
    }
    if (adj_cols) {
        hipLaunchKernelGGL(k_adj_reduce, dim3(adj_cols), dim3(256), 0, stream, (const float*)a.adj_rows,
                           div_up(S, PT_BLOCK), d_dcolor);
        RT_HIP(c, hipGetLastError());
        ++launches;
    }
    if (nrc) {
        RT_HIP(c, hipEventRecord(c->ev[4], stream));
        uint32_t hc[C_WORDS];
        RT_HIP(c, hipMemcpyAsync(hc, a.ctr, sizeof(hc), hipMemcpyDeviceToHost, stream));
        RT_HIP(c, hipStreamSynchronize(stream));
        const uint32_t tn = hc[C_TN];
        nf->stats->records = tn;
        nf->stats->queries = (uint64_t)hc[C_QTOT] | ((uint64_t)hc[C_QTOT + 1] << 32);
        nf->stats->shadow_rays = (uint64_t)hc[C_SHRAYS] | ((uint64_t)hc[C_SHRAYS + 1] << 32);
        const uint32_t avail = tn < a.t_cap ? tn : a.t_cap;
        c->nrc_records = avail;
        c->nrc_tcap = a.t_cap;
        // records in the order of their samples, (view, pixel, sample), so
        // that the batches, and with them the trained network, do not
        // depend on the order in which slots took record indices
        float* sfeat = (float*)c->pt_nrc_sorted.ptr;
        float* stgt = sfeat + 9 * (size_t)a.t_cap;
        uint32_t* skey = (uint32_t*)(stgt + 3 * (size_t)a.t_cap);
        if (avail > 0) {
            uint32_t* keys[2] = {a.t_key, skey};
            uint32_t* vals[2] = {skey + a.t_cap, skey + 2 * (size_t)a.t_cap};
            hipLaunchKernelGGL(k_iota, dim3(div_up(avail, PT_BLOCK)), dim3(PT_BLOCK), 0, stream, vals[0], avail);
            uint32_t bits = 1;
            while (bits < 32 && (P * N - 1) >> bits) ++bits;
            int sc = 0;
            if ((rc = radix_sort_pairs(c, keys, vals, avail, (bits + 7) / 8, &sc))) return rc;
            hipLaunchKernelGGL(k_gather_records, dim3(div_up(avail, PT_BLOCK)), dim3(PT_BLOCK), 0, stream,
                               (const float*)a.t_feat, (const float*)a.t_tgt, (const uint32_t*)vals[sc], avail,
                               sfeat, stgt);
            RT_HIP(c, hipGetLastError());
            launches += 2 + 5 * ((bits + 7) / 8);
        }
        if (!(nf->opt.flags & RT_NRC_NO_TRAIN) && nf->cache) {
            if ((rc = nrc_train_records(nf->cache, sfeat, stgt, avail, nf->opt.batch, nf->opt.max_steps,
                                        &nf->stats->steps, &nf->stats->loss)))
                return rc;
            launches += 2 * nf->stats->steps;
        }
    }
    RT_HIP(c, hipEventRecord(c->ev[1], stream));
    unsigned long long rays = 0;
    RT_HIP(c, hipMemcpyAsync(&rays, &a.ctr[C_RAYS], 8, hipMemcpyDeviceToHost, stream));
    uint32_t err = 0;
    if ((rc = trace_error(c, &err))) return rc;  // synchronises the stream
    if (err) return fail(c, RT_E_DEVICE, "rt_render: traversal exceeded its stack or step bound");
    float ms = 0.0f;
    RT_HIP(c, hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
    if (nrc) {
        RT_HIP(c, hipEventElapsedTime(&nf->stats->render_ms, c->ev[0], c->ev[4]));
        nf->stats->train_ms = ms - nf->stats->render_ms;
    }
    c->timing.last_call_ms = ms;
    c->timing.launches = launches;
    c->timing.iterations = (uint32_t)iter;
    c->timing.rays_traced = rays;
    c->last_op = 2;
    c->render_iters = (uint32_t)iter;
    return RT_OK;
}

int launch_ddgi_shade(rt_ctx* c, const rt_settings* s, const DdgiShade& b) {
    if (b.n == 0) return RT_OK;
    PTArgs a;
    memset(&a, 0, sizeof(a));
    a.geom = (const ShadeGeom*)c->shade_geom.ptr;
    a.nrm = (const ShadeNrm*)c->shade_nrm.ptr;
    a.mat = (const rt_material*)c->mat.ptr;
    a.flags = s->flags;
    a.seed_lo = (uint32_t)s->seed;
    a.seed_hi = (uint32_t)(s->seed >> 32);
    for (int k = 0; k < 3; ++k) {
        a.sky[k] = s->sky[k];
        a.sun_dir[k] = s->sun_dir[k];
        a.sun_rad[k] = s->sun_radiance[k];
    }
    a.sun_cos = s->sun_cos;
    a.sun_on = (s->flags & RT_RENDER_SUN) && s->sun_cos < 1.0f;
    a.nee = c->num_lights > 0 || a.sun_on;
    a.lights = (const LightTri*)c->lights.ptr;
    a.num_lights = c->num_lights;
    a.light_area = c->light_area;
    hipLaunchKernelGGL(k_ddgi_shade, dim3(div_up(b.n, PT_BLOCK)), dim3(PT_BLOCK), 0, c->stream, a, b);
    RT_HIP(c, hipGetLastError());
    return RT_OK;
}

}  // namespace rt
