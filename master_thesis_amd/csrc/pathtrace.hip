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
// order.  Every iteration is two kernels over the live slots:
//   extend  closest hit of every queued ray (trace.hip, device-resident count)
//   shade   the state machine above; a finished sample restarts the slot on
//           the next sample; a finished item takes the next item from a
//           global counter (the analogue of pixelCounter.fetch_add,
//           graphics_engine.cpp:74); surviving slots are compacted into the
//           next queue with one atomic per wave (ballot + popcount).
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

#include "rt_internal.hpp"
#include "rt_core.hpp"

namespace rt {
namespace {

constexpr int PT_BLOCK = 256;
constexpr uint32_t ADJ_MAX_DEPTH = 16;
constexpr uint32_t CULL_BIT = 0x10000u;
// cache mode: front-facing hits so far (bits 17-23), training record open (bit 24)
constexpr uint32_t FH_SHIFT = 17;
constexpr uint32_t FH_MASK = 0x7Fu;
constexpr uint32_t REC_BIT = 1u << 24;

// ctr slots
constexpr int C_ITEM = 0;   // next work item
constexpr int C_N0 = 1;     // live count, queue 0
constexpr int C_N1 = 2;     // live count, queue 1
constexpr int C_RAYS = 4;   // 64-bit count of traced rays (slots 4-5)
constexpr int C_QN = 6;     // cache queries queued since the last flush
constexpr int C_TN = 7;     // training records made
constexpr int C_QDROP = 8;  // queries dropped for want of room (stays 0: the host flushes in time)
constexpr int C_WORDS = 9;

struct PTArgs {
    const ShadeGeom* geom;
    const ShadeNrm* nrm;
    const rt_material* mat;
    const rt_camera* cams;  // V cameras in device memory
    uint32_t W, H, P;       // image size, pixels per view
    uint32_t V, VP;         // views, V * P
    uint32_t K, spc, items, S;
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
    // radiance cache (nrc_q == 0: off)
    uint32_t nrc_q, nrc_every, nrc_query;
    float* q_feat;      // q_cap * 9
    float* q_w;         // q_cap * 3: throughput before the query vertex
    uint32_t* q_idx;    // q_cap: partial-sum index
    uint32_t q_cap;
    float* t_feat;      // t_cap * 9
    float* t_tgt;       // t_cap * 3
    uint32_t t_cap;
    float* s_qf;        // 9 * S: training sample's features at its query vertex
    float* s_Tq;        // 3 * S: its throughput from the query vertex on
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
// constant sky plus the sun inside its cone.
__device__ __forceinline__ f3 miss_radiance(const PTArgs& a, f3 d) {
    f3 L = mk(0.0f, 0.0f, 0.0f);
    if (a.flags & RT_RENDER_SKY) L = mk(a.sky[0], a.sky[1], a.sky[2]);
    if ((a.flags & RT_RENDER_SUN) && dot(d, mk(a.sun_dir[0], a.sun_dir[1], a.sun_dir[2])) >= a.sun_cos)
        L = add(L, mk(a.sun_rad[0], a.sun_rad[1], a.sun_rad[2]));
    return L;
}

template <bool ADJ>
__global__ void __launch_bounds__(PT_BLOCK) k_pt_init(PTArgs a) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        a.ctr[C_ITEM] = a.S;
        a.ctr[C_N0] = a.S;
        a.ctr[C_N1] = 0;
        *(unsigned long long*)&a.ctr[C_RAYS] = 0ull;
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
__device__ __forceinline__ void scatter_grad(const PTArgs& a, uint32_t slot, uint32_t pix, f3 term) {
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
            atomicAdd(&a.dcolor[3 * m[k] + ch], w[ch] * g);
            pre = pre * f[k];
        }
    }
}

template <bool ADJ, bool NRC>
__global__ void __launch_bounds__(PT_BLOCK) k_pt_shade(PTArgs a, int cur) {
    const uint32_t n = a.ctr[cur ? C_N1 : C_N0];
    uint32_t* n_out = &a.ctr[cur ? C_N0 : C_N1];
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k == 0) *(unsigned long long*)&a.ctr[C_RAYS] += n;  // rays the extend step just traced
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    bool alive = false, need_item = false;
    uint32_t slot = 0;
    rt_ray nr;
    // cache mode: a query or a training record to enqueue
    bool push_q = false, push_t = false;
    float qf[9];
    f3 qw = mk(0.0f, 0.0f, 0.0f);
    uint32_t qidx = 0;
    if (k < n) {
        slot = a.q_slot[cur][k];
        const rt_ray r = a.q_ray[cur][k];
        const rt_hit h = a.hits[k];
        const uint32_t item = a.s_item[slot];
        const uint32_t vp = item % a.VP;   // view * P + pixel: index into images and partial sums
        const uint32_t chunk = item / a.VP;
        const uint32_t view = vp / a.P, pix = vp % a.P;
        uint32_t smp = a.s_smp[slot];
        uint32_t st = a.s_state[slot];
        uint32_t depth = st & 0xFFFFu;
        bool cull = (st & CULL_BIT) != 0;
        uint32_t fh = (st >> FH_SHIFT) & FH_MASK;
        bool rec = NRC && (st & REC_BIT) != 0;
        f3 T = mk(a.s_T[3 * slot], a.s_T[3 * slot + 1], a.s_T[3 * slot + 2]);
        f3 Tq = mk(1.0f, 1.0f, 1.0f);
        if (rec) Tq = mk(a.s_Tq[3 * slot], a.s_Tq[3 * slot + 1], a.s_Tq[3 * slot + 2]);
        f3 o = mk(r.ox, r.oy, r.oz), d = mk(r.dx, r.dy, r.dz);
        bool end = false, contrib = false;
        f3 term = mk(1.0f, 1.0f, 1.0f);
        if (h.tri == RT_MISS) {
            end = true;
            if (a.flags & (RT_RENDER_SKY | RT_RENDER_SUN)) {
                contrib = true;
                term = miss_radiance(a, d);
                T = mulv(T, term);
                if (rec) Tq = mulv(Tq, term);
            }
        } else {
            const f3 p = hit_point(o, d, h.t);
            const ShadeGeom g = a.geom[h.tri];
            const ShadeNrm nn = a.nrm[h.tri];
            // Triangle::getBarycentricCoords (triangle.cpp:47-55)
            const f3 v0 = mk(g.v0[0], g.v0[1], g.v0[2]);
            const f3 e0 = mk(g.e0[0], g.e0[1], g.e0[2]), e1 = mk(g.e1[0], g.e1[1], g.e1[2]);
            const f3 e2 = sub(p, v0);
            const float d20 = dot(e2, e0), d21 = dot(e2, e1);
            const float bv = (g.d11 * d20 - g.d01 * d21) / g.denom;
            const float bw = (g.d00 * d21 - g.d01 * d20) / g.denom;
            const float bu = 1.0f - bv - bw;
            // Scene::traceRay normal (scene.cpp:49-53)
            f3 nrm = mk(0.0f, 0.0f, 0.0f);
            nrm = add(nrm, scale(bu, mk(nn.n0[0], nn.n0[1], nn.n0[2])));
            nrm = add(nrm, scale(bv, mk(nn.n1[0], nn.n1[1], nn.n1[2])));
            nrm = add(nrm, scale(bw, mk(nn.n2[0], nn.n2[1], nn.n2[2])));
            nrm = normalize_ref(nrm);
            const float ndotd = dot(nrm, d);
            if (cull && ndotd > 0.0f) {
                o = add(p, scale(0.01f, d));
            } else {
                cull = false;
                const rt_material m = a.mat[g.mat];
                const uint32_t gsmp = a.sample_offset + chunk * a.spc + smp;
                const bool light_end = m.light_source && dot(nrm, d) <= 0.0f;
                if (NRC && fh == a.nrc_q && !light_end && depth + 1 < a.max_depth) {
                    const bool train = a.nrc_every != 0 && train_hash(pix, gsmp, view, a.seed_lo) % a.nrc_every == 0;
                    const float f9[9] = {p.x, p.y, p.z, nrm.x, nrm.y, nrm.z, d.x, d.y, d.z};
                    if (train) {
                        rec = true;
                        Tq = mk(1.0f, 1.0f, 1.0f);
                        for (int j = 0; j < 9; ++j) a.s_qf[9 * (size_t)slot + j] = f9[j];
                    } else if (a.nrc_query) {
                        push_q = true;
                        for (int j = 0; j < 9; ++j) qf[j] = f9[j];
                        qw = T;
                        qidx = chunk * a.VP + vp;
                        end = true;
                    }
                }
                ++fh;
                if (!push_q) {
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
                    if (rec) Tq = mulv(Tq, scale(cs, mk(m.color[0], m.color[1], m.color[2])));
                    if (ADJ) {
                        uint32_t nv = a.s_nv[slot];
                        if (nv < ADJ_MAX_DEPTH) {
                            a.s_amat[nv * a.S + slot] = g.mat;
                            a.s_acos[nv * a.S + slot] = cs;
                        }
                        a.s_nv[slot] = nv + 1;
                    }
                }
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
                if (ADJ) {
                    scatter_grad(a, slot, vp, term);
                } else {
                    a.s_sum[3 * slot] = a.s_sum[3 * slot] + T.x;
                    a.s_sum[3 * slot + 1] = a.s_sum[3 * slot + 1] + T.y;
                    a.s_sum[3 * slot + 2] = a.s_sum[3 * slot + 2] + T.z;
                }
            }
            if (rec) {
                push_t = true;
                for (int j = 0; j < 9; ++j) qf[j] = a.s_qf[9 * (size_t)slot + j];
                qw = contrib ? Tq : mk(0.0f, 0.0f, 0.0f);
            }
            ++smp;
            if (smp < a.spc) {
                a.s_smp[slot] = smp;
                start_sample(a, slot, vp, nr, ADJ);
                alive = true;
            } else {
                if (NRC) {
                    // cache predictions land in the same sums
                    const size_t q = (size_t)chunk * a.VP + vp;
                    atomicAdd(&a.partial[3 * q], a.s_sum[3 * slot]);
                    atomicAdd(&a.partial[3 * q + 1], a.s_sum[3 * slot + 1]);
                    atomicAdd(&a.partial[3 * q + 2], a.s_sum[3 * slot + 2]);
                } else if (!ADJ) {
                    const size_t q = (size_t)chunk * a.VP + vp;
                    a.partial[3 * q] = a.s_sum[3 * slot];
                    a.partial[3 * q + 1] = a.s_sum[3 * slot + 1];
                    a.partial[3 * q + 2] = a.s_sum[3 * slot + 2];
                }
                need_item = true;
            }
        } else {
            a.s_state[slot] = depth | (cull ? CULL_BIT : 0u) | (NRC ? (fh << FH_SHIFT) | (rec ? REC_BIT : 0u) : 0u);
            a.s_T[3 * slot] = T.x; a.s_T[3 * slot + 1] = T.y; a.s_T[3 * slot + 2] = T.z;
            if (rec) {
                a.s_Tq[3 * slot] = Tq.x; a.s_Tq[3 * slot + 1] = Tq.y; a.s_Tq[3 * slot + 2] = Tq.z;
            }
            nr = make_ray(o, d);
            alive = true;
        }
    }
    if (NRC) {
        // enqueue cache queries and training records: one atomic per wave each
        const uint64_t qb = __ballot(push_q);
        if (qb) {
            const int leader = __ffsll((unsigned long long)qb) - 1;
            uint32_t base = 0;
            if ((int)lane == leader) base = atomicAdd(&a.ctr[C_QN], (uint32_t)__popcll(qb));
            base = __shfl(base, leader);
            if (push_q) {
                const uint32_t pos = base + (uint32_t)__popcll(qb & lt);
                if (pos < a.q_cap) {
                    for (int j = 0; j < 9; ++j) a.q_feat[9 * (size_t)pos + j] = qf[j];
                    a.q_w[3 * (size_t)pos] = qw.x; a.q_w[3 * (size_t)pos + 1] = qw.y; a.q_w[3 * (size_t)pos + 2] = qw.z;
                    a.q_idx[pos] = qidx;
                } else {
                    atomicAdd(&a.ctr[C_QDROP], 1u);
                }
            }
        }
        const uint64_t tb = __ballot(push_t);
        if (tb) {
            const int leader = __ffsll((unsigned long long)tb) - 1;
            uint32_t base = 0;
            if ((int)lane == leader) base = atomicAdd(&a.ctr[C_TN], (uint32_t)__popcll(tb));
            base = __shfl(base, leader);
            if (push_t) {
                const uint32_t pos = base + (uint32_t)__popcll(tb & lt);
                if (pos < a.t_cap) {  // records past the capacity are dropped
                    for (int j = 0; j < 9; ++j) a.t_feat[9 * (size_t)pos + j] = qf[j];
                    a.t_tgt[3 * (size_t)pos] = qw.x; a.t_tgt[3 * (size_t)pos + 1] = qw.y; a.t_tgt[3 * (size_t)pos + 2] = qw.z;
                }
            }
        }
    }
    // next work item for slots whose item finished: one atomic per wave
    const uint64_t want = __ballot(need_item);
    if (want) {
        const int leader = __ffsll((unsigned long long)want) - 1;
        uint32_t base = 0;
        if ((int)lane == leader) base = atomicAdd(&a.ctr[C_ITEM], (uint32_t)__popcll(want));
        base = __shfl(base, leader);
        if (need_item) {
            const uint32_t item = base + (uint32_t)__popcll(want & lt);
            if (item < a.items) {
                start_item(a, slot, item, nr, ADJ);
                alive = true;
            }
        }
    }
    // compact the survivors into the next queue: one atomic per wave
    const uint64_t live = __ballot(alive);
    if (live) {
        const int leader = __ffsll((unsigned long long)live) - 1;
        uint32_t base = 0;
        if ((int)lane == leader) base = atomicAdd(n_out, (uint32_t)__popcll(live));
        base = __shfl(base, leader);
        if (alive) {
            const uint32_t pos = base + (uint32_t)__popcll(live & lt);
            a.q_slot[cur ^ 1][pos] = slot;
            a.q_ray[cur ^ 1][pos] = nr;
        }
    }
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

__global__ void k_zero(float* p, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0f;
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
    // samples of a pixel are split into K chunks (K | N) when there are too
    // few pixels to fill the device; deterministic mode keeps one chunk so
    // every pixel's sum has the reference's sample order.
    const uint64_t target_slots = 1ull << 20;
    uint32_t K = 1;
    if (!(s->flags & RT_RENDER_DETERMINISTIC) && !adj) {
        uint64_t want = (target_slots + P - 1) / P;
        if (want > N) want = N;
        for (uint32_t k = (uint32_t)want; k >= 1; --k)
            if (N % k == 0) { K = k; break; }
    }
    const uint64_t items = P * K;
    if (items > 0xFFFFFFFFull) return fail(c, RT_E_UNSUPPORTED, "rt_render: pixels * chunks above 2^32");
    if (nrc && 3 * items > 0xFFFFFFFFull) return fail(c, RT_E_UNSUPPORTED, "rt_render_nrc: pixels * chunks above 2^32 / 3");
    const uint32_t S = (uint32_t)(items < target_slots ? items : target_slots);
    int rc;
    const size_t szS = S;
    // cache mode: queries flushed every 8 iterations (at most S per
    // iteration), training records capped at 4 frames' worth of steps
    const size_t q_cap = nrc ? 8 * szS : 0;
    const size_t t_cap = nrc ? (size_t)4 * nf->opt.batch * nf->opt.max_steps : 0;
    if (nrc && ((rc = ensure(c, c->pt_nrc_query, q_cap * (36 + 12 + 4))) ||
                (rc = ensure(c, c->pt_nrc_train, (t_cap ? t_cap : 1) * (36 + 12)))))
        return rc;
    if ((rc = ensure(c, c->pt_state, szS * (4 * 3 + 12 * 2) + (adj ? szS * (8 * ADJ_MAX_DEPTH + 4) : 0) +
                                         (nrc ? szS * 48 : 0))) ||
        (rc = ensure(c, c->pt_queue[0], 4 * szS + 16 + sizeof(rt_ray) * szS)) ||
        (rc = ensure(c, c->pt_queue[1], 4 * szS + 16 + sizeof(rt_ray) * szS)) ||
        (rc = ensure(c, c->hits, szS * sizeof(rt_hit))) || (rc = ensure(c, c->pt_ctr, 64)) ||
        (rc = ensure(c, c->pt_partial, adj ? 4 : 12 * items)) || (rc = ensure(c, c->pt_cams, sizeof(rt_camera) * V)))
        return rc;
    RT_HIP(c, hipMemcpyAsync(c->pt_cams.ptr, cams, sizeof(rt_camera) * V, hipMemcpyHostToDevice, c->stream));
    PTArgs a;
    a.geom = (const ShadeGeom*)c->shade_geom.ptr;
    a.nrm = (const ShadeNrm*)c->shade_nrm.ptr;
    a.mat = (const rt_material*)c->mat.ptr;
    a.cams = (const rt_camera*)c->pt_cams.ptr;
    a.W = cam->width; a.H = cam->height; a.P = (uint32_t)P1;
    a.V = V; a.VP = (uint32_t)P;
    a.K = K; a.spc = N / K; a.items = (uint32_t)items; a.S = S;
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
    a.s_qf = nrc ? (float*)st : nullptr; st += nrc ? 36 * szS : 0;
    a.s_Tq = nrc ? (float*)st : nullptr;
    a.nrc_q = nrc ? nf->opt.query_vertex : 0;
    a.nrc_every = nrc && !(nf->opt.flags & RT_NRC_NO_RECORDS) ? nf->opt.train_every : 0;
    a.nrc_query = nrc && !(nf->opt.flags & RT_NRC_NO_QUERY);
    a.q_feat = nrc ? (float*)c->pt_nrc_query.ptr : nullptr;
    a.q_w = nrc ? a.q_feat + 9 * q_cap : nullptr;
    a.q_idx = nrc ? (uint32_t*)(a.q_w + 3 * q_cap) : nullptr;
    a.q_cap = (uint32_t)q_cap;
    a.t_feat = nrc ? (float*)c->pt_nrc_train.ptr : nullptr;
    a.t_tgt = nrc ? a.t_feat + 9 * t_cap : nullptr;
    a.t_cap = (uint32_t)t_cap;
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

    hipStream_t stream = c->stream;
    RT_HIP(c, hipEventRecord(c->ev[0], stream));
    if (adj) hipLaunchKernelGGL(k_zero, dim3(div_up(3ull * c->M, 256)), dim3(256), 0, stream, d_dcolor, 3u * c->M);
    if (nrc) hipLaunchKernelGGL(k_zero, dim3(div_up(3ull * items, 256)), dim3(256), 0, stream, a.partial, (uint32_t)(3 * items));
    if (adj) hipLaunchKernelGGL(k_pt_init<true>, dim3(div_up(S, PT_BLOCK)), dim3(PT_BLOCK), 0, stream, a);
    else hipLaunchKernelGGL(k_pt_init<false>, dim3(div_up(S, PT_BLOCK)), dim3(PT_BLOCK), 0, stream, a);
    RT_HIP(c, hipGetLastError());
    // upper bound on iterations: every slot runs ceil(items / S) items of
    // spc samples of at most max_depth extends
    const uint64_t bound = (uint64_t)div_up(items, S) * a.spc * s->max_depth + 2;
    int cur = 0;
    uint64_t iter = 0;
    uint32_t live = S;
    uint32_t launches = 1;
    while (live > 0) {
        for (int rep = 0; rep < 8 && iter < bound; ++rep, ++iter) {
            if ((rc = launch_trace(c, a.q_ray[cur], S, a.hits, false, &a.ctr[cur ? C_N1 : C_N0]))) return rc;
            RT_HIP(c, hipMemsetAsync(&a.ctr[cur ? C_N0 : C_N1], 0, 4, stream));
            if (adj) hipLaunchKernelGGL((k_pt_shade<true, false>), dim3(div_up(S, PT_BLOCK)), dim3(PT_BLOCK), 0, stream, a, cur);
            else if (nrc) hipLaunchKernelGGL((k_pt_shade<false, true>), dim3(div_up(S, PT_BLOCK)), dim3(PT_BLOCK), 0, stream, a, cur);
            else hipLaunchKernelGGL((k_pt_shade<false, false>), dim3(div_up(S, PT_BLOCK)), dim3(PT_BLOCK), 0, stream, a, cur);
            RT_HIP(c, hipGetLastError());
            launches += 2;
            cur ^= 1;
        }
        uint32_t hc[C_WORDS];
        RT_HIP(c, hipMemcpyAsync(hc, a.ctr, sizeof(hc), hipMemcpyDeviceToHost, stream));
        RT_HIP(c, hipStreamSynchronize(stream));
        live = hc[cur ? C_N1 : C_N0];
        if (nrc && hc[C_QDROP] != 0) return fail(c, RT_E_DEVICE, "rt_render_nrc: cache query buffer overflow");
        if (nrc && hc[C_QN] > 0) {
            // the queued samples' cache predictions, added to their pixels
            if ((rc = nrc_infer_scatter(nf->cache, a.q_feat, a.q_w, a.q_idx, hc[C_QN], a.partial))) return rc;
            RT_HIP(c, hipMemsetAsync(&a.ctr[C_QN], 0, 4, stream));
            nf->stats->queries += hc[C_QN];
            ++launches;
        }
        if (live > 0 && iter >= bound) return fail(c, RT_E_DEVICE, "rt_render: paths still live past the iteration bound");
    }
    if (!adj) {
        hipLaunchKernelGGL(k_pt_finish, dim3(div_up(P, PT_BLOCK)), dim3(PT_BLOCK), 0, stream, a);
        RT_HIP(c, hipGetLastError());
        ++launches;
    }
    if (nrc) {
        RT_HIP(c, hipEventRecord(c->ev[4], stream));
        uint32_t tn = 0;
        RT_HIP(c, hipMemcpyAsync(&tn, &a.ctr[C_TN], 4, hipMemcpyDeviceToHost, stream));
        RT_HIP(c, hipStreamSynchronize(stream));
        nf->stats->records = tn;
        if (!(nf->opt.flags & RT_NRC_NO_TRAIN)) {
            const uint32_t avail = tn < a.t_cap ? tn : a.t_cap;
            if ((rc = nrc_train_records(nf->cache, a.t_feat, a.t_tgt, avail, nf->opt.batch, nf->opt.max_steps,
                                        &nf->stats->steps, &nf->stats->loss)))
                return rc;
            launches += 2 * nf->stats->steps;
        }
    }
    RT_HIP(c, hipEventRecord(c->ev[1], stream));
    unsigned long long rays = 0;
    RT_HIP(c, hipMemcpyAsync(&rays, &a.ctr[C_RAYS], 8, hipMemcpyDeviceToHost, stream));
    RT_HIP(c, hipStreamSynchronize(stream));
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

}  // namespace rt
