// bdpt.hip — the bidirectional path tracer, wavefront by depth.
//
// Replaces BidirectionalPathTracer::renderPixel / traceSinglePath /
// getRandomLightSourcePoint (software_renderer/graphic/
// bidirectional_path_tracer.cpp:20-168) driven per pixel by
// GraphicsEngine::render (graphics_engine.cpp:59-97).
//
// Estimator (the reference's, restated): per sample a vision path from the
// fixed primary ray: at depth i < visionJumpCount a miss ends it; while
// culling is on, a back-facing hit steps past (o = p + 0.01 d) and leaves
// slot i empty; otherwise the hit is recorded at slot i (position, normal,
// colour so far) after xi0 picks diffuse / reflect / refract, a diffuse
// bounce multiplying the colour by colour * (n . d_new), and a light hit from
// the front ends the path.  If the last recorded vertex is such a light hit,
// the sample adds its colour.  Otherwise a light path starts at a random
// point of a random light object (object uniformly, then one of its
// triangles uniformly, sqrt-warped barycentrics, interpolated normal): slot
// 0 holds that point with the light's colour, the ray leaves it along a
// uniform hemisphere direction, and at depth i (1 <= i < lightJumpCount) hits
// are recorded without culling, a diffuse bounce multiplying the colour
// (starting from 1) by colour * (n . d_incoming).  Every pair (vi < vision
// depth, li < light depth, vi + li <= maxDepth) whose points offset 0.01
// along their normals see each other (Scene::isOccluded) adds
// vision colour * light colour * (dir . n_vision); if any pair did, the sample
// adds the sum / maxDepth.  Pixel = 2 pi / N * sum over samples, in order.
//
// Scheduling: a batch holds whole pixels (all their samples).  Vision and
// light paths advance one depth per step over a compacted queue (trace.hip
// for the closest hits); all pairs of a sample are then tested in one
// segment-occlusion launch; a combine kernel sums each sample's pairs in the
// reference's loop order and a pixel kernel sums its samples in order.
#include <math.h>

#include "rt_internal.hpp"
#include "rt_core.hpp"

namespace rt {
namespace {

constexpr int BD_BLOCK = 256;
constexpr uint32_t LIGHT_SIDE = 0x40000000u;  // Philox counter word 3 of the light side
constexpr uint32_t V_VALID = 1u, V_DIFFUSE = 2u, V_LIGHTHIT = 4u;

// C_* slots of the batch counters
constexpr int BC_N0 = 0, BC_N1 = 1;

struct BdVert {
    float p[3], n[3], c[3];
    uint32_t flags;
};

struct BdArgs {
    const ShadeGeom* geom;
    const ShadeNrm* nrm;
    const rt_material* mat;
    const float* pos;
    const uint32_t* idx;
    const float* vnrm;          // per-vertex world normals (light points)
    const uint32_t* lobj;       // {first, count} per light object
    const uint32_t* lobj_tris;
    uint32_t num_lobj;
    rt_camera cam;
    uint32_t W, H;
    uint32_t p0, np, spp, B;    // batch: pixels [p0, p0 + np), B = np * spp samples
    uint32_t VJ, LJ, maxd;
    uint32_t sample_offset, seed_lo, seed_hi;
    double scale;
    BdVert* V;                  // B * VJ
    BdVert* Lv;                 // B * LJ
    uint32_t* vdepth;           // B
    uint32_t* ldepth;           // B
    uint32_t* sflags;           // B: bit 0 culling on, bit 1 vision path ended on a light
    float* T;                   // 3 * B
    uint32_t* q_id[2];
    rt_ray* q_ray[2];
    rt_hit* hits;
    uint32_t* ctr;
    float* segs;                // B * VJ * LJ * 6
    uint8_t* occ;               // B * VJ * LJ
    float* sval;                // 3 * B
    float* rgb;
    uint8_t* rgb8;
};

__device__ __forceinline__ rt_ray make_ray(f3 o, f3 d) {
    rt_ray r;
    r.ox = o.x; r.oy = o.y; r.oz = o.z; r.tmin = 0.0f;
    r.dx = d.x; r.dy = d.y; r.dz = d.z; r.tmax = INFINITY;
    return r;
}

__device__ __forceinline__ void put_vert(BdVert& v, f3 p, f3 n, f3 c, uint32_t flags) {
    v.p[0] = p.x; v.p[1] = p.y; v.p[2] = p.z;
    v.n[0] = n.x; v.n[1] = n.y; v.n[2] = n.z;
    v.c[0] = c.x; v.c[1] = c.y; v.c[2] = c.z;
    v.flags = flags;
}

// append the live lanes to queue q: one atomic per workgroup; every thread
// of the workgroup calls it
__device__ __forceinline__ void enqueue(const BdArgs& a, int q, bool alive, uint32_t s, const rt_ray& r) {
    __shared__ uint32_t lds_alloc[BD_BLOCK / 64 + 1];
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const bool f[1] = {alive};
    uint32_t* const c[1] = {&a.ctr[q ? BC_N1 : BC_N0]};
    uint32_t pos[1];
    block_alloc<BD_BLOCK, 1>(f, c, lds_alloc, lane, lt, pos);
    if (alive) {
        a.q_id[q][pos[0]] = s;
        a.q_ray[q][pos[0]] = r;
    }
}

__global__ void __launch_bounds__(BD_BLOCK) k_bd_init(BdArgs a) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    rt_ray r;
    if (s < a.B) {
        const uint32_t pix = a.p0 + s / a.spp;
        const f3 d = primary_dir(a.cam.view_inv, a.cam.proj_inv, pix % a.W, pix / a.W, a.W, a.H);
        r = make_ray(mk(a.cam.origin[0], a.cam.origin[1], a.cam.origin[2]), d);
        a.vdepth[s] = 0;
        a.ldepth[s] = 0;
        a.sflags[s] = 1u;  // culling on
        a.T[3 * s] = 1.0f; a.T[3 * s + 1] = 1.0f; a.T[3 * s + 2] = 1.0f;
        for (uint32_t i = 0; i < a.VJ; ++i) a.V[(size_t)s * a.VJ + i].flags = 0u;
    }
    enqueue(a, 0, s < a.B, s, r);
}

// Depth i of the vision (LIGHT = false) or light (LIGHT = true) paths.
template <bool LIGHT>
__global__ void __launch_bounds__(BD_BLOCK) k_bd_step(BdArgs a, uint32_t i, int cur) {
    const uint32_t n = a.ctr[cur ? BC_N1 : BC_N0];
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    bool alive = false;
    uint32_t s = 0;
    rt_ray nr;
    if (k < n) {
        s = a.q_id[cur][k];
        const rt_ray r = a.q_ray[cur][k];
        const rt_hit h = a.hits[k];
        const uint32_t pix = a.p0 + s / a.spp;
        const uint32_t gsmp = a.sample_offset + s % a.spp;
        if (h.tri != RT_MISS) {
            f3 o = mk(r.ox, r.oy, r.oz), d = mk(r.dx, r.dy, r.dz);
            const f3 p = hit_point(o, d, h.t);
            const ShadeGeom g = a.geom[h.tri];
            const ShadeNrm nn = a.nrm[h.tri];
            // Triangle::getBarycentricCoords and Scene::traceRay's normal
            const f3 v0 = mk(g.v0[0], g.v0[1], g.v0[2]);
            const f3 e0 = mk(g.e0[0], g.e0[1], g.e0[2]), e1 = mk(g.e1[0], g.e1[1], g.e1[2]);
            const f3 e2 = sub(p, v0);
            const float d20 = dot(e2, e0), d21 = dot(e2, e1);
            const float bv = (g.d11 * d20 - g.d01 * d21) / g.denom;
            const float bw = (g.d00 * d21 - g.d01 * d20) / g.denom;
            const float bu = 1.0f - bv - bw;
            f3 nrm = mk(0.0f, 0.0f, 0.0f);
            nrm = add(nrm, scale(bu, mk(nn.n0[0], nn.n0[1], nn.n0[2])));
            nrm = add(nrm, scale(bv, mk(nn.n1[0], nn.n1[1], nn.n1[2])));
            nrm = add(nrm, scale(bw, mk(nn.n2[0], nn.n2[1], nn.n2[2])));
            nrm = normalize_ref(nrm);
            const float ndotd = dot(nrm, d);
            uint32_t sf = a.sflags[s];
            if (!LIGHT && (sf & 1u) && ndotd > 0.0f) {
                o = add(p, scale(0.01f, d));  // back face while culling: slot i stays empty
                nr = make_ray(o, d);
                alive = true;
            } else {
                if (!LIGHT) a.sflags[s] = sf & ~1u;
                const rt_material m = a.mat[g.mat];
                u4 c4;
                c4.x = pix; c4.y = gsmp;
                c4.z = LIGHT ? i + 1u : i;
                c4.w = LIGHT ? LIGHT_SIDE : 0u;
                const u4 rnd = philox4x32_10(c4, a.seed_lo, a.seed_hi);
                const float xi0 = u01(rnd.x);
                const f3 prev = d;
                const bool diffuse = xi0 <= m.diffuse_threshold;
                if (diffuse) d = hemisphere_dir(u01(rnd.y), u01(rnd.z), nrm);
                else if (xi0 <= m.reflect_threshold) d = reflect_ref(d, nrm);
                else if (xi0 <= m.transparent_threshold) d = custom_refract_ref(d, nrm, m.refraction_index);
                o = add(p, scale(0.01f, d));
                f3 T = mk(a.T[3 * s], a.T[3 * s + 1], a.T[3 * s + 2]);
                if (diffuse) {
                    const float cs = dot(nrm, LIGHT ? prev : d);
                    T = mulv(T, scale(cs, mk(m.color[0], m.color[1], m.color[2])));
                    a.T[3 * s] = T.x; a.T[3 * s + 1] = T.y; a.T[3 * s + 2] = T.z;
                }
                const bool light_hit = m.light_source && dot(nrm, prev) <= 0.0f;
                const uint32_t flags = V_VALID | (diffuse ? V_DIFFUSE : 0u) | (light_hit ? V_LIGHTHIT : 0u);
                if (LIGHT) {
                    put_vert(a.Lv[(size_t)s * a.LJ + i], p, nrm, T, flags);
                    a.ldepth[s] = i + 1;
                } else {
                    put_vert(a.V[(size_t)s * a.VJ + i], p, nrm, T, flags);
                    a.vdepth[s] = i + 1;
                }
                if (light_hit) {
                    if (!LIGHT) a.sflags[s] = (sf & ~1u) | 2u;
                } else {
                    nr = make_ray(o, d);
                    alive = true;
                }
            }
        }
    }
    enqueue(a, cur ^ 1, alive, s, nr);
}

// BidirectionalPathTracer::getRandomLightSourcePoint (:139-168) and the light
// path's start (:44-50) for samples whose vision path did not end on a light.
__global__ void __launch_bounds__(BD_BLOCK) k_bd_light_init(BdArgs a, int q) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    bool alive = false;
    rt_ray r;
    if (s < a.B && a.vdepth[s] > 0 && !(a.sflags[s] & 2u) && a.num_lobj > 0) {
        const uint32_t pix = a.p0 + s / a.spp;
        const uint32_t gsmp = a.sample_offset + s % a.spp;
        u4 c4;
        c4.x = pix; c4.y = gsmp; c4.z = 0u; c4.w = LIGHT_SIDE;
        const u4 rnd = philox4x32_10(c4, a.seed_lo, a.seed_hi);
        uint32_t lo = (uint32_t)(u01(rnd.x) * (float)a.num_lobj);
        lo = min(lo, a.num_lobj - 1u);
        const float sq = sqrtf(u01(rnd.y)), r2 = u01(rnd.z);
        const float b0 = 1.0f - sq, b1 = sq * (1.0f - r2), b2 = sq * r2;
        const uint32_t first = a.lobj[2 * lo], count = a.lobj[2 * lo + 1];
        uint32_t ti = (uint32_t)(u01(rnd.w) * (float)count);
        ti = min(ti, count - 1u);
        const uint32_t t = a.lobj_tris[first + ti];
        const uint32_t i0 = a.idx[3 * t], i1 = a.idx[3 * t + 1], i2 = a.idx[3 * t + 2];
        const f3 p0 = mk(a.pos[3 * i0], a.pos[3 * i0 + 1], a.pos[3 * i0 + 2]);
        const f3 p1 = mk(a.pos[3 * i1], a.pos[3 * i1 + 1], a.pos[3 * i1 + 2]);
        const f3 p2 = mk(a.pos[3 * i2], a.pos[3 * i2 + 1], a.pos[3 * i2 + 2]);
        const f3 lp = add(add(scale(b0, p0), scale(b1, p1)), scale(b2, p2));
        f3 ln = mk(0.0f, 0.0f, 0.0f);
        ln = add(ln, scale(b0, mk(a.vnrm[3 * i0], a.vnrm[3 * i0 + 1], a.vnrm[3 * i0 + 2])));
        ln = add(ln, scale(b1, mk(a.vnrm[3 * i1], a.vnrm[3 * i1 + 1], a.vnrm[3 * i1 + 2])));
        ln = add(ln, scale(b2, mk(a.vnrm[3 * i2], a.vnrm[3 * i2 + 1], a.vnrm[3 * i2 + 2])));
        ln = normalize_ref(ln);
        const rt_material lm = a.mat[a.geom[t].mat];
        put_vert(a.Lv[(size_t)s * a.LJ], lp, ln, mk(lm.color[0], lm.color[1], lm.color[2]), V_VALID | V_DIFFUSE);
        c4.z = 1u;
        const u4 rd = philox4x32_10(c4, a.seed_lo, a.seed_hi);
        r = make_ray(lp, hemisphere_dir(u01(rd.y), u01(rd.z), ln));
        a.T[3 * s] = 1.0f; a.T[3 * s + 1] = 1.0f; a.T[3 * s + 2] = 1.0f;
        alive = a.LJ > 1;
    }
    enqueue(a, q, alive, s, r);
}

__device__ __forceinline__ bool pair_valid(const BdArgs& a, uint32_t s, uint32_t vi, uint32_t li) {
    return vi < a.vdepth[s] && li < a.ldepth[s] && vi + li <= a.maxd && !(a.sflags[s] & 2u) &&
           (a.V[(size_t)s * a.VJ + vi].flags & V_VALID);
}

__device__ __forceinline__ void pair_points(const BdArgs& a, uint32_t s, uint32_t vi, uint32_t li, f3& st, f3& en) {
    const BdVert& v = a.V[(size_t)s * a.VJ + vi];
    const BdVert& l = a.Lv[(size_t)s * a.LJ + li];
    st = add(mk(v.p[0], v.p[1], v.p[2]), scale(0.01f, mk(v.n[0], v.n[1], v.n[2])));
    en = add(mk(l.p[0], l.p[1], l.p[2]), scale(0.01f, mk(l.n[0], l.n[1], l.n[2])));
}

// one segment per (sample, vi, li); pairs that are not tested get a zero-length segment
__global__ void __launch_bounds__(BD_BLOCK) k_bd_segments(BdArgs a) {
    const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t per = (uint64_t)a.VJ * a.LJ;
    if (j >= per * a.B) return;
    const uint32_t s = (uint32_t)(j / per), vi = (uint32_t)((j % per) / a.LJ), li = (uint32_t)(j % a.LJ);
    f3 st = mk(0.0f, 0.0f, 0.0f), en = st;
    if (pair_valid(a, s, vi, li)) pair_points(a, s, vi, li, st, en);
    float* g = a.segs + 6 * j;
    g[0] = st.x; g[1] = st.y; g[2] = st.z;
    g[3] = en.x; g[4] = en.y; g[5] = en.z;
}

__global__ void __launch_bounds__(BD_BLOCK) k_bd_combine(BdArgs a) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= a.B) return;
    f3 val = mk(0.0f, 0.0f, 0.0f);
    const uint32_t vd = a.vdepth[s];
    if (vd > 0 && (a.sflags[s] & 2u)) {
        const BdVert& v = a.V[(size_t)s * a.VJ + vd - 1];
        val = mk(v.c[0], v.c[1], v.c[2]);
    } else if (vd > 0) {
        f3 color = mk(0.0f, 0.0f, 0.0f);
        bool done = false;
        for (uint32_t vi = 0; vi < vd; ++vi)
            for (uint32_t li = 0; li < a.ldepth[s]; ++li) {
                if (!pair_valid(a, s, vi, li)) continue;
                if (a.occ[((uint64_t)s * a.VJ + vi) * a.LJ + li]) continue;
                f3 st, en;
                pair_points(a, s, vi, li, st, en);
                const f3 dir = normalize_ref(sub(en, st));
                const BdVert& v = a.V[(size_t)s * a.VJ + vi];
                const BdVert& l = a.Lv[(size_t)s * a.LJ + li];
                const float cs = dot(dir, mk(v.n[0], v.n[1], v.n[2]));
                color = add(color, scale(cs, mulv(mk(v.c[0], v.c[1], v.c[2]), mk(l.c[0], l.c[1], l.c[2]))));
                done = true;
            }
        if (done) val = scale(1.0f / (float)a.maxd, color);
    }
    a.sval[3 * s] = val.x; a.sval[3 * s + 1] = val.y; a.sval[3 * s + 2] = val.z;
}

__global__ void __launch_bounds__(BD_BLOCK) k_bd_pixels(BdArgs a) {
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= a.np) return;
    float sm[3] = {0.0f, 0.0f, 0.0f};
    for (uint32_t j = 0; j < a.spp; ++j)
        for (int ch = 0; ch < 3; ++ch) sm[ch] = sm[ch] + a.sval[3 * ((size_t)q * a.spp + j) + ch];
    const size_t p = a.p0 + q;
    for (int ch = 0; ch < 3; ++ch) {
        // finalColor *= 2.0f * M_PI * (1.0f / float(raysPerPixel))  (double scale)
        const float v = (float)((double)sm[ch] * a.scale);
        if (a.rgb) a.rgb[3 * p + ch] = v;
        if (a.rgb8) a.rgb8[3 * p + ch] = (uint8_t)(int)(fminf(fmaxf(v, 0.0f), 1.0f) * 255.0f);
    }
}

inline uint32_t div_up(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

}  // namespace

int launch_bdpt(rt_ctx* c, const rt_camera* cam, const rt_bdpt_settings* s, float* d_rgb, uint8_t* d_rgb8) {
    const uint64_t P = (uint64_t)cam->width * cam->height;
    const uint32_t VJ = s->vision_jumps, LJ = s->light_jumps;
    // batch: whole pixels, at most 2^20 samples and 2^26 vertex pairs
    uint64_t bmax = (1ull << 26) / ((uint64_t)VJ * LJ);
    if (bmax > (1ull << 20)) bmax = 1ull << 20;
    uint64_t np = bmax / s->spp;
    if (np < 1) np = 1;
    if (np > P) np = P;
    const uint64_t B = np * s->spp;
    int rc;
    if ((rc = ensure(c, c->bd_vert, sizeof(BdVert) * B * (VJ + LJ))) ||
        (rc = ensure(c, c->bd_state, B * (4 + 4 + 4 + 12 + 12))) ||
        (rc = ensure(c, c->bd_queue[0], B * (4 + sizeof(rt_ray)) + 16)) ||
        (rc = ensure(c, c->bd_queue[1], B * (4 + sizeof(rt_ray)) + 16)) || (rc = ensure(c, c->hits, B * sizeof(rt_hit))) ||
        (rc = ensure(c, c->bd_seg, 24 * B * VJ * LJ)) || (rc = ensure(c, c->bd_occ, B * VJ * LJ)) ||
        (rc = ensure(c, c->pt_ctr, 64)))
        return rc;
    BdArgs a;
    a.geom = (const ShadeGeom*)c->shade_geom.ptr;
    a.nrm = (const ShadeNrm*)c->shade_nrm.ptr;
    a.mat = (const rt_material*)c->mat.ptr;
    a.pos = (const float*)c->pos.ptr;
    a.idx = (const uint32_t*)c->idx.ptr;
    a.vnrm = (const float*)c->vnrm.ptr;
    a.lobj = (const uint32_t*)c->light_objs.ptr;
    a.lobj_tris = (const uint32_t*)c->light_obj_tris.ptr;
    a.num_lobj = c->num_light_objs;
    a.cam = *cam;
    a.W = cam->width; a.H = cam->height;
    a.spp = s->spp;
    a.VJ = VJ; a.LJ = LJ; a.maxd = s->max_depth;
    a.sample_offset = s->sample_offset;
    a.seed_lo = (uint32_t)s->seed; a.seed_hi = (uint32_t)(s->seed >> 32);
    a.scale = 2.0 * M_PI * (double)(1.0f / (float)s->spp);
    a.V = (BdVert*)c->bd_vert.ptr;
    a.Lv = a.V + B * VJ;
    char* st = (char*)c->bd_state.ptr;
    a.vdepth = (uint32_t*)st; st += 4 * B;
    a.ldepth = (uint32_t*)st; st += 4 * B;
    a.sflags = (uint32_t*)st; st += 4 * B;
    a.T = (float*)st; st += 12 * B;
    a.sval = (float*)st;
    for (int q = 0; q < 2; ++q) {
        a.q_id[q] = (uint32_t*)c->bd_queue[q].ptr;
        a.q_ray[q] = (rt_ray*)((char*)c->bd_queue[q].ptr + ((4 * B + 15) & ~(uint64_t)15));
    }
    a.hits = (rt_hit*)c->hits.ptr;
    a.ctr = (uint32_t*)c->pt_ctr.ptr;
    a.segs = (float*)c->bd_seg.ptr;
    a.occ = (uint8_t*)c->bd_occ.ptr;
    a.rgb = d_rgb;
    a.rgb8 = d_rgb8;

    hipStream_t stream = c->stream;
    RT_HIP(c, hipEventRecord(c->ev[0], stream));
    uint32_t launches = 0;
    for (uint64_t p0 = 0; p0 < P; p0 += np) {
        a.p0 = (uint32_t)p0;
        a.np = (uint32_t)(P - p0 < np ? P - p0 : np);
        a.B = a.np * a.spp;
        const uint32_t grid = div_up(a.B, BD_BLOCK);
        RT_HIP(c, hipMemsetAsync(a.ctr, 0, 8, stream));
        hipLaunchKernelGGL(k_bd_init, dim3(grid), dim3(BD_BLOCK), 0, stream, a);
        int cur = 0;
        for (uint32_t i = 0; i < VJ; ++i) {
            // the trace clears the queue counter the step appends to
            if ((rc = launch_trace(c, a.q_ray[cur], a.B, a.hits, false, &a.ctr[cur ? BC_N1 : BC_N0],
                                   &a.ctr[cur ? BC_N0 : BC_N1])))
                return rc;
            hipLaunchKernelGGL(k_bd_step<false>, dim3(grid), dim3(BD_BLOCK), 0, stream, a, i, cur);
            cur ^= 1;
            launches += 2;
        }
        RT_HIP(c, hipMemsetAsync(&a.ctr[cur ? BC_N1 : BC_N0], 0, 4, stream));
        hipLaunchKernelGGL(k_bd_light_init, dim3(grid), dim3(BD_BLOCK), 0, stream, a, cur);
        for (uint32_t i = 1; i < LJ; ++i) {
            if ((rc = launch_trace(c, a.q_ray[cur], a.B, a.hits, false, &a.ctr[cur ? BC_N1 : BC_N0],
                                   &a.ctr[cur ? BC_N0 : BC_N1])))
                return rc;
            hipLaunchKernelGGL(k_bd_step<true>, dim3(grid), dim3(BD_BLOCK), 0, stream, a, i, cur);
            cur ^= 1;
            launches += 2;
        }
        const uint64_t pairs = (uint64_t)a.B * VJ * LJ;
        hipLaunchKernelGGL(k_bd_segments, dim3(div_up(pairs, BD_BLOCK)), dim3(BD_BLOCK), 0, stream, a);
        if ((rc = launch_occluded(c, a.segs, (uint32_t)pairs, a.occ))) return rc;
        hipLaunchKernelGGL(k_bd_combine, dim3(grid), dim3(BD_BLOCK), 0, stream, a);
        hipLaunchKernelGGL(k_bd_pixels, dim3(div_up(a.np, BD_BLOCK)), dim3(BD_BLOCK), 0, stream, a);
        RT_HIP(c, hipGetLastError());
        launches += 6;
    }
    RT_HIP(c, hipEventRecord(c->ev[1], stream));
    uint32_t err = 0;
    if ((rc = trace_error(c, &err))) return rc;
    if (err) return fail(c, RT_E_DEVICE, "rt_render_bdpt: traversal exceeded its stack or step bound");
    float ms = 0.0f;
    RT_HIP(c, hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
    c->timing.last_call_ms = ms;
    c->timing.launches = launches;
    c->last_op = 2;
    return RT_OK;
}

}  // namespace rt
