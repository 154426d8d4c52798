/*
 * oracle.c — CPU restatement of the reference software renderer's hot path.
 *
 * TEST INFRASTRUCTURE.  This is the checker: tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg load it; libmtrt.so never does.
 *
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off; plain IEEE float, no
 * fast-math), output oracle/liboracle.so.
 *
 * Reference files restated here (paths under software_renderer/):
 *   graphic/triangle.cpp:4-55        Triangle ctor, Moller-Trumbore, barycentrics
 *   graphic/graphics_object.cpp:64-87 closest hit by squared distance, strict <
 *   graphic/scene.cpp:22-56          two-level query, normal interpolation
 *   graphic/path_tracer.cpp:18-97    the estimator
 *   graphic/graphics_engine.cpp:59-97 per-pixel work loop, clamp, 8-bit store
 *   math/vector.h:202-241,288-345    normalize (double), dot, cross, reflect, refract
 *   math/matrix.h:245-253            row-vector product
 *   math/random.cpp:16-35            sphere / hemisphere sampling
 * Deliberate differences, each documented where it applies: the random
 * numbers come from Philox-4x32-10 instead of a mutex-guarded mt19937 seeded
 * from random_device (the reference is not reproducible run to run), sin/cos
 * of the sampling angle use a fixed polynomial (identical on the GPU), and
 * exact squared-distance ties go to the lowest triangle index (the reference
 * keeps whichever its tree order meets first).
 */
#define _POSIX_C_SOURCE 199309L  /* clock_gettime */
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "../include/rt.h"

typedef struct { float x, y, z; } v3;

static inline v3 V3(float x, float y, float z) { v3 r; r.x = x; r.y = y; r.z = z; return r; }
static inline v3 vld(const float* p) { return V3(p[0], p[1], p[2]); }
static inline v3 vsub(v3 a, v3 b) { return V3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 vadd(v3 a, v3 b) { return V3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 vmul(v3 a, v3 b) { return V3(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline v3 vscl(float s, v3 a) { return V3(s * a.x, s * a.y, s * a.z); }
/* vector.h:235-241 */
static inline float vdot(v3 a, v3 b) {
    float s = 0.0f;
    s = s + a.x * b.x;
    s = s + a.y * b.y;
    s = s + a.z * b.z;
    return s;
}
/* vector.h:288-297 */
static inline v3 vcross(v3 a, v3 b) {
    return V3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
/* vector.h:202-225 */
static inline v3 vnormalize(v3 a) {
    float m2 = vdot(a, a);
    double m = sqrt((double)m2);
    if (m != 0.0) {
        double f = 1.0 / m;
        a.x = (float)((double)a.x * f);
        a.y = (float)((double)a.y * f);
        a.z = (float)((double)a.z * f);
    }
    return a;
}

/* The probe lookup's single-precision normalisation (as the device's
 * normalize_f): a * (1 / sqrt(a . a)), IEEE square root and division.  How
 * the reference's GLSL normalize rounds is not pinned. */
static inline v3 vnormalize_f(v3 a) {
    const float m = sqrtf(vdot(a, a));
    if (m != 0.0f) {
        const float f = 1.0f / m;
        a.x = a.x * f;
        a.y = a.y * f;
        a.z = a.z * f;
    }
    return a;
}

/* ------------------------------------------------------------------ */
/* worker pool: threads pull chunks of [0, n) from an atomic counter   */
typedef void (*range_fn)(void* arg, uint32_t b, uint32_t e, int tid);
typedef struct {
    range_fn fn;
    void* arg;
    uint32_t n, chunk;
    atomic_uint next;
    int tid;
} pool_job;
typedef struct { pool_job* job; int tid; } pool_slot;

static void* pool_worker(void* p) {
    pool_slot* s = (pool_slot*)p;
    pool_job* j = s->job;
    for (;;) {
        uint32_t b = atomic_fetch_add(&j->next, j->chunk);
        if (b >= j->n) break;
        uint32_t e = b + j->chunk < j->n ? b + j->chunk : j->n;
        j->fn(j->arg, b, e, s->tid);
    }
    return NULL;
}

static void run_parallel(int threads, uint32_t n, uint32_t chunk, range_fn fn, void* arg) {
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    pool_job job;
    job.fn = fn; job.arg = arg; job.n = n; job.chunk = chunk ? chunk : 1;
    atomic_init(&job.next, 0u);
    pthread_t th[256];
    pool_slot slot[256];
    for (int t = 0; t < threads; ++t) {
        slot[t].job = &job;
        slot[t].tid = t;
    }
    for (int t = 1; t < threads; ++t) pthread_create(&th[t], NULL, pool_worker, &slot[t]);
    pool_worker(&slot[0]);
    for (int t = 1; t < threads; ++t) pthread_join(th[t], NULL);
}

/* ------------------------------------------------------------------ */
/* Philox-4x32-10 */
void or_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static inline float u01(uint32_t x) { return (float)(x >> 8) * (1.0f / 16777216.0f); }

void or_sincos_2pi(float x, float* s, float* c) {
    float e = x * 8.0f;
    int o = (int)e;
    float r = e - (float)o;
    if (o & 1) r = 1.0f - r;
    float a = r * 0x1.921fb6p-1f;
    float a2 = a * a;
    float ps = a * (1.0f + a2 * (-0x1.555556p-3f + a2 * (0x1.111112p-7f +
               a2 * (-0x1.a01a02p-13f + a2 * 0x1.71de3ap-19f))));
    float pc = 1.0f + a2 * (-0.5f + a2 * (0x1.555556p-5f + a2 * (-0x1.6c16c2p-10f +
               a2 * (0x1.a01a02p-16f + a2 * -0x1.27e4fcp-22f))));
    float sq = (o & 1) ? pc : ps;
    float cq = (o & 1) ? ps : pc;
    switch (o >> 1) {
        case 0: *s = sq;  *c = cq;  break;
        case 1: *s = cq;  *c = -sq; break;
        case 2: *s = -sq; *c = -cq; break;
        default: *s = -cq; *c = sq; break;
    }
}

static inline v3 hemisphere(float xi1, float xi2, v3 n) {
    float st, ct;
    or_sincos_2pi(xi1, &st, &ct);
    float z = 2.0f * xi2 - 1.0f;
    float r = sqrtf(1.0f - z * z);
    v3 d = V3(r * ct, r * st, z);
    if (vdot(d, n) < 0.0f) d = V3(-d.x, -d.y, -d.z);
    return d;
}

void or_hemisphere(float xi1, float xi2, const float n[3], float out[3]) {
    v3 d = hemisphere(xi1, xi2, vld(n));
    out[0] = d.x; out[1] = d.y; out[2] = d.z;
}

/* ------------------------------------------------------------------ */
/* triangle.cpp:21-45 */
static inline int mt(v3 o, v3 d, v3 v0, v3 e0, v3 e1, float* t, float* u, float* v) {
    const float EPS = 1e-7f;
    v3 h = vcross(d, e1);
    float a = vdot(e0, h);
    if (a > -EPS && a < EPS) return 0;
    float f = 1.0f / a;
    v3 s = vsub(o, v0);
    *u = f * vdot(s, h);
    if (*u < 0.0f || *u > 1.0f) return 0;
    v3 q = vcross(s, e0);
    *v = f * vdot(d, q);
    if (*v < 0.0f || *u + *v > 1.0f) return 0;
    *t = f * vdot(e1, q);
    return *t > EPS;
}

int or_intersect(const float o[3], const float d[3], const float v0[3], const float v1[3], const float v2[3],
                 float out[3]) {
    v3 a = vld(v0);
    float t, u, v;
    if (!mt(vld(o), vld(d), a, vsub(vld(v1), a), vsub(vld(v2), a), &t, &u, &v)) return 0;
    out[0] = t; out[1] = u; out[2] = v;
    return 1;
}

/* triangle.cpp:47-55 with the ctor's d00/d01/d11/denom (triangle.cpp:10-13) */
static inline void bary(v3 v0, v3 e0, v3 e1, v3 p, float* b0, float* b1, float* b2) {
    float d00 = vdot(e0, e0), d01 = vdot(e0, e1), d11 = vdot(e1, e1);
    float denom = d00 * d11 - d01 * d01;
    v3 e2 = vsub(p, v0);
    float d20 = vdot(e2, e0), d21 = vdot(e2, e1);
    float v = (d11 * d20 - d01 * d21) / denom;
    float w = (d00 * d21 - d01 * d20) / denom;
    *b0 = 1.0f - v - w;
    *b1 = v;
    *b2 = w;
}

void or_barycentric(const float v0[3], const float v1[3], const float v2[3], const float p[3], float out[3]) {
    v3 a = vld(v0);
    bary(a, vsub(vld(v1), a), vsub(vld(v2), a), vld(p), &out[0], &out[1], &out[2]);
}

/* Hit point and shading normal of a closest hit, scene.cpp:43-53:
 * p = o + t d, barycentrics from p, n = normalize(b0 n0 + b1 n1 + b2 n2). */
void or_hit_shading(const float* pos, const float* nrm, const uint32_t* idx, const void* rays, const void* hits,
                    uint32_t n, float* p_out, float* n_out) {
    const rt_ray* rr = (const rt_ray*)rays;
    const rt_hit* hh = (const rt_hit*)hits;
    for (uint32_t i = 0; i < n; ++i) {
        if (hh[i].tri == RT_MISS) {
            for (int k = 0; k < 3; ++k) { p_out[3 * i + k] = 0.0f; n_out[3 * i + k] = 0.0f; }
            continue;
        }
        uint32_t t = hh[i].tri;
        v3 o = V3(rr[i].ox, rr[i].oy, rr[i].oz), d = V3(rr[i].dx, rr[i].dy, rr[i].dz);
        v3 p = vadd(o, vscl(hh[i].t, d));
        v3 v0 = vld(pos + 3 * idx[3 * t]);
        float b0, b1, b2;
        bary(v0, vsub(vld(pos + 3 * idx[3 * t + 1]), v0), vsub(vld(pos + 3 * idx[3 * t + 2]), v0), p, &b0, &b1, &b2);
        v3 nn = V3(0.0f, 0.0f, 0.0f);
        nn = vadd(nn, vscl(b0, vld(nrm + 3 * idx[3 * t])));
        nn = vadd(nn, vscl(b1, vld(nrm + 3 * idx[3 * t + 1])));
        nn = vadd(nn, vscl(b2, vld(nrm + 3 * idx[3 * t + 2])));
        nn = vnormalize(nn);
        p_out[3 * i] = p.x; p_out[3 * i + 1] = p.y; p_out[3 * i + 2] = p.z;
        n_out[3 * i] = nn.x; n_out[3 * i + 1] = nn.y; n_out[3 * i + 2] = nn.z;
    }
}

/* graphics_object.cpp:75-83: keep the hit with the smallest |o - p|^2,
 * strict <; exact ties to the lowest index (the documented divergence). */
static inline int closer(float d2a, uint32_t ida, float d2b, uint32_t idb) {
    return d2a < d2b || (d2a == d2b && ida < idb);
}

typedef struct {
    float d2, t, u, v;
    uint32_t id;
} best_t;

static inline void consider(v3 o, v3 d, float tmin, float tmax, v3 v0, v3 e0, v3 e1, uint32_t id, best_t* b) {
    float t, u, v;
    if (mt(o, d, v0, e0, e1, &t, &u, &v) && t > tmin && t < tmax) {
        v3 p = vadd(o, vscl(t, d));
        v3 w = vsub(o, p);
        float d2 = vdot(w, w);
        if (closer(d2, id, b->d2, b->id)) {
            b->d2 = d2; b->t = t; b->u = u; b->v = v; b->id = id;
        }
    }
}

static inline void store_hit(rt_hit* h, const best_t* b) {
    if (b->id == RT_MISS) {
        h->tri = RT_MISS; h->t = 0.0f; h->u = 0.0f; h->v = 0.0f;
    } else {
        h->tri = b->id; h->t = b->t; h->u = b->u; h->v = b->v;
    }
}

typedef struct {
    const float* pos;
    const uint32_t* idx;
    uint32_t T;
    const rt_ray* rays;
    rt_hit* hits;
} brute_job;

static void brute_range(void* arg, uint32_t b, uint32_t e, int tid) {
    (void)tid;
    brute_job* j = (brute_job*)arg;
    for (uint32_t r = b; r < e; ++r) {
        const rt_ray* ry = &j->rays[r];
        v3 o = V3(ry->ox, ry->oy, ry->oz), d = V3(ry->dx, ry->dy, ry->dz);
        best_t bt;
        bt.d2 = INFINITY; bt.id = RT_MISS; bt.t = bt.u = bt.v = 0.0f;
        for (uint32_t t = 0; t < j->T; ++t) {
            v3 v0 = vld(j->pos + 3 * j->idx[3 * t]);
            v3 v1 = vld(j->pos + 3 * j->idx[3 * t + 1]);
            v3 v2 = vld(j->pos + 3 * j->idx[3 * t + 2]);
            consider(o, d, ry->tmin, ry->tmax, v0, vsub(v1, v0), vsub(v2, v0), t, &bt);
        }
        store_hit(&j->hits[r], &bt);
    }
}

void or_trace_brute(const float* pos, const uint32_t* idx, uint32_t T, const void* rays, uint32_t n, void* hits,
                    int threads) {
    brute_job j = {pos, idx, T, (const rt_ray*)rays, (rt_hit*)hits};
    run_parallel(threads, n, 64, brute_range, &j);
}

/* Scene::isOccluded (graphic/scene.cpp:58-83) by brute force: direction =
 * end - start, dist2 = |direction|^2 (float), direction normalised as
 * Vector::normalize does (vector.h:214-225, double factor); occluded when an
 * accepted triangle's hit point lies within squared distance dist2. */
typedef struct {
    const float* pos;
    const uint32_t* idx;
    uint32_t T;
    const float* segs;
    uint8_t* out;
} occl_job;

static void occl_range(void* arg, uint32_t b, uint32_t e, int tid) {
    (void)tid;
    occl_job* j = (occl_job*)arg;
    for (uint32_t r = b; r < e; ++r) {
        const float* sg = j->segs + 6 * (size_t)r;
        v3 o = vld(sg);
        v3 d = vsub(vld(sg + 3), o);
        float dist2 = vdot(d, d);
        d = vnormalize(d);
        uint8_t occ = 0;
        for (uint32_t t = 0; t < j->T && !occ; ++t) {
            v3 v0 = vld(j->pos + 3 * j->idx[3 * t]);
            v3 v1 = vld(j->pos + 3 * j->idx[3 * t + 1]);
            v3 v2 = vld(j->pos + 3 * j->idx[3 * t + 2]);
            float th, u, v;
            if (mt(o, d, v0, vsub(v1, v0), vsub(v2, v0), &th, &u, &v) && th > 0.0f) {
                v3 w = vsub(o, vadd(o, vscl(th, d)));
                if (vdot(w, w) <= dist2) occ = 1;
            }
        }
        j->out[r] = occ;
    }
}

void or_occluded_brute(const float* pos, const uint32_t* idx, uint32_t T, const float* segs, uint32_t n,
                       uint8_t* out, int threads) {
    occl_job j = {pos, idx, T, segs, out};
    run_parallel(threads, n, 64, occl_range, &j);
}

/* ------------------------------------------------------------------ */
/* culling shared by both BVH traversals (same rule as the GPU kernel) */
static inline float safe_rcp(float d) {
    const float tiny = 8.27180613e-25f; /* 2^-80 */
    float a = fabsf(d) < tiny ? copysignf(tiny, d) : d;
    return 1.0f / a;
}

typedef struct {
    v3 o, d;
    float tmin, tmax, tlo, sigma;
    float ix, iy, iz, nox, noy, noz;
    int valid;
} ray_pre;

static inline ray_pre prep(const rt_ray* r) {
    ray_pre p;
    p.o = V3(r->ox, r->oy, r->oz);
    p.d = V3(r->dx, r->dy, r->dz);
    float dmax = fmaxf(fabsf(p.d.x), fmaxf(fabsf(p.d.y), fabsf(p.d.z)));
    float omax = fmaxf(fabsf(p.o.x), fmaxf(fabsf(p.o.y), fabsf(p.o.z)));
    p.valid = (dmax > 0.0f) && (dmax < INFINITY) && (omax < INFINITY);
    p.tmin = r->tmin;
    p.tmax = r->tmax;
    p.tlo = fmaxf(r->tmin, 0.0f);
    p.sigma = p.valid ? 3.814697265625e-06f * omax / dmax : 0.0f;
    p.ix = safe_rcp(p.d.x); p.iy = safe_rcp(p.d.y); p.iz = safe_rcp(p.d.z);
    p.nox = -p.o.x * p.ix; p.noy = -p.o.y * p.iy; p.noz = -p.o.z * p.iz;
    return p;
}

static inline float slab(const ray_pre* p, float lx, float ly, float lz, float hx, float hy, float hz, float tcull,
                         int* hit) {
    float x0 = fmaf(lx, p->ix, p->nox), x1 = fmaf(hx, p->ix, p->nox);
    float y0 = fmaf(ly, p->iy, p->noy), y1 = fmaf(hy, p->iy, p->noy);
    float z0 = fmaf(lz, p->iz, p->noz), z1 = fmaf(hz, p->iz, p->noz);
    float tn = fmaxf(fmaxf(fminf(x0, x1), fminf(y0, y1)), fmaxf(fminf(z0, z1), p->tlo));
    float tf = fminf(fminf(fmaxf(x0, x1), fmaxf(y0, y1)), fminf(fmaxf(z0, z1), tcull));
    *hit = tn <= tf;
    return tn;
}

static inline float next_cull(const ray_pre* p, const best_t* b) {
    return fminf(p->tmax, fmaf(b->t, 1.0009765625f, p->sigma));
}

/* ------------------------------------------------------------------ */
/* binary BVH, binned SAH */
typedef struct {
    float lo[3], hi[3];
    uint32_t a;   /* leaf: first triangle record; inner: left child (right = a + 1) */
    uint32_t cnt; /* leaf: triangle count; inner: 0 */
} bnode;

struct or_bvh {
    bnode* nodes;
    uint32_t nn;
    float* rec; /* per leaf-ordered triangle: v0[3], e0[3], e1[3], id (as bits) = 10 floats */
    uint32_t T;
};

typedef struct {
    float lo[3], hi[3];
    float c[3];
} tbox;

uint32_t or_bvh_num_nodes(const or_bvh* b) { return b->nn; }

static void box_empty(float lo[3], float hi[3]) {
    for (int k = 0; k < 3; ++k) { lo[k] = INFINITY; hi[k] = -INFINITY; }
}
static void box_grow(float lo[3], float hi[3], const float l[3], const float h[3]) {
    for (int k = 0; k < 3; ++k) { lo[k] = fminf(lo[k], l[k]); hi[k] = fmaxf(hi[k], h[k]); }
}
static float box_area(const float lo[3], const float hi[3]) {
    float x = hi[0] - lo[0], y = hi[1] - lo[1], z = hi[2] - lo[2];
    if (!(x >= 0.0f)) return 0.0f;
    return x * y + y * z + z * x;
}

or_bvh* or_bvh_build(const float* pos, const uint32_t* idx, uint32_t T, int leaf_size) {
    if (leaf_size < 1) leaf_size = 4;
    or_bvh* b = (or_bvh*)calloc(1, sizeof(or_bvh));
    tbox* tb = (tbox*)malloc(sizeof(tbox) * (T ? T : 1));
    uint32_t* order = (uint32_t*)malloc(4 * (size_t)(T ? T : 1));
    b->nodes = (bnode*)malloc(sizeof(bnode) * (2 * (size_t)T + 1));
    float smax = 0.0f;
    for (uint32_t t = 0; t < T; ++t) {
        for (int k = 0; k < 3; ++k) {
            float a = pos[3 * idx[3 * t] + k], bb = pos[3 * idx[3 * t + 1] + k], c = pos[3 * idx[3 * t + 2] + k];
            tb[t].lo[k] = fminf(a, fminf(bb, c));
            tb[t].hi[k] = fmaxf(a, fmaxf(bb, c));
            tb[t].c[k] = 0.5f * (tb[t].lo[k] + tb[t].hi[k]);
            smax = fmaxf(smax, fmaxf(fabsf(tb[t].lo[k]), fabsf(tb[t].hi[k])));
        }
        order[t] = t;
    }
    typedef struct { uint32_t node, first, count; } item;
    item* stack = (item*)malloc(sizeof(item) * (2 * (size_t)T + 8));
    int sp = 0;
    b->nn = 1;
    stack[sp++] = (item){0, 0, T};
    enum { NB = 16 };
    while (sp > 0) {
        item it = stack[--sp];
        bnode* nd = &b->nodes[it.node];
        float lo[3], hi[3], clo[3], chi[3];
        box_empty(lo, hi);
        box_empty(clo, chi);
        for (uint32_t i = it.first; i < it.first + it.count; ++i) {
            box_grow(lo, hi, tb[order[i]].lo, tb[order[i]].hi);
            box_grow(clo, chi, tb[order[i]].c, tb[order[i]].c);
        }
        memcpy(nd->lo, lo, sizeof(lo));
        memcpy(nd->hi, hi, sizeof(hi));
        if (it.count <= (uint32_t)leaf_size) {
            nd->a = it.first; nd->cnt = it.count;
            continue;
        }
        int axis = 0;
        float ext = chi[0] - clo[0];
        for (int k = 1; k < 3; ++k)
            if (chi[k] - clo[k] > ext) { ext = chi[k] - clo[k]; axis = k; }
        uint32_t mid = 0;
        if (ext > 0.0f) {
            uint32_t cnt[NB] = {0};
            float blo[NB][3], bhi[NB][3];
            for (int q = 0; q < NB; ++q) box_empty(blo[q], bhi[q]);
            float scale = (float)NB / ext;
            for (uint32_t i = it.first; i < it.first + it.count; ++i) {
                int q = (int)((tb[order[i]].c[axis] - clo[axis]) * scale);
                if (q >= NB) q = NB - 1;
                if (q < 0) q = 0;
                cnt[q]++;
                box_grow(blo[q], bhi[q], tb[order[i]].lo, tb[order[i]].hi);
            }
            float best = INFINITY;
            int bsplit = -1;
            for (int s = 1; s < NB; ++s) {
                float l0[3], h0[3], l1[3], h1[3];
                uint32_t n0 = 0, n1 = 0;
                box_empty(l0, h0);
                box_empty(l1, h1);
                for (int q = 0; q < s; ++q) { n0 += cnt[q]; if (cnt[q]) box_grow(l0, h0, blo[q], bhi[q]); }
                for (int q = s; q < NB; ++q) { n1 += cnt[q]; if (cnt[q]) box_grow(l1, h1, blo[q], bhi[q]); }
                if (!n0 || !n1) continue;
                float cost = box_area(l0, h0) * (float)n0 + box_area(l1, h1) * (float)n1;
                if (cost < best) { best = cost; bsplit = s; }
            }
            if (bsplit > 0) {
                uint32_t i = it.first, j = it.first + it.count;
                while (i < j) {
                    int q = (int)((tb[order[i]].c[axis] - clo[axis]) * scale);
                    if (q >= NB) q = NB - 1;
                    if (q < 0) q = 0;
                    if (q < bsplit) ++i;
                    else { --j; uint32_t tmp = order[i]; order[i] = order[j]; order[j] = tmp; }
                }
                mid = i - it.first;
            }
        }
        if (mid == 0 || mid == it.count) mid = it.count / 2; /* coincident centroids: split by index */
        uint32_t l = b->nn;
        b->nn += 2;
        nd->a = l;
        nd->cnt = 0;
        stack[sp++] = (item){l + 1, it.first + mid, it.count - mid};
        stack[sp++] = (item){l, it.first, mid};
    }
    free(stack);
    /* pad every box outward by 2^-16 of the largest coordinate magnitude */
    float pad = smax * 1.52587890625e-05f + 1e-30f;
    for (uint32_t i = 0; i < b->nn; ++i)
        for (int k = 0; k < 3; ++k) { b->nodes[i].lo[k] -= pad; b->nodes[i].hi[k] += pad; }
    b->T = T;
    b->rec = (float*)malloc(sizeof(float) * 10 * (size_t)(T ? T : 1));
    for (uint32_t i = 0; i < T; ++i) {
        uint32_t t = order[i];
        float* r = b->rec + 10 * (size_t)i;
        for (int k = 0; k < 3; ++k) {
            float p0 = pos[3 * idx[3 * t] + k];
            r[k] = p0;
            r[3 + k] = pos[3 * idx[3 * t + 1] + k] - p0;
            r[6 + k] = pos[3 * idx[3 * t + 2] + k] - p0;
        }
        memcpy(&r[9], &t, 4);
    }
    free(tb);
    free(order);
    return b;
}

void or_bvh_free(or_bvh* b) {
    if (!b) return;
    free(b->nodes);
    free(b->rec);
    free(b);
}

static best_t trace_bvh_one(const or_bvh* b, const rt_ray* r, uint64_t* c_nodes, uint64_t* c_tris) {
    best_t bt;
    bt.d2 = INFINITY; bt.id = RT_MISS; bt.t = bt.u = bt.v = 0.0f;
    ray_pre p = prep(r);
    if (!p.valid || b->T == 0) return bt;
    float tcull = p.tmax;
    uint32_t stack[128];
    int sp = 0;
    int h;
    const bnode* root = &b->nodes[0];
    slab(&p, root->lo[0], root->lo[1], root->lo[2], root->hi[0], root->hi[1], root->hi[2], tcull, &h);
    if (!h) return bt;
    uint32_t cur = 0;
    for (;;) {
        const bnode* nd = &b->nodes[cur];
        (*c_nodes)++;
        if (nd->cnt) {
            for (uint32_t i = 0; i < nd->cnt; ++i) {
                const float* rc = b->rec + 10 * (size_t)(nd->a + i);
                uint32_t id;
                memcpy(&id, &rc[9], 4);
                (*c_tris)++;
                float before = bt.d2;
                uint32_t bid = bt.id;
                consider(p.o, p.d, p.tmin, p.tmax, vld(rc), vld(rc + 3), vld(rc + 6), id, &bt);
                if (bt.d2 != before || bt.id != bid) tcull = next_cull(&p, &bt);
            }
        } else {
            const bnode* l = &b->nodes[nd->a];
            const bnode* rr = &b->nodes[nd->a + 1];
            int hl, hr;
            float tl = slab(&p, l->lo[0], l->lo[1], l->lo[2], l->hi[0], l->hi[1], l->hi[2], tcull, &hl);
            float tr = slab(&p, rr->lo[0], rr->lo[1], rr->lo[2], rr->hi[0], rr->hi[1], rr->hi[2], tcull, &hr);
            if (hl && hr) {
                uint32_t nearc = tl <= tr ? nd->a : nd->a + 1, farc = tl <= tr ? nd->a + 1 : nd->a;
                stack[sp++] = farc;
                cur = nearc;
                continue;
            }
            if (hl) { cur = nd->a; continue; }
            if (hr) { cur = nd->a + 1; continue; }
        }
        if (sp == 0) break;
        cur = stack[--sp];
    }
    return bt;
}

typedef struct {
    const or_bvh* b;
    const rt_ray* rays;
    rt_hit* hits;
    uint64_t cnt[256][4];
} bvh_job;

static void bvh_range(void* arg, uint32_t b, uint32_t e, int tid) {
    bvh_job* j = (bvh_job*)arg;
    for (uint32_t r = b; r < e; ++r) {
        best_t bt = trace_bvh_one(j->b, &j->rays[r], &j->cnt[tid][1], &j->cnt[tid][2]);
        store_hit(&j->hits[r], &bt);
        j->cnt[tid][0]++;
        j->cnt[tid][3] += bt.id != RT_MISS;
    }
}

void or_trace_bvh(const or_bvh* b, const void* rays, uint32_t n, void* hits, int threads, uint64_t counters[4]) {
    bvh_job* j = (bvh_job*)calloc(1, sizeof(bvh_job));
    j->b = b;
    j->rays = (const rt_ray*)rays;
    j->hits = (rt_hit*)hits;
    run_parallel(threads, n, 256, bvh_range, j);
    if (counters) {
        for (int k = 0; k < 4; ++k) counters[k] = 0;
        for (int t = 0; t < 256; ++t)
            for (int k = 0; k < 4; ++k) counters[k] += j->cnt[t][k];
    }
    free(j);
}

/* ------------------------------------------------------------------ */
/* BVH4 traversal in the GPU kernel's exact order (trace.hip) */
typedef struct {
    const rt_bvh4q_node* nodes;
    const rt_tri_rec* tris;
    const rt_ray* rays;
    rt_hit* hits;
    uint32_t stack_cap;
    uint64_t cnt[256][4];
} bvh4_job;

static best_t trace_bvh4_one(const bvh4_job* j, const rt_ray* r, uint64_t* c_nodes, uint64_t* c_tris,
                             uint32_t* stack) {
    best_t bt;
    bt.d2 = INFINITY; bt.id = RT_MISS; bt.t = bt.u = bt.v = 0.0f;
    ray_pre p = prep(r);
    if (!p.valid) return bt;
    float tcull = p.tmax;
    uint32_t sp = 0, cur = 0;
    for (;;) {
        if (!(cur & RT_LEAF_BIT)) {
            /* quantised node: plane of byte q on axis k is origin + q * 2^exp,
             * its distance (origin - o) / d + q * (2^exp / d) by one fma */
            const rt_bvh4q_node* nd = &j->nodes[cur];
            (*c_nodes)++;
            const float ax = fmaf(nd->origin[0], p.ix, p.nox), bx = ldexpf(p.ix, nd->exp[0]);
            const float ay = fmaf(nd->origin[1], p.iy, p.noy), by = ldexpf(p.iy, nd->exp[1]);
            const float az = fmaf(nd->origin[2], p.iz, p.noz), bz = ldexpf(p.iz, nd->exp[2]);
            /* sort keys as the kernel forms them: the entry distance's bits
             * (sign cleared) with the two low bits replaced by the child's
             * slot, or 0x7F800000 | slot for a missed or empty child */
            uint32_t pk[4], ref[4];
            for (int c = 0; c < 4; ++c) {
                const float x0 = fmaf((float)nd->lo_x[c], bx, ax), x1 = fmaf((float)nd->hi_x[c], bx, ax);
                const float y0 = fmaf((float)nd->lo_y[c], by, ay), y1 = fmaf((float)nd->hi_y[c], by, ay);
                const float z0 = fmaf((float)nd->lo_z[c], bz, az), z1 = fmaf((float)nd->hi_z[c], bz, az);
                const float tn = fmaxf(fmaxf(fminf(x0, x1), fminf(y0, y1)), fmaxf(fminf(z0, z1), p.tlo));
                const float tf = fminf(fminf(fmaxf(x0, x1), fmaxf(y0, y1)), fminf(fmaxf(z0, z1), tcull));
                ref[c] = nd->child[c];
                uint32_t bits;
                memcpy(&bits, &tn, 4);
                pk[c] = ((tn <= tf && ref[c] != RT_CHILD_EMPTY) ? (bits & 0x7FFFFFFCu) : 0x7F800000u) | (uint32_t)c;
            }
#define CSWAP(i, k)                                                              \
    {                                                                            \
        const uint32_t lo = pk[i] < pk[k] ? pk[i] : pk[k], hi = pk[i] < pk[k] ? pk[k] : pk[i]; \
        pk[i] = lo; pk[k] = hi;                                                  \
    }
            CSWAP(0, 1) CSWAP(2, 3) CSWAP(0, 2) CSWAP(1, 3) CSWAP(1, 2)
#undef CSWAP
            if (pk[0] < 0x7F800000u) {
                for (int c = 3; c >= 1; --c)
                    if (pk[c] < 0x7F800000u && sp < j->stack_cap) stack[sp++] = ref[pk[c] & 3u];
                cur = ref[pk[0] & 3u];
                continue;
            }
        } else {
            uint32_t first = cur & RT_LEAF_FIRST_MASK;
            uint32_t cnt = ((cur >> RT_LEAF_COUNT_SHIFT) & 15u) + 1u;
            for (uint32_t i = 0; i < cnt; ++i) {
                const rt_tri_rec* rc = &j->tris[first + i];
                (*c_tris)++;
                float before = bt.d2;
                uint32_t bid = bt.id;
                consider(p.o, p.d, p.tmin, p.tmax, vld(rc->v0), vld(rc->e0), vld(rc->e1), rc->id, &bt);
                if (bt.d2 != before || bt.id != bid) tcull = next_cull(&p, &bt);
            }
        }
        if (sp == 0) break;
        cur = stack[--sp];
    }
    return bt;
}

static void bvh4_range(void* arg, uint32_t b, uint32_t e, int tid) {
    bvh4_job* j = (bvh4_job*)arg;
    uint32_t* stack = (uint32_t*)malloc(4 * (size_t)j->stack_cap);
    for (uint32_t r = b; r < e; ++r) {
        best_t bt = trace_bvh4_one(j, &j->rays[r], &j->cnt[tid][1], &j->cnt[tid][2], stack);
        store_hit(&j->hits[r], &bt);
        j->cnt[tid][0]++;
        j->cnt[tid][3] += bt.id != RT_MISS;
    }
    free(stack);
}

void or_trace_bvh4(const void* nodes, uint32_t num_nodes, const void* tris, uint32_t T, uint32_t depth,
                   const void* rays, uint32_t n, void* hits, int threads, uint64_t counters[4]) {
    (void)num_nodes;
    (void)T;
    bvh4_job* j = (bvh4_job*)calloc(1, sizeof(bvh4_job));
    j->nodes = (const rt_bvh4q_node*)nodes;
    j->tris = (const rt_tri_rec*)tris;
    j->rays = (const rt_ray*)rays;
    j->hits = (rt_hit*)hits;
    j->stack_cap = 3 * depth + 1;
    run_parallel(threads, n, 256, bvh4_range, j);
    if (counters) {
        for (int k = 0; k < 4; ++k) counters[k] = 0;
        for (int t = 0; t < 256; ++t)
            for (int k = 0; k < 4; ++k) counters[k] += j->cnt[t][k];
    }
    free(j);
}

/* the quantised traversal nodes against the build nodes: same children and
 * parent, every valid child's decoded box (origin + q * 2^exp, evaluated
 * exactly in double) contains its float box, empty children encode as
 * low 255 / high 0.  Returns 0 when sound. */
int or_check_quantized(const void* qnodes_v, const void* nodes_v, uint32_t num_nodes) {
    const rt_bvh4q_node* q = (const rt_bvh4q_node*)qnodes_v;
    const rt_bvh4_node* f = (const rt_bvh4_node*)nodes_v;
    for (uint32_t i = 0; i < num_nodes; ++i) {
        if (q[i].parent != f[i].meta[0]) return 1;
        const uint8_t* qlo[3] = {q[i].lo_x, q[i].lo_y, q[i].lo_z};
        const uint8_t* qhi[3] = {q[i].hi_x, q[i].hi_y, q[i].hi_z};
        const float* flo[3] = {f[i].lo_x, f[i].lo_y, f[i].lo_z};
        const float* fhi[3] = {f[i].hi_x, f[i].hi_y, f[i].hi_z};
        for (int c = 0; c < 4; ++c) {
            if (q[i].child[c] != f[i].child[c]) return 2;
            for (int k = 0; k < 3; ++k) {
                if (f[i].child[c] == RT_CHILD_EMPTY) {
                    if (qlo[k][c] != 255 || qhi[k][c] != 0) return 3;
                    continue;
                }
                const double s = ldexp(1.0, q[i].exp[k]);
                const double lo = (double)q[i].origin[k] + qlo[k][c] * s, hi = (double)q[i].origin[k] + qhi[k][c] * s;
                if (!(lo <= (double)flo[k][c]) || !(hi >= (double)fhi[k][c])) return 4;
                /* tight: one step looser at most */
                if ((double)flo[k][c] - lo >= s || hi - (double)fhi[k][c] >= s) return 5;
            }
        }
    }
    return 0;
}

/* structural check: returns 0 when sound */
int or_check_bvh4(const void* nodes_v, uint32_t num_nodes, const void* tris_v, uint32_t T, const float* pos,
                  const uint32_t* idx, uint32_t* out_depth) {
    const rt_bvh4_node* nodes = (const rt_bvh4_node*)nodes_v;
    const rt_tri_rec* tris = (const rt_tri_rec*)tris_v;
    uint8_t* seen_tri = (uint8_t*)calloc(T ? T : 1, 1);
    uint8_t* seen_rec = (uint8_t*)calloc(T ? T : 1, 1);
    uint8_t* seen_node = (uint8_t*)calloc(num_nodes ? num_nodes : 1, 1);
    typedef struct { uint32_t node, depth; } it;
    it* st = (it*)malloc(sizeof(it) * ((size_t)num_nodes * 4 + 4));
    int sp = 0, rc = 0;
    uint32_t maxd = 0;
    st[sp++] = (it){0, 1};
    while (sp > 0 && !rc) {
        it x = st[--sp];
        if (x.node >= num_nodes) { rc = 1; break; }
        if (seen_node[x.node]) { rc = 2; break; }
        seen_node[x.node] = 1;
        if (x.depth > maxd) maxd = x.depth;
        const rt_bvh4_node* nd = &nodes[x.node];
        for (int c = 0; c < 4 && !rc; ++c) {
            uint32_t ch = nd->child[c];
            if (ch == RT_CHILD_EMPTY) continue;
            float lo[3] = {nd->lo_x[c], nd->lo_y[c], nd->lo_z[c]}, hi[3] = {nd->hi_x[c], nd->hi_y[c], nd->hi_z[c]};
            if (ch & RT_LEAF_BIT) {
                uint32_t first = ch & RT_LEAF_FIRST_MASK, cnt = ((ch >> RT_LEAF_COUNT_SHIFT) & 15u) + 1u;
                if ((uint64_t)first + cnt > T) { rc = 3; break; }
                for (uint32_t i = first; i < first + cnt; ++i) {
                    if (seen_rec[i]) { rc = 4; break; }
                    seen_rec[i] = 1;
                    uint32_t t = tris[i].id;
                    if (t >= T || seen_tri[t]) { rc = 5; break; }
                    seen_tri[t] = 1;
                    for (int v = 0; v < 3 && !rc; ++v)
                        for (int k = 0; k < 3; ++k) {
                            float x = pos[3 * idx[3 * t + v] + k];
                            if (!(x >= lo[k] && x <= hi[k])) { rc = 6; break; }
                        }
                    /* record equals v0 / v1 - v0 / v2 - v0 */
                    for (int k = 0; k < 3 && !rc; ++k) {
                        float p0 = pos[3 * idx[3 * t] + k];
                        if (tris[i].v0[k] != p0 || tris[i].e0[k] != pos[3 * idx[3 * t + 1] + k] - p0 ||
                            tris[i].e1[k] != pos[3 * idx[3 * t + 2] + k] - p0)
                            rc = 7;
                    }
                }
            } else {
                const rt_bvh4_node* cn = &nodes[ch];
                for (int cc = 0; cc < 4 && !rc; ++cc) {
                    if (cn->child[cc] == RT_CHILD_EMPTY) continue;
                    if (cn->lo_x[cc] < lo[0] || cn->lo_y[cc] < lo[1] || cn->lo_z[cc] < lo[2] ||
                        cn->hi_x[cc] > hi[0] || cn->hi_y[cc] > hi[1] || cn->hi_z[cc] > hi[2])
                        rc = 8;
                }
                if (cn->meta[0] != x.node) rc = 9;
                st[sp++] = (it){ch, x.depth + 1};
            }
        }
    }
    if (!rc)
        for (uint32_t t = 0; t < T; ++t)
            if (!seen_tri[t]) { rc = 10; break; }
    if (!rc)
        for (uint32_t i = 0; i < num_nodes; ++i)
            if (!seen_node[i]) { rc = 11; break; }
    if (out_depth) *out_depth = maxd;
    free(seen_tri); free(seen_rec); free(seen_node); free(st);
    return rc;
}

/* ------------------------------------------------------------------ */
/* camera rays, math/matrix.h:245-253 and graphic/path_tracer.cpp:19-23 */
void or_vec4_mat(const float m[16], const float v[4], float out[4]) {
    for (int j = 0; j < 4; ++j) out[j] = 0.0f;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) out[j] = out[j] + m[4 * i + j] * v[i];
}

static v3 primary(const rt_camera* cam, uint32_t px, uint32_t py) {
    float cx = (float)px + 0.5f, cy = (float)py + 0.5f;
    float ux = cx / (float)cam->width, uy = cy / (float)cam->height;
    float dx = 2.0f * ux - 1.0f, dy = 2.0f * uy - 1.0f;
    float v4[4] = {dx, dy, 1.0f, 1.0f}, t4[4];
    or_vec4_mat(cam->proj_inv, v4, t4);
    v3 target = vnormalize(V3(t4[0], t4[1], t4[2]));
    float w4[4] = {target.x, target.y, target.z, 0.0f}, d4[4];
    or_vec4_mat(cam->view_inv, w4, d4);
    return vnormalize(V3(d4[0], d4[1], d4[2]));
}

void or_primary_dir(const float view_inv[16], const float proj_inv[16], uint32_t px, uint32_t py, uint32_t w,
                    uint32_t h, float out[3]) {
    rt_camera cam;
    memcpy(cam.view_inv, view_inv, sizeof(cam.view_inv));
    memcpy(cam.proj_inv, proj_inv, sizeof(cam.proj_inv));
    cam.width = w;
    cam.height = h;
    v3 d = primary(&cam, px, py);
    out[0] = d.x; out[1] = d.y; out[2] = d.z;
}

/* ------------------------------------------------------------------ */
/* vector.h:319-345 */
static v3 reflect_ref(v3 I, v3 N) {
    double k = 2.0 * (double)vdot(N, I);
    v3 kn = V3((float)(k * (double)N.x), (float)(k * (double)N.y), (float)(k * (double)N.z));
    return vsub(I, kn);
}
static v3 refract_ref(v3 I, v3 N, float eta) {
    float ndi = vdot(N, I);
    float k = (float)(1.0 - (double)(eta * eta) * (1.0 - (double)(ndi * ndi)));
    if (k < 0.0f) return V3(0.0f, 0.0f, 0.0f);
    v3 ei = vscl(eta, I);
    double m = (double)(eta * ndi) + sqrt((double)k);
    v3 mn = V3((float)(m * (double)N.x), (float)(m * (double)N.y), (float)(m * (double)N.z));
    return vsub(ei, mn);
}
/* sin(acos(x)) as sqrt((1 - x)(1 + x)) in double: both factors are exact for
 * a float x, so host and device round identically; NaN for |x| > 1 as acos.
 * Within 2^-50 of glibc's sin(acos(x)) on [-1, 1] (tests/test_oracle.py). */
double or_sin_acos(double x) { return sqrt((1.0 - x) * (1.0 + x)); }
static v3 custom_refract_ref(v3 I, v3 N, float eta) {
    if ((double)vdot(N, I) > 0.0) N = V3(-N.x, -N.y, -N.z);
    else eta = (float)(1.0 / (double)eta);
    float angle = (float)(or_sin_acos((double)vdot(N, I)) * (double)eta);
    if (-1.0f < angle && angle < 1.0f) return refract_ref(I, N, eta);
    return reflect_ref(I, N);
}
/* vector.h:319-345 entry points for the tests */
void or_reflect(const float I[3], const float N[3], float out[3]) {
    v3 r = reflect_ref(V3(I[0], I[1], I[2]), V3(N[0], N[1], N[2]));
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
void or_custom_refract(const float I[3], const float N[3], float eta, float out[3]) {
    v3 r = custom_refract_ref(V3(I[0], I[1], I[2]), V3(N[0], N[1], N[2]), eta);
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}

/* ------------------------------------------------------------------ */
/* the path tracer */
typedef struct {
    const float* pos;
    const float* nrm;
    const uint32_t* idx;
    const uint32_t* tri_mat;
    const rt_material* mat;
    uint32_t M;
    const rt_camera* cams; /* num_views cameras of one size */
    uint32_t num_views;
    const rt_settings* s;
    or_bvh* bvh;
    double scale;
    float* rgb;
    uint8_t* rgb8;
    const float* dimg;
    double* dcol; /* [256 threads][3M] */
    /* radiance-cache mode (cache_rgb != NULL): a constant cache */
    const float* cache_rgb;
    uint32_t q_vertex, train_every, query_on, nee;
    /* probe-field mode (rt_render_ddgi): the field answers the queries */
    const rt_ddgi_config* ddgi;
    const float* ddgi_irr;
    const float* ddgi_dep;
    uint64_t* n_queries; /* [256 threads] */
    uint64_t* n_records; /* [256 threads] */
    float* rec;          /* per (vp, sample): 9 features, 3 targets, valid flag (13 floats), or NULL */
    /* light sampling: emissive triangles of positive area, cumulative area fraction */
    uint32_t num_lights;
    uint32_t* light_tri;
    float* light_cdf;
    float light_area;
} pt_job;

/* Training-sample choice of the cache mode: lowbias32 (C. Wellons) over
 * pixel, sample, view and the low seed word. */
static uint32_t train_hash(uint32_t pix, uint32_t smp, uint32_t view, uint32_t seed) {
    uint32_t h = pix * 0x9E3779B1u ^ (smp + 0x7F4A7C15u) * 0x85EBCA77u ^ (view + 0x632BE5ABu) * 0xC2B2AE3Du ^ seed;
    h ^= h >> 16;
    h *= 0x7FEB352Du;
    h ^= h >> 15;
    h *= 0x846CA68Bu;
    h ^= h >> 16;
    return h;
}

#define OR_MAX_DEPTH 64

/* One sample of graphic/path_tracer.cpp:42-97.  Returns 1 and the
 * contribution when the path ends on a light (or escapes to the sky);
 * nv / vm / vc receive the diffuse vertices for the adjoint. */
static int trace_path(const pt_job* j, uint32_t view, uint32_t pix, uint32_t gsmp, v3 dir0, v3* contrib, v3* term,
                      uint32_t* nv, uint32_t* vm, float* vc) {
    const rt_settings* s = j->s;
    const rt_camera* cam = &j->cams[view];
    v3 o = V3(cam->origin[0], cam->origin[1], cam->origin[2]);
    v3 d = dir0;
    v3 T = V3(1.0f, 1.0f, 1.0f);
    int cull = 1;
    *nv = 0;
    *term = V3(1.0f, 1.0f, 1.0f);
    uint32_t key[2] = {(uint32_t)s->seed, (uint32_t)(s->seed >> 32)};
    for (uint32_t depth = 0; depth < s->max_depth; ++depth) {
        rt_ray r;
        r.ox = o.x; r.oy = o.y; r.oz = o.z; r.tmin = 0.0f;
        r.dx = d.x; r.dy = d.y; r.dz = d.z; r.tmax = INFINITY;
        uint64_t cn = 0, ct = 0;
        best_t h = trace_bvh_one(j->bvh, &r, &cn, &ct);
        if (h.id == RT_MISS) {
            /* extension: sky and sun radiance for escaped rays (the reference's miss is black) */
            if (s->flags & (RT_RENDER_SKY | RT_RENDER_SUN)) {
                v3 L = V3(0.0f, 0.0f, 0.0f);
                if (s->flags & RT_RENDER_SKY) L = V3(s->sky[0], s->sky[1], s->sky[2]);
                if ((s->flags & RT_RENDER_SUN) && vdot(d, vld(s->sun_dir)) >= s->sun_cos)
                    L = vadd(L, vld(s->sun_radiance));
                *term = L;
                *contrib = vmul(T, L);
                return 1;
            }
            return 0;
        }
        v3 p = vadd(o, vscl(h.t, d));
        uint32_t t = h.id;
        v3 v0 = vld(j->pos + 3 * j->idx[3 * t]);
        v3 e0 = vsub(vld(j->pos + 3 * j->idx[3 * t + 1]), v0);
        v3 e1 = vsub(vld(j->pos + 3 * j->idx[3 * t + 2]), v0);
        float b0, b1, b2;
        bary(v0, e0, e1, p, &b0, &b1, &b2);
        /* scene.cpp:49-53 */
        v3 n = V3(0.0f, 0.0f, 0.0f);
        n = vadd(n, vscl(b0, vld(j->nrm + 3 * j->idx[3 * t])));
        n = vadd(n, vscl(b1, vld(j->nrm + 3 * j->idx[3 * t + 1])));
        n = vadd(n, vscl(b2, vld(j->nrm + 3 * j->idx[3 * t + 2])));
        n = vnormalize(n);
        float ndotd = vdot(n, d);
        if (cull && ndotd > 0.0f) {
            o = vadd(p, vscl(0.01f, d));
            continue;
        }
        cull = 0;
        const rt_material* m = &j->mat[j->tri_mat[t]];
        uint32_t ctr[4] = {pix, gsmp, depth, view}, rnd[4];
        or_philox(ctr, key, rnd);
        float xi0 = u01(rnd[0]);
        v3 prev = d;
        int diffuse = xi0 <= m->diffuse_threshold;
        if (diffuse) d = hemisphere(u01(rnd[1]), u01(rnd[2]), n);
        else if (xi0 <= m->reflect_threshold) d = reflect_ref(d, n);
        else if (xi0 <= m->transparent_threshold) d = custom_refract_ref(d, n, m->refraction_index);
        o = vadd(p, vscl(0.01f, d));
        if (diffuse) {
            float cs = vdot(n, d);
            T = vmul(T, vscl(cs, V3(m->color[0], m->color[1], m->color[2])));
            if (*nv < OR_MAX_DEPTH) { vm[*nv] = j->tri_mat[t]; vc[*nv] = cs; }
            (*nv)++;
        }
        if (m->light_source && vdot(n, prev) <= 0.0f) {
            *contrib = T;
            return 1;
        }
    }
    return 0;
}

/* Light sample of the cache mode at vertex p (the expected value of the
 * diffuse branch's next-vertex light hit, or of its escape inside the sun's
 * cone): the GPU's nee_sample restated with the same float operations. */
static int nee_sample(const pt_job* j, v3 p, v3 nrm, const rt_material* m, const uint32_t rnd[4], rt_ray* sh, v3* C) {
    const rt_settings* s = j->s;
    const int sun_on = (s->flags & RT_RENDER_SUN) && s->sun_cos < 1.0f;
    const uint32_t nls = (j->num_lights > 0 ? 1u : 0u) + (sun_on ? 1u : 0u);
    if (nls == 0 || !(m->diffuse_threshold > 0.0f)) return 0;
    const int use_light = j->num_lights > 0 && (!sun_on || u01(rnd[0]) < 0.5f);
    const v3 col = V3(m->color[0], m->color[1], m->color[2]);
    v3 w;
    if (use_light) {
        const float u = u01(rnd[1]);
        uint32_t lo = 0, hi = j->num_lights - 1;
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (j->light_cdf[mid] > u) hi = mid;
            else lo = mid + 1;
        }
        const uint32_t t = j->light_tri[lo];
        const v3 v0 = vld(j->pos + 3 * j->idx[3 * t]);
        const v3 e0 = vsub(vld(j->pos + 3 * j->idx[3 * t + 1]), v0);
        const v3 e1 = vsub(vld(j->pos + 3 * j->idx[3 * t + 2]), v0);
        const float su = sqrtf(u01(rnd[2])), u2 = u01(rnd[3]);
        const float b1 = su * (1.0f - u2), b2 = su * u2, b0 = 1.0f - su;
        const v3 y = vadd(vadd(v0, vscl(b1, e0)), vscl(b2, e1));
        v3 ny = vscl(b0, vld(j->nrm + 3 * j->idx[3 * t]));
        ny = vadd(ny, vscl(b1, vld(j->nrm + 3 * j->idx[3 * t + 1])));
        ny = vadd(ny, vscl(b2, vld(j->nrm + 3 * j->idx[3 * t + 2])));
        ny = vnormalize(ny);
        const v3 wv = vsub(y, p);
        const float d2 = vdot(wv, wv);
        const float dist = sqrtf(d2);
        if (!(dist > 0.02f)) return 0;
        w = vscl(1.0f / dist, wv);
        const float cx = vdot(nrm, w), cy = -vdot(ny, w);
        if (!(cx > 0.0f) || !(cy > 0.0f)) return 0;
        const rt_material* lm = &j->mat[j->tri_mat[t]];
        const float tdl = lm->diffuse_threshold;
        const v3 RL = V3(tdl * lm->color[0] * 0.5f + (1.0f - tdl), tdl * lm->color[1] * 0.5f + (1.0f - tdl),
                         tdl * lm->color[2] * 0.5f + (1.0f - tdl));
        const float f = m->diffuse_threshold * cx * cy * j->light_area * (float)nls / (6.28318530717958647f * d2);
        *C = vscl(f, vmul(col, RL));
        sh->tmax = (dist - 0.01f) * 0.9990234375f;
    } else {
        const v3 sd = vld(s->sun_dir);
        const float ct = 1.0f - u01(rnd[1]) * (1.0f - s->sun_cos);
        const float st = sqrtf(fmaxf(0.0f, 1.0f - ct * ct));
        float sp, cp;
        or_sincos_2pi(u01(rnd[2]), &sp, &cp);
        /* orthonormal basis around the sun direction (Duff et al., JCGT 2017) */
        const float sg = copysignf(1.0f, sd.z);
        const float ia = -1.0f / (sg + sd.z);
        const float ib = sd.x * sd.y * ia;
        const v3 t1 = V3(1.0f + sg * sd.x * sd.x * ia, sg * ib, -sg * sd.x);
        const v3 t2 = V3(ib, sg + sd.y * sd.y * ia, -sd.y);
        w = vadd(vscl(ct, sd), vscl(st, vadd(vscl(cp, t1), vscl(sp, t2))));
        const float cx = vdot(nrm, w);
        if (!(cx > 0.0f)) return 0;
        const float f = m->diffuse_threshold * cx * (1.0f - s->sun_cos) * (float)nls;
        *C = vscl(f, vmul(col, vld(s->sun_radiance)));
        sh->tmax = INFINITY;
    }
    const v3 o = vadd(p, vscl(0.01f, w));
    sh->ox = o.x; sh->oy = o.y; sh->oz = o.z; sh->tmin = 0.0f;
    sh->dx = w.x; sh->dy = w.y; sh->dz = w.z;
    return 1;
}

static v3 ddgi_lookup(const rt_ddgi_config* g, const float* irr, const float* dep, v3 x, v3 n, v3 vd);

/* One sample of the cache mode (rt_render_nrc with a constant cache):
 * returns its summed pixel contribution; a training sample fills rec
 * (9 features, 3 targets, flag 1) when rec is not NULL. */
static v3 trace_path_cached(const pt_job* j, uint32_t view, uint32_t pix, uint32_t gsmp, v3 dir0, int tid, float* rec) {
    const rt_settings* s = j->s;
    const rt_camera* cam = &j->cams[view];
    v3 o = V3(cam->origin[0], cam->origin[1], cam->origin[2]);
    v3 d = dir0;
    v3 T = V3(1.0f, 1.0f, 1.0f), Tq = V3(1.0f, 1.0f, 1.0f), sum = V3(0.0f, 0.0f, 0.0f), tgt = V3(0.0f, 0.0f, 0.0f);
    int cull = 1, has_rec = 0, prev_diff = 0;
    uint32_t fh = 0;
    float feats[9];
    uint32_t key[2] = {(uint32_t)s->seed, (uint32_t)(s->seed >> 32)};
    if (rec) memset(rec, 0, 13 * sizeof(float));
    for (uint32_t depth = 0; depth < s->max_depth; ++depth) {
        rt_ray r;
        r.ox = o.x; r.oy = o.y; r.oz = o.z; r.tmin = 0.0f;
        r.dx = d.x; r.dy = d.y; r.dz = d.z; r.tmax = INFINITY;
        uint64_t cn = 0, ct = 0;
        best_t h = trace_bvh_one(j->bvh, &r, &cn, &ct);
        if (h.id == RT_MISS) {
            if (s->flags & (RT_RENDER_SKY | RT_RENDER_SUN)) {
                v3 L = V3(0.0f, 0.0f, 0.0f);
                if (s->flags & RT_RENDER_SKY) L = V3(s->sky[0], s->sky[1], s->sky[2]);
                if (!(j->nee && prev_diff) && (s->flags & RT_RENDER_SUN) && vdot(d, vld(s->sun_dir)) >= s->sun_cos)
                    L = vadd(L, vld(s->sun_radiance));
                sum = vadd(sum, vmul(T, L));
                if (has_rec) tgt = vadd(tgt, vmul(Tq, L));
            }
            break;
        }
        v3 p = vadd(o, vscl(h.t, d));
        uint32_t t = h.id;
        v3 v0 = vld(j->pos + 3 * j->idx[3 * t]);
        v3 e0 = vsub(vld(j->pos + 3 * j->idx[3 * t + 1]), v0);
        v3 e1 = vsub(vld(j->pos + 3 * j->idx[3 * t + 2]), v0);
        float b0, b1, b2;
        bary(v0, e0, e1, p, &b0, &b1, &b2);
        v3 n = V3(0.0f, 0.0f, 0.0f);
        n = vadd(n, vscl(b0, vld(j->nrm + 3 * j->idx[3 * t])));
        n = vadd(n, vscl(b1, vld(j->nrm + 3 * j->idx[3 * t + 1])));
        n = vadd(n, vscl(b2, vld(j->nrm + 3 * j->idx[3 * t + 2])));
        n = vnormalize(n);
        const float ndotd = vdot(n, d);
        if (cull && ndotd > 0.0f) {
            o = vadd(p, vscl(0.01f, d));
            continue;
        }
        cull = 0;
        const rt_material* m = &j->mat[j->tri_mat[t]];
        const int light_end = m->light_source && ndotd <= 0.0f;
        if (light_end && j->nee) {
            /* after a diffuse bounce: estimated by the previous vertex's light sample; else the light's
             * expected factor td_L * c_L / 2 + (1 - td_L) */
            if (!prev_diff) {
                const float tdl = m->diffuse_threshold;
                const v3 RL = V3(tdl * m->color[0] * 0.5f + (1.0f - tdl), tdl * m->color[1] * 0.5f + (1.0f - tdl),
                                 tdl * m->color[2] * 0.5f + (1.0f - tdl));
                sum = vadd(sum, vmul(T, RL));
                if (has_rec) tgt = vadd(tgt, vmul(Tq, RL));
            }
            break;
        }
        if (!light_end && fh == j->q_vertex && depth + 1 < s->max_depth &&
            (!j->ddgi || m->diffuse_threshold >= 1.0f)) {
            const int train = j->train_every != 0 && train_hash(pix, gsmp, view, key[0]) % j->train_every == 0;
            if (train) {
                j->n_records[tid]++;
                has_rec = 1;
                Tq = V3(1.0f, 1.0f, 1.0f);
                tgt = V3(0.0f, 0.0f, 0.0f);
                feats[0] = p.x; feats[1] = p.y; feats[2] = p.z;
                feats[3] = n.x; feats[4] = n.y; feats[5] = n.z;
                feats[6] = d.x; feats[7] = d.y; feats[8] = d.z;
            } else if (j->query_on) {
                j->n_queries[tid]++;
                if (j->ddgi) {
                    /* probe field: the vertex's own light sample, then albedo / 2 * lookup */
                    if (j->nee) {
                        uint32_t c5[4] = {pix, gsmp, depth, view | 0x80000000u}, r5[4];
                        or_philox(c5, key, r5);
                        rt_ray sh;
                        v3 C;
                        if (nee_sample(j, p, n, m, r5, &sh, &C)) {
                            uint64_t a1 = 0, a2 = 0;
                            best_t oh = trace_bvh_one(j->bvh, &sh, &a1, &a2);
                            if (oh.id == RT_MISS) sum = vadd(sum, vmul(T, C));
                        }
                    }
                    const v3 e = ddgi_lookup(j->ddgi, j->ddgi_irr, j->ddgi_dep, p, n, d);
                    const v3 qw = vmul(T, vscl(0.5f, V3(m->color[0], m->color[1], m->color[2])));
                    sum = vadd(sum, vmul(qw, e));
                } else {
                    sum = vadd(sum, vmul(T, V3(fmaxf(j->cache_rgb[0], 0.0f), fmaxf(j->cache_rgb[1], 0.0f),
                                                fmaxf(j->cache_rgb[2], 0.0f))));
                }
                break;
            }
        }
        ++fh;
        uint32_t ctr[4] = {pix, gsmp, depth, view}, rnd[4];
        if (j->nee && !light_end && depth + 1 < s->max_depth) {
            uint32_t c5[4] = {pix, gsmp, depth, view | 0x80000000u}, r5[4];
            or_philox(c5, key, r5);
            rt_ray sh;
            v3 C;
            if (nee_sample(j, p, n, m, r5, &sh, &C)) {
                uint64_t a1 = 0, a2 = 0;
                best_t oh = trace_bvh_one(j->bvh, &sh, &a1, &a2);
                if (oh.id == RT_MISS) {
                    sum = vadd(sum, vmul(T, C));
                    if (has_rec) tgt = vadd(tgt, vmul(Tq, C));
                }
            }
        }
        or_philox(ctr, key, rnd);
        float xi0 = u01(rnd[0]);
        int diffuse = xi0 <= m->diffuse_threshold;
        if (diffuse) d = hemisphere(u01(rnd[1]), u01(rnd[2]), n);
        else if (xi0 <= m->reflect_threshold) d = reflect_ref(d, n);
        else if (xi0 <= m->transparent_threshold) d = custom_refract_ref(d, n, m->refraction_index);
        o = vadd(p, vscl(0.01f, d));
        if (diffuse) {
            const v3 f = vscl(vdot(n, d), V3(m->color[0], m->color[1], m->color[2]));
            T = vmul(T, f);
            Tq = vmul(Tq, f);
        }
        prev_diff = diffuse;
        if (light_end) {
            sum = vadd(sum, T);
            if (has_rec) tgt = vadd(tgt, Tq);
            break;
        }
    }
    if (rec && has_rec) {
        memcpy(rec, feats, sizeof(feats));
        rec[9] = tgt.x; rec[10] = tgt.y; rec[11] = tgt.z;
        rec[12] = 1.0f;
    }
    return sum;
}

static void pt_range(void* arg, uint32_t b, uint32_t e, int tid) {
    pt_job* j = (pt_job*)arg;
    const uint32_t P = j->cams[0].width * j->cams[0].height;
    uint32_t vm[OR_MAX_DEPTH];
    float vc[OR_MAX_DEPTH];
    /* vp = view * P + pixel; images are view-major */
    for (uint32_t vp = b; vp < e; ++vp) {
        const uint32_t view = vp / P, pix = vp % P;
        const rt_camera* cam = &j->cams[view];
        v3 dir = primary(cam, pix % cam->width, pix / cam->width);
        float sum[3] = {0.0f, 0.0f, 0.0f};
        for (uint32_t smp = 0; smp < j->s->spp; ++smp) {
            v3 c, term = V3(0.0f, 0.0f, 0.0f);
            uint32_t nv = 0;
            if (j->cache_rgb) {
                /* cache mode: every sample's summed contribution */
                c = trace_path_cached(j, view, pix, j->s->sample_offset + smp, dir, tid,
                                      j->rec ? j->rec + 13 * ((size_t)vp * j->s->spp + smp) : NULL);
            } else if (!trace_path(j, view, pix, j->s->sample_offset + smp, dir, &c, &term, &nv, vm, vc)) {
                continue;
            }
            if (j->dimg) {
                float w[3], tm[3] = {term.x, term.y, term.z};
                for (int ch = 0; ch < 3; ++ch) w[ch] = j->dimg[3 * (size_t)vp + ch] * (float)j->scale;
                for (int ch = 0; ch < 3; ++ch) {
                    if (w[ch] == 0.0f) continue;
                    float f[OR_MAX_DEPTH], suf[OR_MAX_DEPTH + 1], pre = 1.0f;
                    for (uint32_t k = 0; k < nv; ++k) f[k] = j->mat[vm[k]].color[ch] * vc[k];
                    suf[nv] = tm[ch];
                    for (int k = (int)nv - 1; k >= 0; --k) suf[k] = f[k] * suf[k + 1];
                    for (uint32_t k = 0; k < nv; ++k) {
                        float g = pre * suf[k + 1] * vc[k];
                        j->dcol[(size_t)tid * 3 * j->M + 3 * vm[k] + ch] += (double)(w[ch] * g);
                        pre = pre * f[k];
                    }
                }
            } else {
                sum[0] = sum[0] + c.x;
                sum[1] = sum[1] + c.y;
                sum[2] = sum[2] + c.z;
            }
        }
        if (!j->dimg) {
            for (int ch = 0; ch < 3; ++ch) {
                float v = (float)((double)sum[ch] * j->scale);
                j->rgb[3 * (size_t)vp + ch] = v;
                if (j->rgb8) {
                    float cv = fminf(fmaxf(v, 0.0f), 1.0f);
                    j->rgb8[3 * (size_t)vp + ch] = (uint8_t)(int)(cv * 255.0f);
                }
            }
        }
    }
}

typedef struct {
    const float* rgb;
    uint32_t q_vertex, train_every, query_on, nee;
    float* rec;
    uint64_t queries, records;
    const rt_ddgi_config* ddgi;
    const float* ddgi_irr;
    const float* ddgi_dep;
} cache_cfg;

/* emissive triangles by cumulative area, float, in triangle order */
static void setup_lights(pt_job* j, const float* pos, const uint32_t* idx, const uint32_t* tri_mat, uint32_t T) {
    j->light_tri = (uint32_t*)malloc(sizeof(uint32_t) * (T ? T : 1));
    j->light_cdf = (float*)malloc(sizeof(float) * (T ? T : 1));
    j->num_lights = 0;
    float acc = 0.0f;
    for (uint32_t t = 0; t < T; ++t) {
        if (!j->mat[tri_mat[t]].light_source) continue;
        const v3 p0 = vld(pos + 3 * idx[3 * t]), p1 = vld(pos + 3 * idx[3 * t + 1]), p2 = vld(pos + 3 * idx[3 * t + 2]);
        const v3 c = vcross(vsub(p1, p0), vsub(p2, p0));
        const float area = 0.5f * sqrtf(vdot(c, c));
        if (!(area > 0.0f)) continue;
        acc = acc + area;
        j->light_tri[j->num_lights] = t;
        j->light_cdf[j->num_lights] = acc;
        j->num_lights++;
    }
    for (uint32_t k = 0; k < j->num_lights; ++k) j->light_cdf[k] = j->light_cdf[k] / acc;
    if (j->num_lights) j->light_cdf[j->num_lights - 1] = 1.0f;
    j->light_area = acc;
}

/* Wall time of the last path-tracing pass, BVH build excluded (for timing
 * the oracle as a CPU baseline). */
static double g_render_seconds;

double or_last_render_seconds(void) { return g_render_seconds; }

static double now_seconds(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static int pt_run(const float* pos, const float* nrm, const uint32_t* idx, const uint32_t* tri_mat, const void* mat,
                  uint32_t V, uint32_t T, uint32_t M, const void* cams, uint32_t num_views, const void* settings,
                  float* rgb, uint8_t* rgb8, const float* dimg, float* dcol, int threads, cache_cfg* cc) {
    (void)V;
    const rt_settings* s = (const rt_settings*)settings;
    if (s->spp == 0 || s->max_depth == 0 || s->max_depth > OR_MAX_DEPTH) return -1;
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    pt_job j;
    memset(&j, 0, sizeof(j));
    j.pos = pos; j.nrm = nrm; j.idx = idx; j.tri_mat = tri_mat; j.mat = (const rt_material*)mat; j.M = M;
    j.cams = (const rt_camera*)cams;
    j.num_views = num_views;
    j.s = s;
    j.bvh = or_bvh_build(pos, idx, T, 4);
    /* finalColor *= 2.0f * M_PI * (1.0f / float(raysPerPixel)), path_tracer.cpp:37 */
    j.scale = 2.0 * 3.14159265358979323846 * (double)(1.0f / (float)s->spp);
    j.rgb = rgb; j.rgb8 = rgb8; j.dimg = dimg;
    if (dimg) j.dcol = (double*)calloc((size_t)threads * 3 * M, sizeof(double));
    uint64_t nq[256], nr[256];
    memset(nq, 0, sizeof(nq));
    memset(nr, 0, sizeof(nr));
    j.n_queries = nq;
    j.n_records = nr;
    if (cc) {
        j.cache_rgb = cc->rgb;
        j.q_vertex = cc->q_vertex;
        j.train_every = cc->train_every;
        j.query_on = cc->query_on;
        j.nee = cc->nee;
        j.rec = cc->rec;
        j.ddgi = cc->ddgi;
        j.ddgi_irr = cc->ddgi_irr;
        j.ddgi_dep = cc->ddgi_dep;
        setup_lights(&j, pos, idx, tri_mat, T);
    }
    const double t0 = now_seconds();
    run_parallel(threads, j.cams[0].width * j.cams[0].height * num_views, 16, pt_range, &j);
    g_render_seconds = now_seconds() - t0;
    if (cc) {
        free(j.light_tri);
        free(j.light_cdf);
        cc->queries = cc->records = 0;
        for (int t = 0; t < 256; ++t) {
            cc->queries += nq[t];
            cc->records += nr[t];
        }
    }
    if (dimg) {
        for (uint32_t i = 0; i < 3 * M; ++i) {
            double acc = 0.0;
            for (int t = 0; t < threads; ++t) acc += j.dcol[(size_t)t * 3 * M + i];
            dcol[i] = (float)acc;
        }
        free(j.dcol);
    }
    or_bvh_free(j.bvh);
    return 0;
}

int or_render(const float* pos, const float* nrm, const uint32_t* idx, const uint32_t* tri_mat, const void* mat,
              uint32_t V, uint32_t T, uint32_t M, const void* cam, const void* settings, float* rgb, uint8_t* rgb8,
              int threads) {
    return pt_run(pos, nrm, idx, tri_mat, mat, V, T, M, cam, 1, settings, rgb, rgb8, NULL, NULL, threads, NULL);
}

int or_render_views(const float* pos, const float* nrm, const uint32_t* idx, const uint32_t* tri_mat, const void* mat,
                    uint32_t V, uint32_t T, uint32_t M, const void* cams, uint32_t num_views, const void* settings,
                    float* rgb, uint8_t* rgb8, int threads) {
    return pt_run(pos, nrm, idx, tri_mat, mat, V, T, M, cams, num_views, settings, rgb, rgb8, NULL, NULL, threads, NULL);
}

int or_render_adjoint(const float* pos, const float* nrm, const uint32_t* idx, const uint32_t* tri_mat,
                      const void* mat, uint32_t V, uint32_t T, uint32_t M, const void* cam, const void* settings,
                      const float* d_image, float* d_color, int threads) {
    return pt_run(pos, nrm, idx, tri_mat, mat, V, T, M, cam, 1, settings, NULL, NULL, d_image, d_color, threads, NULL);
}

int or_render_cached(const float* pos, const float* nrm, const uint32_t* idx, const uint32_t* tri_mat,
                     const void* mat, uint32_t V, uint32_t T, uint32_t M, const void* cams, uint32_t num_views,
                     const void* settings, uint32_t q_vertex, uint32_t train_every, int query_on, int nee,
                     const float cache_rgb[3], float* rgb, float* records, uint64_t counts[2], int threads) {
    cache_cfg cc;
    memset(&cc, 0, sizeof(cc));
    cc.rgb = cache_rgb;
    cc.q_vertex = q_vertex;
    cc.train_every = train_every;
    cc.query_on = (uint32_t)(query_on != 0);
    cc.nee = (uint32_t)(nee != 0);
    cc.rec = records;
    int rc = pt_run(pos, nrm, idx, tri_mat, mat, V, T, M, cams, num_views, settings, rgb, NULL, NULL, NULL, threads,
                    &cc);
    if (counts) {
        counts[0] = cc.queries;
        counts[1] = cc.records;
    }
    return rc;
}

/* ------------------------------------------------------------------ */
/* Irradiance probes: the Vulkan renderer's Majercik2019 pipeline
 * (res/shader/majercik2019_probe_raygen.rgen, majercik2019_shading_update_comp.comp,
 * majercik2019_final_raygen.rgen, helper.glsl) in the path tracer's units.
 * Atlas: probe p owns the (side + 2)^2 tile at tile column p mod (cx cy),
 * tile row p div (cx cy); texels 1..side hold the octahedral map, the border
 * the texels across its folded edges. */

/* helper.glsl:79-86 */
static void oct_enc(v3 v, float* ox, float* oy) {
    const float l1 = fabsf(v.x) + fabsf(v.y) + fabsf(v.z);
    const float inv = 1.0f / l1;
    float rx = v.x * inv, ry = v.y * inv;
    if (v.z < 0.0f) {
        const float nx = (1.0f - fabsf(ry)) * (rx >= 0.0f ? 1.0f : -1.0f);
        const float ny = (1.0f - fabsf(rx)) * (ry >= 0.0f ? 1.0f : -1.0f);
        rx = nx;
        ry = ny;
    }
    *ox = rx;
    *oy = ry;
}

/* helper.glsl:88-94 */
static v3 oct_dec(float ox, float oy) {
    v3 v = V3(ox, oy, 1.0f - fabsf(ox) - fabsf(oy));
    if (v.z < 0.0f) {
        const float nx = (1.0f - fabsf(oy)) * (ox >= 0.0f ? 1.0f : -1.0f);
        const float ny = (1.0f - fabsf(ox)) * (oy >= 0.0f ? 1.0f : -1.0f);
        v.x = nx;
        v.y = ny;
    }
    return vnormalize(v);
}

void or_ddgi_oct_encode(const float v[3], float o[2]) { oct_enc(vld(v), &o[0], &o[1]); }
void or_ddgi_oct_decode(const float o[2], float v[3]) {
    const v3 r = oct_dec(o[0], o[1]);
    v[0] = r.x; v[1] = r.y; v[2] = r.z;
}

/* helper.glsl:59-70 sphericalFibonacci */
static v3 fib(uint32_t i, uint32_t n) {
    const float fi = (float)i;
    const float iphi = fi * 2.2360680f * 0.5f + 0.5f - 1.0f;
    float fr = iphi - floorf(iphi);
    if (!(fr < 1.0f)) fr = 0.0f;
    float sn, cs;
    or_sincos_2pi(fr, &sn, &cs);
    const float ct = 1.0f - (2.0f * fi + 1.0f) * (1.0f / (float)n);
    const float st = sqrtf(1.0f - ct * ct);
    return V3(cs * st, sn * st, ct);
}

void or_ddgi_fib(uint32_t i, uint32_t n, float out[3]) {
    const v3 r = fib(i, n);
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}

static uint32_t dg_tiles_x(const rt_ddgi_config* g) { return g->probe_count[0] * g->probe_count[1]; }
static size_t dg_texel(const rt_ddgi_config* g, uint32_t p, uint32_t tx, uint32_t ty) {
    const uint32_t ts = g->side + 2, tiles = dg_tiles_x(g);
    return (size_t)((p / tiles) * ts + ty) * (tiles * ts) + (p % tiles) * ts + tx;
}

/* the interior texel a tile texel holds (mirrored across a folded edge,
 * opposite corner at a corner) */
static void dg_source(uint32_t side, uint32_t tx, uint32_t ty, uint32_t* ix, uint32_t* iy) {
    const uint32_t e = side + 1;
    const int bx = tx == 0 || tx == e, by = ty == 0 || ty == e;
    if (bx && by) {
        *ix = tx == 0 ? side : 1;
        *iy = ty == 0 ? side : 1;
    } else if (bx) {
        *ix = tx == 0 ? 1 : side;
        *iy = e - ty;
    } else if (by) {
        *ix = e - tx;
        *iy = ty == 0 ? 1 : side;
    } else {
        *ix = tx;
        *iy = ty;
    }
}

/* Atlas update (majercik2019_shading_update_comp.comp:25-94) from surfels
 * (8 floats: direction, radiance, distance, hit); irr / dep hold the
 * previous atlases and receive the new ones (blended unless first). */
void or_ddgi_update(const void* cfg, const float* surfels, int first, float* irr, float* dep) {
    const rt_ddgi_config* g = (const rt_ddgi_config*)cfg;
    const uint32_t P = g->probe_count[0] * g->probe_count[1] * g->probe_count[2];
    const uint32_t R = g->rays_per_probe, ts = g->side + 2;
    const float sc = 2.0f / (float)g->side;
    for (uint32_t p = 0; p < P; ++p) {
        const float* sp = surfels + (size_t)p * 8 * R;
        for (uint32_t ty = 0; ty < ts; ++ty) {
            for (uint32_t tx = 0; tx < ts; ++tx) {
                uint32_t ix, iy;
                dg_source(g->side, tx, ty, &ix, &iy);
                const v3 nt = oct_dec(((float)(ix - 1) + 0.5f) * sc - 1.0f, ((float)(iy - 1) + 0.5f) * sc - 1.0f);
                float sr = 0.0f, sg = 0.0f, sb = 0.0f, tw = 0.0f, d0 = 0.0f, d1 = 0.0f, tdw = 0.0f;
                for (uint32_t i = 0; i < R; ++i) {
                    const float* q = sp + 8 * i;
                    float rd = g->max_ray_distance;
                    if (q[7] != 0.0f) rd = fminf(rd, q[6]);
                    const float w = fmaxf(0.0f, vdot(nt, V3(q[0], q[1], q[2])));
                    const float dw = powf(w, g->depth_sharpness);
                    if (w > 1e-6f) {
                        sr = sr + w * q[3];
                        sg = sg + w * q[4];
                        sb = sb + w * q[5];
                        tw = tw + w;
                    }
                    if (dw > 1e-6f) {
                        d0 = d0 + dw * rd;
                        d1 = d1 + dw * (rd * rd);
                        tdw = tdw + dw;
                    }
                }
                float nr = 0.0f, ng = 0.0f, nb = 0.0f, nd0 = 0.0f, nd1 = 0.0f;
                if (tw > 0.0f) { nr = sr / tw; ng = sg / tw; nb = sb / tw; }
                if (tdw > 0.0f) { nd0 = d0 / tdw; nd1 = d1 / tdw; }
                const size_t t = dg_texel(g, p, tx, ty);
                if (!first) {
                    const float h = g->hysteresis, h1 = 1.0f - g->hysteresis;
                    nr = h * irr[4 * t] + h1 * nr;
                    ng = h * irr[4 * t + 1] + h1 * ng;
                    nb = h * irr[4 * t + 2] + h1 * nb;
                    nd0 = h * dep[2 * t] + h1 * nd0;
                    nd1 = h * dep[2 * t + 1] + h1 * nd1;
                }
                irr[4 * t] = nr; irr[4 * t + 1] = ng; irr[4 * t + 2] = nb; irr[4 * t + 3] = 0.0f;
                dep[2 * t] = nd0; dep[2 * t + 1] = nd1;
            }
        }
    }
}

/* bilinear read of the octahedral coordinate (ox, oy) in probe p's tile,
 * texel centres at k + 0.5, the map over tile coordinates [1, side + 1] */
static void dg_bilinear(const rt_ddgi_config* g, uint32_t p, float ox, float oy, size_t t[4], float* ax, float* ay) {
    const float fs = (float)g->side;
    float fx = (ox + 1.0f) * 0.5f * fs + 0.5f;
    float fy = (oy + 1.0f) * 0.5f * fs + 0.5f;
    fx = fminf(fmaxf(fx, 0.0f), fs + 0.5f);
    fy = fminf(fmaxf(fy, 0.0f), fs + 0.5f);
    const float x0 = floorf(fx), y0 = floorf(fy);
    *ax = fx - x0;
    *ay = fy - y0;
    const uint32_t ix = (uint32_t)x0, iy = (uint32_t)y0;
    const uint32_t ix1 = ix + 1 > g->side + 1 ? g->side + 1 : ix + 1;
    const uint32_t iy1 = iy + 1 > g->side + 1 ? g->side + 1 : iy + 1;
    t[0] = dg_texel(g, p, ix, iy);
    t[1] = dg_texel(g, p, ix1, iy);
    t[2] = dg_texel(g, p, ix, iy1);
    t[3] = dg_texel(g, p, ix1, iy1);
}
static float dg_blend(const float* a, int stride, int c, const size_t t[4], float ax, float ay) {
    const float top = (1.0f - ax) * a[stride * t[0] + c] + ax * a[stride * t[1] + c];
    const float bot = (1.0f - ax) * a[stride * t[2] + c] + ax * a[stride * t[3] + c];
    return (1.0f - ay) * top + ay * bot;
}

/* getIrradiance (majercik2019_final_raygen.rgen:43-91) at x with normal n,
 * reached along vd; energy_preservation applied */
static v3 ddgi_lookup(const rt_ddgi_config* g, const float* irr, const float* dep, v3 x, v3 n, v3 vd) {
    const float xs[3] = {x.x, x.y, x.z};
    uint32_t bc[3];
    float al[3];
    for (int k = 0; k < 3; ++k) {
        const float q = (xs[k] - g->probe_start[k]) / g->probe_spacing[k];
        const float top = (float)(g->probe_count[k] - 1);
        bc[k] = q > 0.0f ? (uint32_t)fminf(q, top) : 0u;
        const float bp = g->probe_spacing[k] * (float)bc[k] + g->probe_start[k];
        al[k] = fminf(fmaxf(fabsf(xs[k] - bp) / g->probe_spacing[k], 0.0f), 1.0f);
    }
    const v3 od = V3(-vd.x, -vd.y, -vd.z);
    const v3 bias = vscl(g->normal_bias, vadd(n, vscl(3.0f, od)));
    const int linear = (g->flags & RT_DDGI_LINEAR_BLENDING) != 0;
    const int ref_vis = (g->flags & RT_DDGI_REFERENCE_VISIBILITY) != 0;
    float sr = 0.0f, sg = 0.0f, sb = 0.0f, sw = 0.0f;
    for (uint32_t i = 0; i < 8; ++i) {
        const uint32_t off[3] = {i & 1u, (i >> 1) & 1u, (i >> 2) & 1u};
        uint32_t pc[3];
        for (int k = 0; k < 3; ++k)
            pc[k] = bc[k] + off[k] > g->probe_count[k] - 1 ? g->probe_count[k] - 1 : bc[k] + off[k];
        const uint32_t pidx = pc[0] + pc[1] * g->probe_count[0] + pc[2] * g->probe_count[0] * g->probe_count[1];
        const v3 pp = V3(g->probe_spacing[0] * (float)pc[0] + g->probe_start[0],
                         g->probe_spacing[1] * (float)pc[1] + g->probe_start[1],
                         g->probe_spacing[2] * (float)pc[2] + g->probe_start[2]);
        const v3 ptp = vadd(vsub(x, pp), bias);
        const v3 dtp = vnormalize_f(V3(-ptp.x, -ptp.y, -ptp.z));
        float w = (vdot(dtp, n) + 1.0f) * 0.5f;
        const float tx = off[0] ? al[0] : 1.0f - al[0];
        const float ty = off[1] ? al[1] : 1.0f - al[1];
        const float tz = off[2] ? al[2] : 1.0f - al[2];
        w = w * (tx * ty * tz);
        float ox, oy, ax, ay;
        size_t t[4];
        oct_enc(vnormalize_f(V3(-dtp.x, -dtp.y, -dtp.z)), &ox, &oy);
        dg_bilinear(g, pidx, ox, oy, t, &ax, &ay);
        const float dist = sqrtf(vdot(ptp, ptp));
        const float mean = dg_blend(dep, 2, 0, t, ax, ay);
        const float m2 = dg_blend(dep, 2, 1, t, ax, ay);
        const float var = fabsf(mean * mean - m2);
        const float dd = fmaxf(dist - mean, 0.0f);
        const float den = var + dd * dd;
        float cheb = den > 0.0f ? var / den : 1.0f;
        cheb = fmaxf(cheb * cheb * cheb, 0.0f);
        if (ref_vis ? dist <= mean : dist > mean) w = w * cheb;
        w = fmaxf(0.0001f, w);
        float ir = dg_blend(irr, 4, 0, t, ax, ay), ig = dg_blend(irr, 4, 1, t, ax, ay), ib = dg_blend(irr, 4, 2, t, ax, ay);
        if (!linear) {
            ir = sqrtf(fmaxf(ir, 0.0f));
            ig = sqrtf(fmaxf(ig, 0.0f));
            ib = sqrtf(fmaxf(ib, 0.0f));
        }
        sr = sr + w * ir;
        sg = sg + w * ig;
        sb = sb + w * ib;
        sw = sw + w;
    }
    v3 net = V3(sr / sw, sg / sw, sb / sw);
    if (!linear) net = vmul(net, net);
    return vscl(g->energy_preservation, net);
}

void or_ddgi_eval(const void* cfg, const float* irr, const float* dep, const float* samples, uint32_t n, float* out) {
    const rt_ddgi_config* g = (const rt_ddgi_config*)cfg;
    for (uint32_t i = 0; i < n; ++i) {
        const float* q = samples + 9 * (size_t)i;
        const v3 e = ddgi_lookup(g, irr, dep, vld(q), vld(q + 3), vld(q + 6));
        out[3 * (size_t)i] = e.x; out[3 * (size_t)i + 1] = e.y; out[3 * (size_t)i + 2] = e.z;
    }
}

typedef struct {
    pt_job j;
    const rt_ddgi_config* g;
    const float* rot;
    uint32_t update, multi;
    const float* irr;
    const float* dep;
    float* surf;
} dg_job;

/* One probe ray (majercik2019_probe_raygen.rgen:8-48 in the path tracer's
 * units): miss -> sky without the sun; a light seen from the front or any
 * back face -> 0; a surface -> td albedo / 2 * lookup (multi-bounce) plus the
 * mean of shadow_count light samples, sample k drawing Philox (ray, update,
 * k, 0x50000000); distance from the probe to the hit moved normal_bias along
 * its normal. */
static void dg_range(void* arg, uint32_t b, uint32_t e, int tid) {
    (void)tid;
    const dg_job* d = (const dg_job*)arg;
    const pt_job* j = &d->j;
    const rt_ddgi_config* g = d->g;
    const uint32_t R = g->rays_per_probe, cx = g->probe_count[0], cy = g->probe_count[1];
    const uint32_t key[2] = {(uint32_t)j->s->seed, (uint32_t)(j->s->seed >> 32)};
    const uint32_t count = g->shadow_count ? g->shadow_count : 1;
    for (uint32_t r = b; r < e; ++r) {
        const uint32_t p = r / R, i = r % R;
        const v3 f = fib(i, R);
        const float* m9 = d->rot;
        const v3 dir = V3(m9[0] * f.x + m9[1] * f.y + m9[2] * f.z, m9[3] * f.x + m9[4] * f.y + m9[5] * f.z,
                          m9[6] * f.x + m9[7] * f.y + m9[8] * f.z);
        const uint32_t px = p % cx, py = (p / cx) % cy, pz = p / (cx * cy);
        const v3 o = V3(g->probe_spacing[0] * (float)px + g->probe_start[0],
                        g->probe_spacing[1] * (float)py + g->probe_start[1],
                        g->probe_spacing[2] * (float)pz + g->probe_start[2]);
        rt_ray ray;
        ray.ox = o.x; ray.oy = o.y; ray.oz = o.z; ray.tmin = 0.0f;
        ray.dx = dir.x; ray.dy = dir.y; ray.dz = dir.z; ray.tmax = INFINITY;
        uint64_t a1 = 0, a2 = 0;
        const best_t h = trace_bvh_one(j->bvh, &ray, &a1, &a2);
        v3 rad = V3(0.0f, 0.0f, 0.0f);
        float dist = g->max_ray_distance, hit = 0.0f;
        if (h.id == RT_MISS) {
            if (j->s->flags & RT_RENDER_SKY) rad = vld(j->s->sky);
        } else {
            const v3 pt = vadd(o, vscl(h.t, dir));
            const uint32_t t = h.id;
            const v3 v0 = vld(j->pos + 3 * j->idx[3 * t]);
            const v3 e0 = vsub(vld(j->pos + 3 * j->idx[3 * t + 1]), v0);
            const v3 e1 = vsub(vld(j->pos + 3 * j->idx[3 * t + 2]), v0);
            float b0, b1, b2;
            bary(v0, e0, e1, pt, &b0, &b1, &b2);
            v3 n = V3(0.0f, 0.0f, 0.0f);
            n = vadd(n, vscl(b0, vld(j->nrm + 3 * j->idx[3 * t])));
            n = vadd(n, vscl(b1, vld(j->nrm + 3 * j->idx[3 * t + 1])));
            n = vadd(n, vscl(b2, vld(j->nrm + 3 * j->idx[3 * t + 2])));
            n = vnormalize(n);
            const v3 dv = vsub(o, vadd(pt, vscl(g->normal_bias, n)));
            dist = sqrtf(vdot(dv, dv));
            hit = 1.0f;
            const rt_material* m = &j->mat[j->tri_mat[t]];
            if (!(vdot(n, dir) > 0.0f) && !m->light_source) {
                if (d->multi) {
                    const v3 lk = ddgi_lookup(g, d->irr, d->dep, pt, n, dir);
                    rad = vmul(vscl(0.5f * m->diffuse_threshold, V3(m->color[0], m->color[1], m->color[2])), lk);
                }
                float ar = 0.0f, ag = 0.0f, ab = 0.0f;
                for (uint32_t k = 0; k < count; ++k) {
                    const uint32_t ctr[4] = {r, d->update, k, 0x50000000u};
                    uint32_t rnd[4];
                    or_philox(ctr, key, rnd);
                    rt_ray sh;
                    v3 C;
                    if (!j->nee || !nee_sample(j, pt, n, m, rnd, &sh, &C)) continue;
                    uint64_t c1 = 0, c2 = 0;
                    if (trace_bvh_one(j->bvh, &sh, &c1, &c2).id != RT_MISS) continue;
                    ar = ar + C.x;
                    ag = ag + C.y;
                    ab = ab + C.z;
                }
                const float fc = (float)count;
                rad = V3(rad.x + ar / fc, rad.y + ag / fc, rad.z + ab / fc);
            }
        }
        float* q = d->surf + 8 * (size_t)r;
        q[0] = dir.x; q[1] = dir.y; q[2] = dir.z;
        q[3] = rad.x; q[4] = rad.y; q[5] = rad.z;
        q[6] = dist; q[7] = hit;
    }
}

int or_ddgi_surfels(const float* pos, const float* nrm, const uint32_t* idx, const uint32_t* tri_mat, const void* mat,
                    uint32_t V, uint32_t T, uint32_t M, const void* cfg, const void* settings, const float rot[9],
                    uint32_t update, const float* irr, const float* dep, int multi, float* surfels, int threads) {
    (void)V;
    const rt_ddgi_config* g = (const rt_ddgi_config*)cfg;
    const rt_settings* s = (const rt_settings*)settings;
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    dg_job d;
    memset(&d, 0, sizeof(d));
    d.j.pos = pos; d.j.nrm = nrm; d.j.idx = idx; d.j.tri_mat = tri_mat; d.j.mat = (const rt_material*)mat; d.j.M = M;
    d.j.s = s;
    d.j.bvh = or_bvh_build(pos, idx, T, 4);
    setup_lights(&d.j, pos, idx, tri_mat, T);
    d.j.nee = d.j.num_lights > 0 || ((s->flags & RT_RENDER_SUN) && s->sun_cos < 1.0f);
    d.g = g;
    d.rot = rot;
    d.update = update;
    d.multi = (uint32_t)(multi != 0);
    d.irr = irr;
    d.dep = dep;
    d.surf = surfels;
    const uint32_t n = g->probe_count[0] * g->probe_count[1] * g->probe_count[2] * g->rays_per_probe;
    run_parallel(threads, n, 64, dg_range, &d);
    free(d.j.light_tri);
    free(d.j.light_cdf);
    or_bvh_free(d.j.bvh);
    return 0;
}

int or_render_ddgi(const float* pos, const float* nrm, const uint32_t* idx, const uint32_t* tri_mat, const void* mat,
                   uint32_t V, uint32_t T, uint32_t M, const void* cams, uint32_t num_views, const void* settings,
                   uint32_t q_vertex, const void* cfg, const float* irr, const float* dep, float* rgb,
                   uint64_t* queries, int threads) {
    const rt_settings* s = (const rt_settings*)settings;
    const float zero[3] = {0.0f, 0.0f, 0.0f};
    cache_cfg cc;
    memset(&cc, 0, sizeof(cc));
    cc.rgb = zero;
    cc.q_vertex = q_vertex;
    cc.query_on = 1;
    cc.ddgi = (const rt_ddgi_config*)cfg;
    cc.ddgi_irr = irr;
    cc.ddgi_dep = dep;
    /* light sampling whenever there is something to sample */
    int nl = 0;
    for (uint32_t t = 0; t < T && !nl; ++t) nl = ((const rt_material*)mat)[tri_mat[t]].light_source != 0;
    cc.nee = (uint32_t)(nl || ((s->flags & RT_RENDER_SUN) && s->sun_cos < 1.0f));
    int rc = pt_run(pos, nrm, idx, tri_mat, mat, V, T, M, cams, num_views, settings, rgb, NULL, NULL, NULL, threads,
                    &cc);
    if (queries) *queries = cc.queries;
    return rc;
}

/* ------------------------------------------------------------------ */
/* BidirectionalPathTracer (graphic/bidirectional_path_tracer.cpp:20-168) */

#define OR_LIGHT_SIDE 0x40000000u
#define OR_BD_MAX 16

typedef struct {
    v3 p, n, c;
    int valid;
} bd_vert;

typedef struct {
    const pt_job* j;
    uint32_t vj, lj, maxd;
    uint32_t num_lobj;
    const uint32_t* lobj;      /* {first, count} per light object */
    const uint32_t* lobj_tris;
} bd_job;

/* surface at the closest hit: hit point, normalised interpolated normal */
static void bd_surface(const pt_job* j, v3 o, v3 d, const best_t* h, v3* p, v3* n) {
    *p = vadd(o, vscl(h->t, d));
    uint32_t t = h->id;
    v3 v0 = vld(j->pos + 3 * j->idx[3 * t]);
    v3 e0 = vsub(vld(j->pos + 3 * j->idx[3 * t + 1]), v0);
    v3 e1 = vsub(vld(j->pos + 3 * j->idx[3 * t + 2]), v0);
    float b0, b1, b2;
    bary(v0, e0, e1, *p, &b0, &b1, &b2);
    v3 nn = V3(0.0f, 0.0f, 0.0f);
    nn = vadd(nn, vscl(b0, vld(j->nrm + 3 * j->idx[3 * t])));
    nn = vadd(nn, vscl(b1, vld(j->nrm + 3 * j->idx[3 * t + 1])));
    nn = vadd(nn, vscl(b2, vld(j->nrm + 3 * j->idx[3 * t + 2])));
    *n = vnormalize(nn);
}

/* Scene::isOccluded (scene.cpp:58-83) on the oracle's BVH: the closest hit's
 * squared distance against |end - start|^2 */
static int bd_occluded(const pt_job* j, v3 st, v3 en) {
    v3 d = vsub(en, st);
    float dist2 = vdot(d, d);
    d = vnormalize(d);
    rt_ray r;
    r.ox = st.x; r.oy = st.y; r.oz = st.z; r.tmin = 0.0f;
    r.dx = d.x; r.dy = d.y; r.dz = d.z; r.tmax = INFINITY;
    uint64_t a1 = 0, a2 = 0;
    best_t h = trace_bvh_one(j->bvh, &r, &a1, &a2);
    return h.id != RT_MISS && h.d2 <= dist2;
}

static v3 bdpt_sample(const bd_job* b, uint32_t pix, uint32_t gsmp, v3 dir0) {
    const pt_job* j = b->j;
    const rt_settings* s = j->s;
    const rt_camera* cam = &j->cams[0];
    uint32_t key[2] = {(uint32_t)s->seed, (uint32_t)(s->seed >> 32)};
    bd_vert V[OR_BD_MAX], L[OR_BD_MAX];
    memset(V, 0, sizeof(V));
    memset(L, 0, sizeof(L));
    /* vision path (traceSinglePath, isLightRay = false) */
    v3 o = V3(cam->origin[0], cam->origin[1], cam->origin[2]), d = dir0, T = V3(1.0f, 1.0f, 1.0f);
    int cull = 1, lasthit = 0;
    uint32_t vdepth = 0, ldepth = 0;
    for (uint32_t i = 0; i < b->vj; ++i) {
        rt_ray r;
        r.ox = o.x; r.oy = o.y; r.oz = o.z; r.tmin = 0.0f;
        r.dx = d.x; r.dy = d.y; r.dz = d.z; r.tmax = INFINITY;
        uint64_t a1 = 0, a2 = 0;
        best_t h = trace_bvh_one(j->bvh, &r, &a1, &a2);
        if (h.id == RT_MISS) break;
        v3 p, n;
        bd_surface(j, o, d, &h, &p, &n);
        if (cull && vdot(n, d) > 0.0f) {
            o = vadd(p, vscl(0.01f, d));
            continue;
        }
        cull = 0;
        vdepth = i + 1;
        const rt_material* m = &j->mat[j->tri_mat[h.id]];
        uint32_t ctr[4] = {pix, gsmp, i, 0u}, rnd[4];
        or_philox(ctr, key, rnd);
        float xi0 = u01(rnd[0]);
        v3 prev = d;
        int diffuse = xi0 <= m->diffuse_threshold;
        if (diffuse) d = hemisphere(u01(rnd[1]), u01(rnd[2]), n);
        else if (xi0 <= m->reflect_threshold) d = reflect_ref(d, n);
        else if (xi0 <= m->transparent_threshold) d = custom_refract_ref(d, n, m->refraction_index);
        o = vadd(p, vscl(0.01f, d));
        if (diffuse) T = vmul(T, vscl(vdot(n, d), V3(m->color[0], m->color[1], m->color[2])));
        V[i].p = p; V[i].n = n; V[i].c = T; V[i].valid = 1;
        if (m->light_source && vdot(n, prev) <= 0.0f) {
            lasthit = 1;
            break;
        }
    }
    if (vdepth == 0) return V3(0.0f, 0.0f, 0.0f);
    if (lasthit) return V[vdepth - 1].c;
    /* light point (getRandomLightSourcePoint) and light path (isLightRay = true) */
    if (b->num_lobj > 0) {
        uint32_t ctr[4] = {pix, gsmp, 0u, OR_LIGHT_SIDE}, rnd[4];
        or_philox(ctr, key, rnd);
        uint32_t lo = (uint32_t)(u01(rnd[0]) * (float)b->num_lobj);
        if (lo > b->num_lobj - 1u) lo = b->num_lobj - 1u;
        const float sq = sqrtf(u01(rnd[1])), r2 = u01(rnd[2]);
        const float b0 = 1.0f - sq, b1 = sq * (1.0f - r2), b2 = sq * r2;
        const uint32_t first = b->lobj[2 * lo], count = b->lobj[2 * lo + 1];
        uint32_t ti = (uint32_t)(u01(rnd[3]) * (float)count);
        if (ti > count - 1u) ti = count - 1u;
        const uint32_t t = b->lobj_tris[first + ti];
        const uint32_t i0 = j->idx[3 * t], i1 = j->idx[3 * t + 1], i2 = j->idx[3 * t + 2];
        v3 lp = vadd(vadd(vscl(b0, vld(j->pos + 3 * i0)), vscl(b1, vld(j->pos + 3 * i1))), vscl(b2, vld(j->pos + 3 * i2)));
        v3 ln = V3(0.0f, 0.0f, 0.0f);
        ln = vadd(ln, vscl(b0, vld(j->nrm + 3 * i0)));
        ln = vadd(ln, vscl(b1, vld(j->nrm + 3 * i1)));
        ln = vadd(ln, vscl(b2, vld(j->nrm + 3 * i2)));
        ln = vnormalize(ln);
        const rt_material* lm = &j->mat[j->tri_mat[t]];
        L[0].p = lp; L[0].n = ln; L[0].c = V3(lm->color[0], lm->color[1], lm->color[2]); L[0].valid = 1;
        ctr[2] = 1u;
        or_philox(ctr, key, rnd);
        o = lp;
        d = hemisphere(u01(rnd[1]), u01(rnd[2]), ln);
        T = V3(1.0f, 1.0f, 1.0f);
        for (uint32_t i = 1; i < b->lj; ++i) {
            rt_ray r;
            r.ox = o.x; r.oy = o.y; r.oz = o.z; r.tmin = 0.0f;
            r.dx = d.x; r.dy = d.y; r.dz = d.z; r.tmax = INFINITY;
            uint64_t a1 = 0, a2 = 0;
            best_t h = trace_bvh_one(j->bvh, &r, &a1, &a2);
            if (h.id == RT_MISS) break;
            v3 p, n;
            bd_surface(j, o, d, &h, &p, &n);
            ldepth = i + 1;
            const rt_material* m = &j->mat[j->tri_mat[h.id]];
            uint32_t c2[4] = {pix, gsmp, i + 1u, OR_LIGHT_SIDE}, rn[4];
            or_philox(c2, key, rn);
            float xi0 = u01(rn[0]);
            v3 prev = d;
            int diffuse = xi0 <= m->diffuse_threshold;
            if (diffuse) d = hemisphere(u01(rn[1]), u01(rn[2]), n);
            else if (xi0 <= m->reflect_threshold) d = reflect_ref(d, n);
            else if (xi0 <= m->transparent_threshold) d = custom_refract_ref(d, n, m->refraction_index);
            o = vadd(p, vscl(0.01f, d));
            if (diffuse) T = vmul(T, vscl(vdot(n, prev), V3(m->color[0], m->color[1], m->color[2])));
            L[i].p = p; L[i].n = n; L[i].c = T; L[i].valid = 1;
            if (m->light_source && vdot(n, prev) <= 0.0f) break;
        }
    }
    /* connections, in the reference's loop order */
    v3 color = V3(0.0f, 0.0f, 0.0f);
    int done = 0;
    for (uint32_t vi = 0; vi < vdepth; ++vi)
        for (uint32_t li = 0; li < ldepth; ++li) {
            if (vi + li > b->maxd || !V[vi].valid) continue;
            v3 st = vadd(V[vi].p, vscl(0.01f, V[vi].n));
            v3 en = vadd(L[li].p, vscl(0.01f, L[li].n));
            if (bd_occluded(j, st, en)) continue;
            v3 dir = vnormalize(vsub(en, st));
            color = vadd(color, vscl(vdot(dir, V[vi].n), vmul(V[vi].c, L[li].c)));
            done = 1;
        }
    return done ? vscl(1.0f / (float)b->maxd, color) : V3(0.0f, 0.0f, 0.0f);
}

static void bd_range(void* arg, uint32_t lo, uint32_t hi, int tid) {
    (void)tid;
    const bd_job* b = (const bd_job*)arg;
    const pt_job* j = b->j;
    const rt_camera* cam = &j->cams[0];
    for (uint32_t pix = lo; pix < hi; ++pix) {
        v3 dir = primary(cam, pix % cam->width, pix / cam->width);
        float sum[3] = {0.0f, 0.0f, 0.0f};
        for (uint32_t smp = 0; smp < j->s->spp; ++smp) {
            v3 c = bdpt_sample(b, pix, j->s->sample_offset + smp, dir);
            sum[0] = sum[0] + c.x;
            sum[1] = sum[1] + c.y;
            sum[2] = sum[2] + c.z;
        }
        for (int ch = 0; ch < 3; ++ch) j->rgb[3 * (size_t)pix + ch] = (float)((double)sum[ch] * j->scale);
    }
}

int or_render_bdpt(const float* pos, const float* nrm, const uint32_t* idx, const uint32_t* tri_mat, const void* mat,
                   uint32_t V, uint32_t T, uint32_t M, const void* cam, const void* settings, uint32_t vision_jumps,
                   uint32_t light_jumps, uint32_t max_depth, float* rgb, int threads) {
    (void)V;
    const rt_settings* s = (const rt_settings*)settings;
    if (s->spp == 0 || vision_jumps == 0 || vision_jumps > OR_BD_MAX || light_jumps == 0 || light_jumps > OR_BD_MAX ||
        max_depth == 0)
        return -1;
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    pt_job j;
    memset(&j, 0, sizeof(j));
    j.pos = pos; j.nrm = nrm; j.idx = idx; j.tri_mat = tri_mat; j.mat = (const rt_material*)mat; j.M = M;
    j.cams = (const rt_camera*)cam;
    j.num_views = 1;
    j.s = s;
    j.bvh = or_bvh_build(pos, idx, T, 4);
    j.scale = 2.0 * 3.14159265358979323846 * (double)(1.0f / (float)s->spp);
    j.rgb = rgb;
    /* light objects: every emissive material with its triangles, in order */
    uint32_t* lobj = (uint32_t*)malloc(sizeof(uint32_t) * 2 * (M ? M : 1));
    uint32_t* ltris = (uint32_t*)malloc(sizeof(uint32_t) * (T ? T : 1));
    uint32_t nl = 0, nt = 0;
    for (uint32_t m = 0; m < M; ++m) {
        if (!j.mat[m].light_source) continue;
        uint32_t first = nt;
        for (uint32_t t = 0; t < T; ++t)
            if (tri_mat[t] == m) ltris[nt++] = t;
        if (nt > first) {
            lobj[2 * nl] = first;
            lobj[2 * nl + 1] = nt - first;
            ++nl;
        }
    }
    bd_job b = {&j, vision_jumps, light_jumps, max_depth, nl, lobj, ltris};
    run_parallel(threads, j.cams[0].width * j.cams[0].height, 4, bd_range, &b);
    free(lobj);
    free(ltris);
    or_bvh_free(j.bvh);
    return 0;
}

/* ------------------------------------------------------------------ */
/* Post-filters of the Vulkan renderer (src/graphic/denoiser/). */

/* GaussDenoiser::updateUniforms (gauss_denoiser.cpp:31-47): exp(-(i - (k - 1))^2
 * / (2 sigma^2)), scaled by 1 / sum; exp in double, rounded to float. */
int or_gauss_weights(uint32_t k, float sigma, float* g) {
    if (k == 0 || k > 19 || k % 2 == 0) return -1;
    float sum = 0.0f;
    for (uint32_t i = 0; i < k; ++i) {
        float a = (float)i - ((float)k - 1.0f);
        a = a * a;
        const float b = 2.0f * sigma * sigma;
        const float r = (float)exp((double)(-a / b));
        g[i] = r;
        sum += r;
    }
    sum = 1.0f / sum;
    for (uint32_t i = 0; i < k; ++i) g[i] *= sum;
    return 0;
}

/* denoise_gauss_comp.comp:17-33 on a float RGB image (h rows of w pixels) */
int or_denoise_gauss(const float* in, float* out, uint32_t w, uint32_t h, uint32_t k, float sigma) {
    float g[20];
    if (or_gauss_weights(k, sigma, g)) return -1;
    const int off = (int)(k / 2);
    for (uint32_t py = 0; py < h; ++py)
        for (uint32_t px = 0; px < w; ++px) {
            float c[3] = {0.0f, 0.0f, 0.0f};
            for (int x = -off; x <= off; ++x)
                for (int y = -off; y <= off; ++y) {
                    const int qx = (int)px + x, qy = (int)py + y;
                    if (qx < 0 || qx >= (int)w || qy < 0 || qy >= (int)h) continue;
                    const float* v = in + 3 * ((size_t)qy * w + (size_t)qx);
                    for (int ch = 0; ch < 3; ++ch) c[ch] = c[ch] + v[ch] * g[x + off] * g[y + off];
                }
            for (int ch = 0; ch < 3; ++ch) out[3 * ((size_t)py * w + px) + ch] = c[ch];
        }
    return 0;
}

/* denoise_median_comp.comp:22-46: the shader's exchange network on the 3x3
 * taps v[(dx + 1) * 3 + (dy + 1)] of each channel, taps outside the image 0 */
#define OR_S2(a, b) do { const float t_ = (a); (a) = fminf((a), (b)); (b) = fmaxf(t_, (b)); } while (0)
int or_denoise_median(const float* in, float* out, uint32_t w, uint32_t h) {
    for (uint32_t py = 0; py < h; ++py)
        for (uint32_t px = 0; px < w; ++px)
            for (int ch = 0; ch < 3; ++ch) {
                float v[9];
                for (int dx = -1; dx <= 1; ++dx)
                    for (int dy = -1; dy <= 1; ++dy) {
                        const int qx = (int)px + dx, qy = (int)py + dy;
                        const int inside = qx >= 0 && qx < (int)w && qy >= 0 && qy < (int)h;
                        v[(dx + 1) * 3 + (dy + 1)] = inside ? in[3 * ((size_t)qy * w + (size_t)qx) + ch] : 0.0f;
                    }
                /* mnmx6(v0..v5) */
                OR_S2(v[0], v[3]); OR_S2(v[1], v[4]); OR_S2(v[2], v[5]);
                OR_S2(v[0], v[1]); OR_S2(v[0], v[2]); OR_S2(v[4], v[5]); OR_S2(v[3], v[5]);
                /* mnmx5(v1, v2, v3, v4, v6) */
                OR_S2(v[1], v[2]); OR_S2(v[3], v[4]); OR_S2(v[1], v[3]); OR_S2(v[1], v[6]);
                OR_S2(v[4], v[6]); OR_S2(v[2], v[6]);
                /* mnmx4(v2, v3, v4, v7) */
                OR_S2(v[2], v[3]); OR_S2(v[4], v[7]); OR_S2(v[2], v[4]); OR_S2(v[3], v[7]);
                /* mnmx3(v3, v4, v8) */
                OR_S2(v[4], v[8]); OR_S2(v[3], v[8]); OR_S2(v[3], v[4]);
                out[3 * ((size_t)py * w + px) + ch] = v[4];
            }
    return 0;
}

