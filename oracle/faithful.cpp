// faithful.cpp — the reference software renderer's hot loop restated with its
// cost structure, as the CPU baseline of the path-tracing workloads.
//
// TEST INFRASTRUCTURE: loaded only by tests/ and bench.py's cpu_baseline leg
// (through liboracle.so).  Nothing in the product path calls it.
//
// What it keeps from the reference (paths relative to software_renderer/),
// because those are where the reference spends its time:
//  - two-level binary BVHs (one per object over its triangles, one over the
//    objects) built by agglomerative clustering of the closest pair of box
//    centres on a linked list (math/bounding_volume_hierachy.cpp:9-27,47-75);
//  - queries that return every leaf whose whole root path passes the slab
//    test, recursively, one std::vector per visited node, no t-limit
//    (bounding_volume_hierachy.cpp:81-101; aabb.cpp:39-53);
//  - per object, Moller-Trumbore on every candidate and the squared distance
//    of the hit point against one shared minimum (graphics_object.cpp:64-87,
//    triangle.cpp:21-45); barycentrics and the interpolated normal normalised
//    in double on the winner (scene.cpp:22-56, triangle.cpp:47-55);
//  - one std::mt19937 behind one global mutex for every random number
//    (math/random.cpp:9-35);
//  - PathTracer's estimator (graphic/path_tracer.cpp:18-97) on worker
//    threads that take one pixel at a time from an atomic counter, clamp and
//    truncate to 8 bits (graphic/graphics_engine.cpp:59-97).
// Its images equal the reference's statistically (tests/test_reference.py);
// the seed is explicit so that a baseline run is repeatable.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <list>
#include <mutex>
#include <random>
#include <thread>
#include <vector>

#include "../include/rt.h"
#include "oracle.h"

namespace {

struct F3 {
    float x = 0.0f, y = 0.0f, z = 0.0f;
};
inline F3 f3(float x, float y, float z) { F3 r; r.x = x; r.y = y; r.z = z; return r; }
inline F3 f3p(const float* p) { return f3(p[0], p[1], p[2]); }
inline F3 operator-(F3 a, F3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline F3 operator+(F3 a, F3 b) { return f3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline float dot(F3 a, F3 b) {
    float s = 0.0f;
    s += a.x * b.x;
    s += a.y * b.y;
    s += a.z * b.z;
    return s;
}
inline F3 cross(F3 a, F3 b) { return f3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
inline float at(F3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }
inline F3 normalized(F3 a) {  // vector.h:214-225: magnitude and factor in double
    double m = std::sqrt((double)dot(a, a));
    if (m != 0) {
        double f = 1.0 / m;
        a = f3((float)(a.x * f), (float)(a.y * f), (float)(a.z * f));
    }
    return a;
}

// An axis-aligned box; the first point is padded by 1e-4 (aabb.cpp:4,27-30).
struct Box {
    bool empty = true;
    F3 lo, hi;
    void add(F3 p) {
        if (empty) {
            lo = f3(p.x - 1e-4f, p.y - 1e-4f, p.z - 1e-4f);
            hi = f3(p.x + 1e-4f, p.y + 1e-4f, p.z + 1e-4f);
            empty = false;
        } else {
            lo = f3(std::min(lo.x, p.x), std::min(lo.y, p.y), std::min(lo.z, p.z));
            hi = f3(std::max(hi.x, p.x), std::max(hi.y, p.y), std::max(hi.z, p.z));
        }
    }
    static Box join(const Box& a, const Box& b) {
        Box r;
        r.empty = false;
        r.lo = f3(std::min(a.lo.x, b.lo.x), std::min(a.lo.y, b.lo.y), std::min(a.lo.z, b.lo.z));
        r.hi = f3(std::max(a.hi.x, b.hi.x), std::max(a.hi.y, b.hi.y), std::max(a.hi.z, b.hi.z));
        return r;
    }
    F3 centre() const {
        return f3(0.5f * (lo.x + hi.x), 0.5f * (lo.y + hi.y), 0.5f * (lo.z + hi.z));
    }
};

struct Ray {
    F3 o, d, inv;
};
inline Ray make_ray(F3 o, F3 d) {
    Ray r;
    r.o = o;
    r.d = d;
    r.inv = f3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    return r;
}

inline bool slab(const Box& b, const Ray& r) {
    if (b.empty) return false;
    float tmin = 0.0f, tmax = INFINITY;
    for (int k = 0; k < 3; ++k) {
        const float t1 = (at(b.lo, k) - at(r.o, k)) * at(r.inv, k);
        const float t2 = (at(b.hi, k) - at(r.o, k)) * at(r.inv, k);
        tmin = std::max(tmin, std::min(t1, t2));
        tmax = std::min(tmax, std::max(t1, t2));
    }
    return tmin < tmax;
}

// Binary BVH node with owning child pointers.
struct Node {
    Box box;
    Node* kid[2] = {nullptr, nullptr};
    size_t elem = 0;
    ~Node() {
        delete kid[0];
        delete kid[1];
    }
};

// Repeatedly joins the two list entries whose box centres are closest (first
// minimum in list order); the joined node goes to the end of the list.
Node* cluster(const std::vector<Box>& boxes) {
    std::list<Node*> open;
    for (size_t i = 0; i < boxes.size(); ++i) {
        Node* n = new Node();
        n->box = boxes[i];
        n->elem = i;
        open.push_back(n);
    }
    if (open.empty()) return nullptr;
    while (open.size() > 1) {
        auto best_a = open.end(), best_b = open.end();
        float best = INFINITY;
        for (auto i = open.begin(); i != open.end(); ++i) {
            const F3 ci = (*i)->box.centre();
            for (auto j = std::next(i); j != open.end(); ++j) {
                const F3 dv = ci - (*j)->box.centre();
                const float d2 = dot(dv, dv);
                if (d2 < best) {
                    best = d2;
                    best_a = i;
                    best_b = j;
                }
            }
        }
        Node* n = new Node();
        n->kid[0] = *best_a;
        n->kid[1] = *best_b;
        n->box = Box::join(n->kid[0]->box, n->kid[1]->box);
        open.erase(best_a);
        open.erase(best_b);
        open.push_back(n);
    }
    return open.front();
}

// Every leaf under a path of passing boxes, left subtree first.
std::vector<size_t> candidates(const Node* n, const Ray& r) {
    if (!n || !slab(n->box, r)) return {};
    if (!n->kid[0]) return {n->elem};
    std::vector<size_t> out;
    std::vector<size_t> a = candidates(n->kid[0], r);
    std::vector<size_t> b = candidates(n->kid[1], r);
    out.insert(out.end(), a.begin(), a.end());
    out.insert(out.end(), b.begin(), b.end());
    return out;
}

struct Tri {
    uint32_t v[3];  // object-local vertex indices
    F3 v0, e0, e1;
    float d00, d01, d11, denom;
};

struct Object {
    std::vector<F3> pos, nrm;
    std::vector<Tri> tris;
    Node* bvh = nullptr;
    Box box;
    rt_material mat{};
    ~Object() { delete bvh; }

    // Moller-Trumbore on every candidate; keeps a hit strictly closer (in
    // squared distance of the hit point) than the shared minimum.
    bool trace(const Ray& r, F3& hit, const Tri*& best, float& min_d2) const {
        std::vector<size_t> c = candidates(bvh, r);
        if (c.empty()) return false;
        std::vector<const Tri*> list(c.size());
        for (size_t i = 0; i < c.size(); ++i) list[i] = &tris[c[i]];
        bool any = false;
        for (const Tri* t : list) {
            const F3 h = cross(r.d, t->e1);
            const float a = dot(t->e0, h);
            if (a > -1e-7f && a < 1e-7f) continue;
            const float f = 1.0f / a;
            const F3 s = r.o - t->v0;
            const float u = f * dot(s, h);
            if (u < 0.0f || u > 1.0f) continue;
            const F3 q = cross(s, t->e0);
            const float v = f * dot(r.d, q);
            if (v < 0.0f || u + v > 1.0f) continue;
            const float tt = f * dot(t->e1, q);
            if (!(tt > 1e-7f)) continue;
            const F3 p = r.o + f3(tt * r.d.x, tt * r.d.y, tt * r.d.z);
            const F3 w = r.o - p;
            const float d2 = dot(w, w);
            if (d2 < min_d2) {
                hit = p;
                min_d2 = d2;
                best = t;
                any = true;
            }
        }
        return any;
    }
};

struct World {
    std::vector<Object*> objs;
    Node* top = nullptr;
    ~World() {
        delete top;
        for (Object* o : objs) delete o;
    }
    bool trace(const Ray& r, F3& p, F3& n, const Object*& hit_obj) const {
        std::vector<size_t> c = candidates(top, r);
        if (c.empty()) return false;
        std::vector<const Object*> list(c.size());
        for (size_t i = 0; i < c.size(); ++i) list[i] = objs[c[i]];
        float min_d2 = INFINITY;
        const Tri* t = nullptr;
        hit_obj = nullptr;
        for (const Object* o : list)
            if (o->trace(r, p, t, min_d2)) hit_obj = o;
        if (!hit_obj) return false;
        const F3 e2 = p - t->v0;
        const float d20 = dot(e2, t->e0), d21 = dot(e2, t->e1);
        const float bv = (t->d11 * d20 - t->d01 * d21) / t->denom;
        const float bw = (t->d00 * d21 - t->d01 * d20) / t->denom;
        const float b[3] = {1.0f - bv - bw, bv, bw};
        F3 s = f3(0.0f, 0.0f, 0.0f);
        for (int k = 0; k < 3; ++k) {
            const F3 nk = hit_obj->nrm[t->v[k]];
            s = s + f3(b[k] * nk.x, b[k] * nk.y, b[k] * nk.z);
        }
        n = normalized(s);
        return true;
    }
};

// One generator for every thread, serialised by a mutex (random.cpp:9-14).
struct SharedRng {
    std::mutex m;
    std::mt19937 g;
    std::uniform_real_distribution<float> u{0.0f, 1.0f};
    explicit SharedRng(uint32_t seed) : g(seed) {}
    float draw() {
        std::lock_guard<std::mutex> lk(m);
        return u(g);
    }
    // random.cpp:16-35: theta and the square root through double libm calls
    F3 hemisphere(F3 n) {
        const float theta = 2.0f * M_PI * draw();
        const float u = 2.0f * draw() - 1.0f;
        const float r = std::sqrt((double)(1.0f - u * u));
        const float st = std::sin((double)theta), ct = std::cos((double)theta);
        F3 d = f3(r * ct, r * st, u);
        if (dot(d, n) < 0.0f) d = f3(d.x * -1.0f, d.y * -1.0f, d.z * -1.0f);
        return d;
    }
};

inline F3 offset(F3 p, F3 d) { return p + f3(0.01f * d.x, 0.01f * d.y, 0.01f * d.z); }

// path_tracer.cpp:42-97 for one sample; returns true with the path's colour
// when it ends on a light.
bool sample_path(const World& w, SharedRng& rng, F3 o, F3 d, uint32_t max_depth, F3& colour) {
    F3 c = f3(1.0f, 1.0f, 1.0f);
    bool cull = true, lit = false;
    Ray r = make_ray(o, d);
    for (uint32_t i = 0; i < max_depth; ++i) {
        F3 p, n;
        const Object* obj;
        if (!w.trace(r, p, n, obj)) break;
        if (cull && dot(n, r.d) > 0.0f) {
            r = make_ray(offset(p, r.d), r.d);
            continue;
        }
        cull = false;
        lit = false;
        const float xi = rng.draw();
        const rt_material& m = obj->mat;
        const F3 prev = r.d;
        F3 nd = r.d;
        const float I[3] = {r.d.x, r.d.y, r.d.z}, N[3] = {n.x, n.y, n.z};
        float out[3];
        if (xi <= m.diffuse_threshold) {
            nd = rng.hemisphere(n);
        } else if (xi <= m.reflect_threshold) {
            or_reflect(I, N, out);
            nd = f3p(out);
        } else if (xi <= m.transparent_threshold) {
            or_custom_refract(I, N, m.refraction_index, out);
            nd = f3p(out);
        }
        r = make_ray(offset(p, nd), nd);
        if (xi <= m.diffuse_threshold) {
            const float k = dot(n, r.d);
            c = f3(c.x * (m.color[0] * k), c.y * (m.color[1] * k), c.z * (m.color[2] * k));
        }
        if (m.light_source && dot(n, prev) <= 0.0f) {
            lit = true;
            break;
        }
    }
    colour = c;
    return lit;
}

double g_last_seconds = 0.0;

}  // namespace

extern "C" {

// The reference's renderer in faithful mode on a world-space scene whose
// triangles are grouped into objects (tri_obj: object per triangle; obj_mat:
// material per object).  Renders pixels [0, pixel_limit) of the image in the
// order of an atomic counter taken by `threads` workers, spp samples of at
// most max_depth vertices each, and writes their float values (rgb, W*H*3,
// untouched beyond the limit) and 8-bit values (rgb8, may be NULL).  The
// clustering build is excluded from the time or_faithful_seconds reports.
int or_render_faithful(const float* pos, const float* nrm, const uint32_t* idx, const uint32_t* tri_obj,
                       uint32_t T, const uint32_t* obj_mat, uint32_t O, const rt_material* mats, const void* camv,
                       uint32_t spp, uint32_t max_depth, uint32_t threads, uint32_t seed, uint32_t pixel_limit,
                       float* rgb, uint8_t* rgb8) {
    const rt_camera* cam = (const rt_camera*)camv;
    if (!pos || !nrm || !idx || !tri_obj || !obj_mat || !mats || !cam || !rgb || O == 0 || spp == 0 || threads == 0)
        return -1;
    World w;
    w.objs.resize(O);
    std::vector<std::vector<uint32_t>> remap(O);
    for (uint32_t o = 0; o < O; ++o) {
        w.objs[o] = new Object();
        w.objs[o]->mat = mats[obj_mat[o]];
    }
    for (uint32_t t = 0; t < T; ++t) {
        const uint32_t o = tri_obj[t];
        if (o >= O) return -1;
        Object* ob = w.objs[o];
        Tri tr;
        for (int k = 0; k < 3; ++k) {
            const uint32_t v = idx[3 * t + k];
            tr.v[k] = (uint32_t)ob->pos.size();
            ob->pos.push_back(f3p(pos + 3 * (size_t)v));
            ob->nrm.push_back(f3p(nrm + 3 * (size_t)v));
            ob->box.add(ob->pos.back());
        }
        const F3 a = ob->pos[tr.v[0]], b = ob->pos[tr.v[1]], c = ob->pos[tr.v[2]];
        tr.v0 = a;
        tr.e0 = b - a;
        tr.e1 = c - a;
        tr.d00 = dot(tr.e0, tr.e0);
        tr.d01 = dot(tr.e0, tr.e1);
        tr.d11 = dot(tr.e1, tr.e1);
        tr.denom = tr.d00 * tr.d11 - tr.d01 * tr.d01;
        ob->tris.push_back(tr);
    }
    std::vector<Box> obj_boxes(O);
    for (uint32_t o = 0; o < O; ++o) {
        Object* ob = w.objs[o];
        std::vector<Box> tb(ob->tris.size());
        for (size_t i = 0; i < ob->tris.size(); ++i)
            for (int k = 0; k < 3; ++k) tb[i].add(ob->pos[ob->tris[i].v[k]]);
        ob->bvh = cluster(tb);
        obj_boxes[o] = ob->box;
    }
    w.top = cluster(obj_boxes);

    const uint32_t W = cam->width, H = cam->height;
    const uint32_t limit = std::min<uint64_t>(pixel_limit, (uint64_t)W * H);
    SharedRng rng(seed);
    std::atomic<uint32_t> next(0);
    auto worker = [&]() {
        for (;;) {
            const uint32_t i = next.fetch_add(1);
            if (i >= limit) break;
            float d3[3];
            or_primary_dir(cam->view_inv, cam->proj_inv, i % W, i / W, W, H, d3);
            const F3 o = f3p(cam->origin), d = f3p(d3);
            F3 sum = f3(0.0f, 0.0f, 0.0f);
            for (uint32_t s = 0; s < spp; ++s) {
                F3 c;
                if (sample_path(w, rng, o, d, max_depth, c)) sum = sum + c;
            }
            // finalColor *= 2.0f * M_PI * (1.0f / float(raysPerPixel)): double scale
            const double k = 2.0 * M_PI * (double)(1.0f / (float)spp);
            const float v[3] = {(float)(sum.x * k), (float)(sum.y * k), (float)(sum.z * k)};
            for (int ch = 0; ch < 3; ++ch) {
                rgb[3 * (size_t)i + ch] = v[ch];
                if (rgb8) rgb8[3 * (size_t)i + ch] = (uint8_t)(int)(std::min(std::max(v[ch], 0.0f), 1.0f) * 255.0f);
            }
        }
    };
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<std::thread> pool;
    for (uint32_t k = 1; k < threads; ++k) pool.emplace_back(worker);
    worker();
    for (auto& th : pool) th.join();
    g_last_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return 0;
}

double or_faithful_seconds(void) { return g_last_seconds; }

}  // extern "C"
