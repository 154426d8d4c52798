// ref_driver.cpp — C entry points over the reference software renderer,
// compiled together with the reference's own sources where they lie
// (oracle/Makefile target `ref`, output oracle/_ref/libref.so, this container
// only).  TEST INFRASTRUCTURE: it exists to pin the oracle and the host layer
// against the reference itself.  No reference source is copied; this file
// only calls the reference's classes.
#include <atomic>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include <unistd.h>

#include "camera.h"
#include "graphic/graphics_object.h"
#include "graphic/mesh.h"
#include "graphic/bidirectional_path_tracer.h"
#include "graphic/path_tracer.h"
#include "graphic/scene.h"
#include "graphic/triangle.h"
#include "input_parser.h"
#include "mesh_manager.h"
#include "stl_loader.h"

namespace {

Vector3f v3(const float* p) { return Vector3f({p[0], p[1], p[2]}); }

struct RefScene {
    std::vector<Mesh*> meshes;
    std::vector<GraphicsObject*> objects;
    std::vector<GraphicsObject*> lights;
    MeshManager* manager = nullptr;  // owns meshes and objects of a scene read from a file
    Scene scene;
    ~RefScene() {
        if (manager) {
            delete manager;
            return;
        }
        for (auto* o : objects) delete o;
        for (auto* m : meshes) delete m;
    }
};

InputEntry entry(const std::string& name, const std::vector<std::pair<std::string, std::vector<std::string>>>& kv) {
    InputEntry e(name);
    for (const auto& p : kv) e.insert(p);
    return e;
}

std::string fstr(float x) {
    char b[64];
    snprintf(b, sizeof(b), "%.9g", (double)x);
    return b;
}

}  // namespace

extern "C" {

// Triangle::rayIntersects + getBarycentricCoords (triangle.cpp:21-55).
int ref_intersect(const float* v0, const float* v1, const float* v2, const float* o, const float* d, float* p,
                  float* bary) {
    Triangle t(Vector3u({0u, 1u, 2u}), v3(v0), v3(v1), v3(v2));
    Ray r(v3(o), v3(d));
    Vector3f hp;
    if (!t.rayIntersects(r, hp)) return 0;
    Vector3f b = t.getBarycentricCoords(hp);
    for (int k = 0; k < 3; ++k) {
        p[k] = hp[k];
        bary[k] = b[k];
    }
    return 1;
}

// reflect and customRefract (math/vector.h:319-345), the reference's own
// templates on Vector3f.
void ref_reflect(const float* I, const float* N, float* out) {
    Vector3f r = reflect(v3(I), v3(N));
    for (int k = 0; k < 3; ++k) out[k] = r[k];
}
void ref_custom_refract(const float* I, const float* N, float eta, float* out) {
    Vector3f r = customRefract(v3(I), v3(N), eta);
    for (int k = 0; k < 3; ++k) out[k] = r[k];
}

// A reference Scene from world-space arrays: one Mesh + GraphicsObject per
// material index (identity transform), material thresholds from weights.
// weights (may be NULL: diffuse only): per material diffuse, reflect and
// transparent weight and refraction index, 4 floats, given to the object
// before GraphicsObject::init derives its thresholds (graphics_object.cpp:48-51).
void* ref_scene_create(const float* pos, const float* nrm, const uint32_t* idx, const uint32_t* tri_mat, uint32_t T,
                       const float* colors, const uint32_t* is_light, uint32_t M, const float* weights) {
    auto* s = new RefScene();
    for (uint32_t m = 0; m < M; ++m) {
        Mesh* mesh = new Mesh();
        std::vector<int> remap;
        for (uint32_t t = 0; t < T; ++t) {
            if (tri_mat[t] != m) continue;
            uint32_t loc[3];
            for (int k = 0; k < 3; ++k) {
                uint32_t v = idx[3 * t + k];
                mesh->addVertex(v3(pos + 3 * v), v3(nrm + 3 * v));
                loc[k] = (uint32_t)mesh->vertices.size() - 1;
            }
            mesh->addIndex(Vector3u({loc[0], loc[1], loc[2]}));
        }
        if (mesh->indices.empty()) {
            delete mesh;
            continue;
        }
        mesh->init();
        s->meshes.push_back(mesh);
        auto* o = new GraphicsObject(mesh, Vector3f({0.0f, 0.0f, 0.0f}));
        o->color = v3(colors + 3 * m);
        if (weights) {
            o->diffuseWeight = weights[4 * m + 0];
            o->reflectWeight = weights[4 * m + 1];
            o->transparentWeight = weights[4 * m + 2];
            o->refractionIndex = weights[4 * m + 3];
        }
        if (is_light[m]) {
            o->lightSource = true;
            s->lights.push_back(o);
        }
        o->init();
        s->objects.push_back(o);
        s->scene.addObject(o);
    }
    s->scene.init();
    return s;
}

// A reference Scene read from a .scene file by MeshManager, each object
// initialised and added as GraphicsEngine::init does (graphics_engine.cpp:20-30).
void* ref_scene_load(const char* scene_path, const char* basepath) {
    auto* s = new RefScene();
    s->manager = new MeshManager(basepath);
    s->manager->createObjectsFromFile(scene_path);
    for (GraphicsObject* o : s->manager->getCreatedObjects()) {
        o->init();
        s->objects.push_back(o);
        if (o->lightSource) s->lights.push_back(o);
        s->scene.addObject(o);
    }
    s->scene.init();
    return s;
}

void ref_scene_free(void* h) { delete (RefScene*)h; }

// Scene::traceRay per ray (scene.cpp:22-56): hit flag, position, normal,
// object index (in creation order).
void ref_scene_trace(void* h, const float* rays /* 8 floats each */, uint32_t n, int* hit, float* p, float* nr,
                     int* obj) {
    auto* s = (RefScene*)h;
    for (uint32_t i = 0; i < n; ++i) {
        const float* r = rays + 8 * i;
        Ray ray(Vector3f({r[0], r[1], r[2]}), Vector3f({r[4], r[5], r[6]}));
        Mesh::Vertex hv;
        const GraphicsObject* o = nullptr;
        hit[i] = s->scene.traceRay(ray, hv, o) ? 1 : 0;
        obj[i] = -1;
        if (hit[i]) {
            for (int k = 0; k < 3; ++k) {
                p[3 * i + k] = hv.pos[k];
                nr[3 * i + k] = hv.normal[k];
            }
            for (size_t q = 0; q < s->objects.size(); ++q)
                if (s->objects[q] == o) obj[i] = (int)q;
        }
    }
}

// Scene::isOccluded per segment (scene.cpp:58-83), 6 floats each (start, end).
void ref_scene_occluded(void* h, const float* segs, uint32_t n, uint8_t* out) {
    auto* s = (RefScene*)h;
    for (uint32_t i = 0; i < n; ++i) {
        const float* g = segs + 6 * i;
        out[i] = s->scene.isOccluded(Vector3f({g[0], g[1], g[2]}), Vector3f({g[3], g[4], g[5]})) ? 1 : 0;
    }
}

// Camera::parseInput + GraphicsEngine::render's matrices (camera.cpp:19-51,
// graphics_engine.cpp:42-47).  The reference's fov is fixed at pi/4.
void ref_camera(const float* position, float angle0, float angle1, uint32_t w, uint32_t h, float* view_inv,
                float* proj_inv, float* origin) {
    Camera cam;
    cam.parseInput(entry("Camera", {{"position", {fstr(position[0]), fstr(position[1]), fstr(position[2])}},
                                    {"angle", {fstr(angle0), fstr(angle1)}}}));
    Matrix4f vi = cam.getViewMatrix().inverseMatrix();
    Matrix4f pi = cam.getProjectionMatrix(float(w) / float(h)).inverseMatrix();
    Vector3f o = cutVector(vi * Vector4f({0.0f, 0.0f, 0.0f, 1.0f}));
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            view_inv[4 * i + j] = vi[i][j];
            proj_inv[4 * i + j] = pi[i][j];
        }
    for (int k = 0; k < 3; ++k) origin[k] = o[k];
}

// MeshManager::createObjectsFromFile + GraphicsObject::init: world-space
// vertices of every object (flattened, object order), plus counts.  With
// pos == NULL only the counts are returned.
int ref_load_scene(const char* scene_path, const char* basepath, float* pos, float* nrm, uint32_t* idx,
                   uint32_t* tri_obj, float* colors, uint32_t* counts /* V, T, O */) {
    MeshManager mm(basepath);
    mm.createObjectsFromFile(scene_path);
    std::vector<GraphicsObject*> objs = mm.getCreatedObjects();
    uint32_t V = 0, T = 0;
    for (size_t oi = 0; oi < objs.size(); ++oi) {
        GraphicsObject* o = objs[oi];
        o->init();
        if (pos) {
            for (size_t v = 0; v < o->vertices.size(); ++v)
                for (int k = 0; k < 3; ++k) {
                    pos[3 * (V + v) + k] = o->vertices[v].pos[k];
                    nrm[3 * (V + v) + k] = o->vertices[v].normal[k];
                }
            for (size_t i = 0; i < o->mesh->indices.size(); ++i) idx[3 * T + i] = V + (uint32_t)o->mesh->indices[i];
            for (size_t t = 0; t < o->mesh->indices.size() / 3; ++t) tri_obj[T + t] = (uint32_t)oi;
            for (int k = 0; k < 3; ++k) colors[3 * oi + k] = o->color[k];
        }
        V += (uint32_t)o->vertices.size();
        T += (uint32_t)(o->mesh->indices.size() / 3);
    }
    counts[0] = V;
    counts[1] = T;
    counts[2] = (uint32_t)objs.size();
    return 0;
}

// STLLoader::load + get_mesh (stl_loader.cpp:18-123) on <workdir>/../res/stl/<name>
// (the loader opens "../res/stl/" + name relative to the working directory,
// which is switched to workdir for the call).  Vertices in the loader's
// order; with pos == NULL only the counts {V, T}.
int ref_stl_load(const char* workdir, const char* name, float* pos, float* nrm, uint32_t* idx, uint32_t* counts) {
    char prev[4096];
    if (!getcwd(prev, sizeof(prev))) return -1;
    if (chdir(workdir) != 0) return -2;
    Mesh* m = nullptr;
    try {
        STLLoader loader("");
        loader.load(name);
        m = loader.get_mesh();
    } catch (...) {
        m = nullptr;
    }
    if (chdir(prev) != 0) return -3;
    if (!m) return -4;
    counts[0] = (uint32_t)m->vertices.size();
    counts[1] = (uint32_t)(m->indices.size() / 3);
    if (pos) {
        for (size_t v = 0; v < m->vertices.size(); ++v)
            for (int k = 0; k < 3; ++k) {
                pos[3 * v + k] = m->vertices[v].pos[k];
                nrm[3 * v + k] = m->vertices[v].normal[k];
            }
        for (size_t i = 0; i < m->indices.size(); ++i) idx[i] = (uint32_t)m->indices[i];
    }
    delete m;
    return 0;
}

// PathTracer::renderPixel over a whole image (path_tracer.cpp:18-40), driven
// by `threads` workers pulling pixels from an atomic counter like
// GraphicsEngine::render.  The float pixel values, unclamped.
void ref_render(void* h, const float* view_inv, const float* proj_inv, const float* origin, uint32_t w,
                uint32_t hgt, uint32_t spp, uint32_t depth, int threads, float* rgb) {
    auto* s = (RefScene*)h;
    PathTracer pt;
    pt.parseInput(entry("PathTracer", {{"visionJumpCount", {std::to_string(depth)}},
                                       {"raysPerPixel", {std::to_string(spp)}}}));
    Renderer::PixelRenderData prd;
    prd.scene = &s->scene;
    prd.objects = &s->objects;
    prd.lightSources = &s->lights;
    prd.imageSize = Vector2u({w, hgt});
    prd.origin = v3(origin);
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            prd.viewInverse[i][j] = view_inv[4 * i + j];
            prd.projInverse[i][j] = proj_inv[4 * i + j];
        }
    std::atomic<uint32_t> next(0);
    auto work = [&]() {
        Renderer::PixelRenderData local = prd;
        for (;;) {
            uint32_t i = next.fetch_add(1);
            if (i >= w * hgt) break;
            local.pixel = Vector2u({i % w, i / w});
            Vector3f c = pt.renderPixel(local);
            for (int k = 0; k < 3; ++k) rgb[3 * i + k] = c[k];
        }
    };
    std::vector<std::thread> th;
    for (int t = 1; t < threads; ++t) th.emplace_back(work);
    work();
    for (auto& t : th) t.join();
}

// BidirectionalPathTracer::renderPixel over the image (graphic/
// bidirectional_path_tracer.cpp:13-18 keys), `threads` workers.
void ref_render_bdpt(void* h, const float* view_inv, const float* proj_inv, const float* origin, uint32_t w,
                     uint32_t hgt, uint32_t spp, uint32_t vision, uint32_t light, uint32_t maxd, int threads,
                     float* rgb) {
    auto* s = (RefScene*)h;
    BidirectionalPathTracer bt;
    bt.parseInput(entry("BidirectionalPathTracer", {{"visionJumpCount", {std::to_string(vision)}},
                                                    {"lightJumpCount", {std::to_string(light)}},
                                                    {"maxDepth", {std::to_string(maxd)}},
                                                    {"raysPerPixel", {std::to_string(spp)}}}));
    Renderer::PixelRenderData prd;
    prd.scene = &s->scene;
    prd.objects = &s->objects;
    prd.lightSources = &s->lights;
    prd.imageSize = Vector2u({w, hgt});
    prd.origin = v3(origin);
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            prd.viewInverse[i][j] = view_inv[4 * i + j];
            prd.projInverse[i][j] = proj_inv[4 * i + j];
        }
    std::atomic<uint32_t> next(0);
    auto work = [&]() {
        Renderer::PixelRenderData local = prd;
        for (;;) {
            uint32_t i = next.fetch_add(1);
            if (i >= w * hgt) break;
            local.pixel = Vector2u({i % w, i / w});
            Vector3f c = bt.renderPixel(local);
            for (int k = 0; k < 3; ++k) rgb[3 * i + k] = c[k];
        }
    };
    std::vector<std::thread> th;
    for (int t = 1; t < threads; ++t) th.emplace_back(work);
    work();
    for (auto& t : th) t.join();
}

}  // extern "C"
