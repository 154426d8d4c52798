// host_capi.cpp — extern "C" surface of the host layer (include/rt_host.h).
#include <cstdio>
#include <cstring>
#include <exception>
#include <string>

#include "../../include/rt_host.h"
#include "config.hpp"
#include "scene_io.hpp"

struct mth_scene {
    mth::SceneDesc desc;
    mth::FlatScene flat;
};

static thread_local std::string g_err;

static int fail(int code, const std::string& m) {
    g_err = m;
    return code;
}

extern "C" {

const char* mth_last_error(void) { return g_err.c_str(); }

int mth_scene_load(const char* scene_path, const char* mesh_dir, mth_scene** out) {
    if (!scene_path || !mesh_dir || !out) return fail(RT_E_INVALID, "mth_scene_load: NULL argument");
    *out = nullptr;
    try {
        auto* s = new mth_scene();
        s->desc.load(scene_path, mesh_dir);
        s->flat = s->desc.flatten();
        if (s->flat.idx.empty()) {
            delete s;
            return fail(RT_E_INVALID, std::string("mth_scene_load: no triangles in \"") + scene_path + "\"");
        }
        *out = s;
        return RT_OK;
    } catch (const std::exception& e) {
        return fail(RT_E_INVALID, e.what());
    }
}

int mth_mesh_load(const char* mesh_path, mth_scene** out) {
    if (!mesh_path || !out) return fail(RT_E_INVALID, "mth_mesh_load: NULL argument");
    *out = nullptr;
    try {
        const std::string p(mesh_path);
        auto ends = [&](const char* suf) {
            const std::string x(suf);
            return p.size() >= x.size() && p.compare(p.size() - x.size(), x.size(), x) == 0;
        };
        mth::MeshData m;
        if (ends(".stl")) m = mth::load_stl(p);
        else if (ends(".obj")) m = mth::load_obj(p);
        else return fail(RT_E_INVALID, "mth_mesh_load: unsupported mesh format \"" + p + "\"");
        auto* s = new mth_scene();
        s->flat.pos = std::move(m.pos);
        s->flat.nrm = std::move(m.nrm);
        s->flat.idx = std::move(m.idx);
        s->flat.tri_mat.assign(s->flat.idx.size() / 3, 0u);
        s->flat.mat.push_back(mth::SceneObject{}.material());
        *out = s;
        return RT_OK;
    } catch (const std::exception& e) {
        return fail(RT_E_INVALID, e.what());
    }
}

int mth_scene_desc(const mth_scene* s, rt_scene_desc* out) {
    if (!s || !out) return fail(RT_E_INVALID, "mth_scene_desc: NULL argument");
    *out = s->flat.desc();
    return RT_OK;
}

void mth_scene_free(mth_scene* s) { delete s; }

int mth_camera_load(const char* path, uint32_t width, uint32_t height, rt_camera* out) {
    if (!path || !out) return fail(RT_E_INVALID, "mth_camera_load: NULL argument");
    if (!width || !height) return fail(RT_E_INVALID, "mth_camera_load: empty image");
    try {
        auto entries = mth::parse_config_file(path);
        if (entries.empty()) return fail(RT_E_INVALID, std::string("mth_camera_load: no entry in \"") + path + "\"");
        mth::CameraDesc c;
        c.parse(entries[0]);
        *out = c.rt(width, height);
        return RT_OK;
    } catch (const std::exception& e) {
        return fail(RT_E_INVALID, e.what());
    }
}

int mth_camera_make(const float position[3], float angle0, float angle1, float fov, uint32_t width,
                    uint32_t height, rt_camera* out) {
    if (!position || !out) return fail(RT_E_INVALID, "mth_camera_make: NULL argument");
    if (!width || !height) return fail(RT_E_INVALID, "mth_camera_make: empty image");
    if (!(fov > 0.0f && fov < 3.14159f)) return fail(RT_E_INVALID, "mth_camera_make: fov outside (0, pi)");
    // same path as a parsed file: values go through the parser's conversions
    mth::InputEntry e("Camera");
    char buf[64];
    std::vector<std::string> p;
    for (int k = 0; k < 3; ++k) {
        snprintf(buf, sizeof(buf), "%.9g", (double)position[k]);
        p.push_back(buf);
    }
    e.insert("position", p);
    std::vector<std::string> a;
    snprintf(buf, sizeof(buf), "%.9g", (double)angle0);
    a.push_back(buf);
    snprintf(buf, sizeof(buf), "%.9g", (double)angle1);
    a.push_back(buf);
    e.insert("angle", a);
    mth::CameraDesc c;
    c.parse(e);
    c.fov = fov;
    *out = c.rt(width, height);
    return RT_OK;
}

int mth_renderer_load(const char* path, char* name, size_t cap, rt_settings* s) {
    if (!path || !name || !cap || !s) return fail(RT_E_INVALID, "mth_renderer_load: NULL argument");
    try {
        auto entries = mth::parse_config_file(path);
        if (entries.empty()) return fail(RT_E_INVALID, std::string("mth_renderer_load: no entry in \"") + path + "\"");
        const mth::InputEntry& e = entries[0];
        std::snprintf(name, cap, "%s", e.name.c_str());
        std::memset(s, 0, sizeof(*s));
        s->max_depth = e.get_uint("visionJumpCount");
        s->spp = e.get_uint("raysPerPixel");
        s->seed = e.has("seed") ? (uint64_t)std::stoull(e.get_string("seed")) : 1ull;
        s->num_views = 1;
        return RT_OK;
    } catch (const std::exception& e) {
        return fail(RT_E_INVALID, e.what());
    }
}

int mth_write_ppm(const char* path, const uint8_t* rgb8, uint32_t width, uint32_t height) {
    if (!path || !rgb8) return fail(RT_E_INVALID, "mth_write_ppm: NULL argument");
    FILE* f = std::fopen(path, "wb");
    if (!f) return fail(RT_E_INVALID, std::string("mth_write_ppm: cannot open \"") + path + "\"");
    std::fprintf(f, "P6\n%u\n%u\n255\n", width, height);
    size_t n = (size_t)width * height * 3;
    size_t w = std::fwrite(rgb8, 1, n, f);
    std::fclose(f);
    if (w != n) return fail(RT_E_INVALID, "mth_write_ppm: short write");
    return RT_OK;
}

}  // extern "C"
