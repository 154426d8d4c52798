// scene_io.cpp — OBJ and binary STL readers, scene-file objects, world-space flattening and
// the camera.
#include "scene_io.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <tuple>
#include <unordered_map>

namespace mth {

// ---------------------------------------------------------------- OBJ
// Numbers go through strtod and are rounded to float, as std::atof does in
// obj_loader.cpp:68-105.  Only position / normal / face lines matter for the
// tracer; texture coordinates are parsed to keep vertex identity right.
MeshData load_obj(const std::string& path) {
    std::ifstream in(path);
    if (!in.is_open()) throw InitError("ObjLoader", "Could not load OBJ file \"" + path + "\"!");
    std::vector<Vec3> points, normals;
    std::vector<std::vector<std::tuple<unsigned, unsigned, unsigned>>> faces;
    std::string line;
    auto three = [](std::istringstream& ls) {
        std::string a, b, c;
        std::getline(ls, a, ' ');
        std::getline(ls, b, ' ');
        std::getline(ls, c);
        return Vec3((float)std::strtod(a.c_str(), nullptr), (float)std::strtod(b.c_str(), nullptr),
                    (float)std::strtod(c.c_str(), nullptr));
    };
    while (std::getline(in, line)) {
        std::istringstream ls(line);
        std::string mode;
        std::getline(ls, mode, ' ');
        if (mode == "v") {
            points.push_back(three(ls));
        } else if (mode == "vn") {
            normals.push_back(three(ls));
        } else if (mode == "f") {
            std::vector<std::tuple<unsigned, unsigned, unsigned>> face;
            std::string tok;
            while (std::getline(ls, tok, ' ')) {
                std::istringstream fs(tok);
                std::string p, t, n;
                std::getline(fs, p, '/');
                std::getline(fs, t, '/');
                std::getline(fs, n);
                face.emplace_back((unsigned)std::atoi(p.c_str()), (unsigned)std::atoi(t.c_str()),
                                  (unsigned)std::atoi(n.c_str()));
            }
            faces.push_back(std::move(face));
        }
    }
    MeshData m;
    std::map<std::tuple<unsigned, unsigned, unsigned>, uint32_t> unique;
    for (const auto& f : faces)
        for (const auto& v : f) {
            if (unique.count(v)) continue;
            unsigned pi = std::get<0>(v), ni = std::get<2>(v);
            if (pi == 0 || pi > points.size() || ni == 0 || ni > normals.size())
                throw InitError("ObjLoader", "face index out of range in \"" + path + "\"");
            const Vec3& p = points[pi - 1];
            const Vec3& n = normals[ni - 1];
            for (int k = 0; k < 3; ++k) {
                m.pos.push_back(p[k]);
                m.nrm.push_back(n[k]);
            }
            unique.emplace(v, (uint32_t)(m.pos.size() / 3 - 1));
        }
    for (const auto& f : faces)
        for (size_t i = 2; i < f.size(); ++i) {
            m.idx.push_back(unique[f[0]]);
            m.idx.push_back(unique[f[i - 1]]);
            m.idx.push_back(unique[f[i]]);
        }
    return m;
}

// ---------------------------------------------------------------- objects
Mat4 SceneObject::matrix() const {
    Mat4 m;
    m = mul(m, scale_matrix(scale));
    m = mul(m, rotation.matrix());
    m = mul(m, translation_matrix(position));
    return m;
}

rt_material SceneObject::material() const {
    rt_material r;
    for (int k = 0; k < 3; ++k) r.color[k] = color[k];
    float total = diffuse + reflect + transparent;
    r.diffuse_threshold = diffuse / total;
    r.reflect_threshold = r.diffuse_threshold + (reflect / total);
    r.transparent_threshold = r.reflect_threshold + (transparent / total);
    r.refraction_index = refraction_index;
    r.light_source = light_source ? 1u : 0u;
    return r;
}

rt_scene_desc FlatScene::desc() const {
    rt_scene_desc d;
    d.pos = pos.data();
    d.nrm = nrm.data();
    d.idx = idx.data();
    d.tri_mat = tri_mat.data();
    d.mat = mat.data();
    d.num_vertices = (uint32_t)(pos.size() / 3);
    d.num_triangles = (uint32_t)(idx.size() / 3);
    d.num_materials = (uint32_t)mat.size();
    return d;
}

// Binary STL (STLLoader::load / get_mesh, software_renderer/stl_loader.cpp:
// 18-123): 80-byte header, uint32 count, 50-byte records (facet normal,
// three vertices, attribute), y and z swapped.  The corners meeting at one
// exact position are split into groups: the last corner opens the first
// group, every other corner (in triangle order) joins the first group whose
// opening corner's facet normal is within pi/3 of its own, else opens a new
// one; a group becomes one vertex with the mean of its facet normals.  The
// vertices are numbered in first-seen position order (the reference numbers
// them in its hash map's order); every triangle corner gets the same
// position and normal as in the reference.
MeshData load_stl(const std::string& path) {
    std::ifstream in(path, std::ios::binary);
    if (!in.is_open()) throw InitError("STLLoader", "Could not load STL file \"" + path + "\"!");
    char header[80];
    uint32_t count = 0;
    in.read(header, 80);
    in.read(reinterpret_cast<char*>(&count), 4);
    if (!in) throw InitError("STLLoader", "truncated STL header in \"" + path + "\"");
    in.seekg(0, std::ios::end);
    const uint64_t size = (uint64_t)in.tellg();
    if (size < 84 + 50ull * count) throw InitError("STLLoader", "truncated STL file \"" + path + "\"");
    in.seekg(84);
    struct Tri {
        Vec3 n;
        Vec3 v[3];
    };
    std::vector<Tri> tris(count);
    std::vector<char> rec(50);
    for (uint32_t t = 0; t < count; ++t) {
        in.read(rec.data(), 50);
        float f[12];
        std::memcpy(f, rec.data(), 48);
        for (float x : f)
            if (std::isnan(x)) throw InitError("STLLoader", "NaN coordinate in \"" + path + "\"");
        tris[t].n = Vec3{f[0], f[2], f[1]};
        for (int k = 0; k < 3; ++k) tris[t].v[k] = Vec3{f[3 + 3 * k], f[3 + 3 * k + 2], f[3 + 3 * k + 1]};
    }
    // corners by exact position (float equality: -0 and +0 meet), first-seen order
    struct Key {
        float x, y, z;
        bool operator==(const Key& o) const { return x == o.x && y == o.y && z == o.z; }
    };
    struct KeyHash {
        size_t operator()(const Key& k) const {
            auto h = [](float f) { return f == 0.0f ? (size_t)0 : std::hash<float>()(f); };
            return h(k.x) ^ (h(k.y) << 1) ^ (h(k.z) << 2);
        }
    };
    std::unordered_map<Key, uint32_t, KeyHash> slot;
    std::vector<std::vector<uint32_t>> corners;  // per position: corner ids 3 t + k in triangle order
    std::vector<Vec3> where;
    for (uint32_t t = 0; t < count; ++t)
        for (int k = 0; k < 3; ++k) {
            const Vec3& v = tris[t].v[k];
            auto it = slot.emplace(Key{v[0], v[1], v[2]}, (uint32_t)corners.size());
            if (it.second) {
                corners.emplace_back();
                where.push_back(v);
            }
            corners[it.first->second].push_back(3 * t + (uint32_t)k);
        }
    // within pi/3 (stl_loader.cpp:13-16): Vector::angle = acos(clamp(dot / (|a| |b|))) in double, kept as float
    auto fits = [](const Vec3& a, const Vec3& b) {
        float d = 0.0f, ma = 0.0f, mb = 0.0f;
        for (int k = 0; k < 3; ++k) d += a[k] * b[k];
        for (int k = 0; k < 3; ++k) ma += a[k] * a[k];
        for (int k = 0; k < 3; ++k) mb += b[k] * b[k];
        const double c = (double)d / (std::sqrt((double)ma) * std::sqrt((double)mb));
        const float angle = (float)std::acos(std::clamp(c, -1.0, 1.0));
        return angle < 1.0 / 3.0 * M_PI;
    };
    MeshData m;
    m.idx.assign(3 * (size_t)count, 0);
    uint32_t next = 0;
    for (size_t g = 0; g < corners.size(); ++g) {
        std::vector<uint32_t> pre = corners[g];
        std::vector<std::vector<uint32_t>> groups;
        groups.push_back({pre.back()});
        pre.pop_back();
        for (uint32_t c : pre) {
            bool added = false;
            for (auto& grp : groups)
                if (fits(tris[grp[0] / 3].n, tris[c / 3].n)) {
                    grp.push_back(c);
                    added = true;
                    break;
                }
            if (!added) groups.push_back({c});
        }
        for (const auto& grp : groups) {
            Vec3 n{0.0f, 0.0f, 0.0f};
            for (uint32_t c : grp)
                for (int k = 0; k < 3; ++k) n[k] += tris[c / 3].n[k];
            for (int k = 0; k < 3; ++k) n[k] /= (float)grp.size();
            for (int k = 0; k < 3; ++k) m.pos.push_back(where[g][k]);
            for (int k = 0; k < 3; ++k) m.nrm.push_back(n[k]);
            for (uint32_t c : grp) m.idx[c] = next;
            ++next;
        }
    }
    return m;
}

std::shared_ptr<MeshData> SceneDesc::mesh(const std::string& name, const std::string& mesh_dir) {
    auto it = meshes_.find(name);
    if (it != meshes_.end()) return it->second;
    auto ends = [&](const char* suf) {
        std::string s(suf);
        return name.size() >= s.size() && name.compare(name.size() - s.size(), s.size(), s) == 0;
    };
    std::shared_ptr<MeshData> m;
    // MeshManager::getMesh (mesh_manager.cpp:17-28): .obj from the mesh directory, .stl from its
    // sibling stl directory (the reference's res/obj and res/stl)
    if (ends(".obj")) m = std::make_shared<MeshData>(load_obj(mesh_dir + "/" + name));
    else if (ends(".stl")) m = std::make_shared<MeshData>(load_stl(mesh_dir + "/../stl/" + name));
    else throw InitError("MeshManager", "unsupported mesh format \"" + name + "\"");
    meshes_[name] = m;
    return m;
}

void SceneDesc::add(const SceneObject& o, std::shared_ptr<MeshData> mesh) {
    objects_.push_back(o);
    object_mesh_.push_back(std::move(mesh));
}

void SceneDesc::load(const std::string& scene_path, const std::string& mesh_dir) {
    for (const InputEntry& e : parse_config_file(scene_path)) {
        if (e.name == "Probes") {  // the probe grid, not geometry (mesh_manager.cpp:36-40)
            for (int k = 0; k < 3; ++k) {
                probes_.count[k] = e.get_uint("probeCount", (size_t)k);
                probes_.start[k] = e.get_float("probeStartCorner", (size_t)k);
                probes_.spacing[k] = e.get_float("betweenProbeDistance", (size_t)k);
            }
            probes_.present = true;
            continue;
        }
        SceneObject o;
        o.mesh = e.name;
        o.position = e.get_vec3("position");
        if (e.has("color")) o.color = e.get_vec3("color");
        if (e.has("scale")) o.scale = e.get_vec3("scale");
        if (e.has("rotation")) o.rotation = e.get_rotation("rotation");
        if (e.has("diffuse")) o.diffuse = e.get_float("diffuse");
        if (e.has("reflect")) o.reflect = e.get_float("reflect");
        if (e.has("transparent")) o.transparent = e.get_float("transparent");
        if (e.has("refractionIndex")) o.refraction_index = e.get_float("refractionIndex");
        if (e.has("lightSource")) {
            o.light_source = true;
            o.light_strength = e.get_float("lightSource", 0);
        }
        add(o, mesh(e.name, mesh_dir));
    }
}

FlatScene SceneDesc::flatten() const {
    FlatScene f;
    for (size_t oi = 0; oi < objects_.size(); ++oi) {
        const SceneObject& o = objects_[oi];
        const MeshData& m = *object_mesh_[oi];
        const Mat4 M = o.matrix();
        const uint32_t base = (uint32_t)(f.pos.size() / 3);
        for (size_t v = 0; v < m.pos.size() / 3; ++v) {
            float p4[4] = {m.pos[3 * v], m.pos[3 * v + 1], m.pos[3 * v + 2], 1.0f};
            float n4[4] = {m.nrm[3 * v], m.nrm[3 * v + 1], m.nrm[3 * v + 2], 0.0f};
            float po[4], no[4];
            vec4_mul(M, p4, po);
            vec4_mul(M, n4, no);
            for (int k = 0; k < 3; ++k) {
                f.pos.push_back(po[k]);
                f.nrm.push_back(no[k]);
            }
        }
        for (size_t i = 0; i < m.idx.size(); ++i) f.idx.push_back(base + m.idx[i]);
        for (size_t t = 0; t < m.idx.size() / 3; ++t) f.tri_mat.push_back((uint32_t)oi);
        f.mat.push_back(o.material());
    }
    return f;
}

// ---------------------------------------------------------------- camera
void CameraDesc::parse(const InputEntry& e) {
    position = e.get_vec3("position");
    theta = e.get_float("angle", 0);
    phi = e.get_float("angle", 1);
    theta = (float)(theta * M_PI);
    phi = (float)(phi * -M_PI_2 + M_PI_2);
    if (phi < 0.01f) phi = 0.01f;
    if (phi > M_PI - 0.01f) phi = (float)(M_PI - 0.01f);
    if (theta < -M_PI) theta = (float)(theta + M_PI * 2.0f);
    if (theta > M_PI) theta = (float)(theta - M_PI * 2.0f);
    if (e.has("fov")) fov = (float)(e.get_float("fov") * M_PI / 180.0);  // extension, degrees
}

Mat4 CameraDesc::view() const {
    float sp = (float)std::sin((double)phi), cp = (float)std::cos((double)phi);
    float st = (float)std::sin((double)theta), ct = (float)std::cos((double)theta);
    Vec3 look(sp * ct, cp, sp * st);
    return look_at(position, look + position, Vec3(0.0f, 1.0f, 0.0f));
}

Mat4 CameraDesc::projection(float aspect) const { return perspective(aspect, fov, near_z, far_z, invert_y); }

rt_camera CameraDesc::rt(uint32_t width, uint32_t height) const {
    rt_camera c;
    Mat4 vi = inverse(view());
    Mat4 pi = inverse(projection(float(width) / float(height)));
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            c.view_inv[4 * i + j] = vi.m[i][j];
            c.proj_inv[4 * i + j] = pi.m[i][j];
        }
    float o4[4] = {0.0f, 0.0f, 0.0f, 1.0f}, out[4];
    vec4_mul(vi, o4, out);
    for (int k = 0; k < 3; ++k) c.origin[k] = out[k];
    c.width = width;
    c.height = height;
    return c;
}

}  // namespace mth
