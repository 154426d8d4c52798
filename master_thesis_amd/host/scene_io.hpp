// scene_io.hpp — scene, mesh and camera interface of the software renderer,
// kept as the drop-in boundary, and its flattening into the world-space
// arrays the C ABI takes (include/rt.h rt_scene_desc).
//
//   MeshData          Mesh (graphic/mesh.h:9-29): vertices (pos, normal) + indices
//   load_obj          ObjLoader::load / get_mesh (obj_loader.cpp:19-56,156-184):
//                     v / vn / f lines, unique (pos, tex, normal) vertices in
//                     first-seen order, faces fan-triangulated
//   load_stl          STLLoader::load / get_mesh (stl_loader.cpp:18-123)
//   SceneObject       GraphicsObject's scene-file fields (graphics_object.h:28-52)
//   SceneDesc::load   MeshManager::createObjectsFromFile (mesh_manager.cpp:30-60),
//                     the Probes entry included (ProbeGrid)
//   SceneDesc::flatten GraphicsObject::init (graphics_object.cpp:12-52): v * (S R T)
//                     for positions (w = 1) and normals (w = 0), thresholds
//   CameraDesc        Camera (camera.h:9-27, camera.cpp:13-51) + the matrices
//                     GraphicsEngine::render derives (graphics_engine.cpp:42-49)
#pragma once

#include <cstdint>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/rt.h"
#include "config.hpp"
#include "mth_math.hpp"

namespace mth {

struct MeshData {
    std::vector<float> pos;      // 3 * V (object space)
    std::vector<float> nrm;      // 3 * V
    std::vector<uint32_t> idx;   // 3 * T
};

MeshData load_obj(const std::string& path);
// binary STL (stl_loader.cpp:18-123): y and z swapped, normals smoothed
// across facets within pi/3 of each other
MeshData load_stl(const std::string& path);

struct SceneObject {
    std::string mesh;            // mesh file name as written in the scene file
    Vec3 position;
    Vec3 scale{1.0f, 1.0f, 1.0f};
    Rotation rotation;
    Vec3 color{1.0f, 1.0f, 1.0f};
    bool light_source = false;
    float light_strength = 0.0f;  // parsed, unused by the path tracer (as in the reference)
    float diffuse = 1.0f, reflect = 0.0f, transparent = 0.0f;
    float refraction_index = 1.0f;

    Mat4 matrix() const;         // I * S * R * T (graphics_object.cpp:54-62)
    rt_material material() const;
};

// The scene file's Probes entry (mesh_manager.cpp:36-40): the irradiance
// probe grid of the Majercik2019 renderer.
struct ProbeGrid {
    bool present = false;
    uint32_t count[3] = {0, 0, 0};   // probeCount
    float start[3] = {0, 0, 0};      // probeStartCorner
    float spacing[3] = {0, 0, 0};    // betweenProbeDistance
};

struct FlatScene {
    std::vector<float> pos, nrm;
    std::vector<uint32_t> idx, tri_mat;
    std::vector<rt_material> mat;
    rt_scene_desc desc() const;
};

class SceneDesc {
  public:
    // mesh_dir: directory the scene's mesh files are read from
    void load(const std::string& scene_path, const std::string& mesh_dir);
    void add(const SceneObject& o, std::shared_ptr<MeshData> mesh);
    FlatScene flatten() const;
    const std::vector<SceneObject>& objects() const { return objects_; }
    const ProbeGrid& probes() const { return probes_; }

  private:
    ProbeGrid probes_;
    std::shared_ptr<MeshData> mesh(const std::string& name, const std::string& mesh_dir);
    std::vector<SceneObject> objects_;
    std::vector<std::shared_ptr<MeshData>> object_mesh_;
    std::map<std::string, std::shared_ptr<MeshData>> meshes_;
};

struct CameraDesc {
    Vec3 position{4.0f, 0.0f, 0.0f};
    float theta = (float)M_PI, phi = (float)M_PI_2;
    float fov = (float)M_PI_4;       // the reference fixes pi/4; a camera file may set fov(degrees)
    float near_z = 0.01f, far_z = 1000.0f;
    bool invert_y = true;

    void parse(const InputEntry& e);  // Camera::parseInput, camera.cpp:19-31 (+ optional fov)
    Mat4 view() const;                 // camera.cpp:33-47
    Mat4 projection(float aspect) const;
    rt_camera rt(uint32_t width, uint32_t height) const;
};

}  // namespace mth
