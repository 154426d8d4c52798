/*
 * rt_host.h — C ABI of the host-side scene / camera / config layer
 * (libmtrt_host.so, plain C++, no GPU).  It keeps the reference software
 * renderer's file formats and object model as the drop-in interface and turns
 * them into the world-space arrays rt_scene_upload takes.
 *
 * Status codes are those of rt.h; the message of the last failure on the
 * calling thread is returned by mth_last_error().
 */
#ifndef MT_RT_HOST_H
#define MT_RT_HOST_H

#include <stddef.h>
#include <stdint.h>

#include "rt.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mth_scene mth_scene;

/* Reads a .scene file and its meshes (OBJ from mesh_dir, binary STL from
 * mesh_dir/../stl, as the reference's res/obj and res/stl) and flattens every
 * object to world space.  Replaces MeshManager::createObjectsFromFile
 * (software_renderer/mesh_manager.cpp:30-60) + GraphicsObject::init
 * (graphic/graphics_object.cpp:12-52).  One material per object, in file
 * order; the "Probes" entry (the probe grid) adds no geometry. */
int  mth_scene_load(const char* scene_path, const char* mesh_dir, mth_scene** out);
/* One mesh file as loaded, in its own coordinates (no transform), one
 * default material: OBJ (ObjLoader, obj_loader.cpp:19-56,156-184) or binary
 * STL (STLLoader::load / get_mesh, stl_loader.cpp:18-123: y and z swapped,
 * corners at one position merged into one vertex per group of facets whose
 * normals lie within pi/3 of the group's first, normal = mean facet
 * normal).  Vertex numbering is first-seen; each triangle corner's position
 * and normal equal the reference's.  Free with mth_scene_free. */
int  mth_mesh_load(const char* mesh_path, mth_scene** out);
/* Pointers into the loaded scene, valid until mth_scene_free. */
int  mth_scene_desc(const mth_scene* s, rt_scene_desc* out);
void mth_scene_free(mth_scene* s);

/* Camera from a .camera file (Camera::parseInput, camera.cpp:19-31; an
 * optional fov(degrees) key extends it) and the matrices GraphicsEngine::render
 * derives (graphics_engine.cpp:42-49). */
int  mth_camera_load(const char* camera_path, uint32_t width, uint32_t height, rt_camera* out);
/* Same from values: angle0/angle1 as in the file (theta/pi, phi mapping of
 * camera.cpp:23-31), fov in radians (the reference's is pi/4). */
int  mth_camera_make(const float position[3], float angle0, float angle1, float fov, uint32_t width,
                     uint32_t height, rt_camera* out);

/* Reads a .renderer file: the first entry's name (e.g. "PathTracer") into
 * name (capacity cap) and its PathTracer keys visionJumpCount / raysPerPixel
 * (path_tracer.cpp:13-16; optional seed) into *s. */
int  mth_renderer_load(const char* renderer_path, char* name, size_t cap, rt_settings* s);

/* Binary PPM P6 writer, GraphicsEngine::saveImage (graphics_engine.cpp:32-40). */
int  mth_write_ppm(const char* path, const uint8_t* rgb8, uint32_t width, uint32_t height);

const char* mth_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
