/*
 * rt.h — C ABI of the MI355X (gfx950) software-renderer hot path.
 *
 * The reference renderer (vikkre/master_thesis, software_renderer/) has no
 * C ABI: its hot loop is reached through C++ classes.  Each entry point below
 * replaces one of those interfaces; the comment above it names the reference
 * file:line it stands in for.  Everything here is plain C: pointers, sizes,
 * status codes.  No torch or HIP types cross this boundary (a HIP stream is
 * passed as an opaque void*).
 *
 * Conventions
 *  - Every function returns RT_OK (0) or a negative RT_E_* code.  The message
 *    for the last failure on a context is returned by rt_last_error(ctx);
 *    rt_last_error(NULL) returns the last failure of rt_create on this thread.
 *    Nothing throws across the boundary.
 *  - ptr_kind says where caller buffers live.  RT_PTR_HOST: the call copies
 *    through device staging and returns when the result is in host memory.
 *    RT_PTR_DEVICE: the pointers are device pointers (e.g. torch tensors on
 *    the same GPU); the work is enqueued on the context stream and the call
 *    returns without waiting.
 *  - One context is used by one host thread at a time.  Contexts are
 *    independent of each other.
 *  - Matrices are row-major float[16], m[i][j] at [4*i+j], and act on ROW
 *    vectors (v' = v * M), matching software_renderer/math/matrix.h:245-253.
 */
#ifndef MT_RT_H
#define MT_RT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes -------------------------------------------------------- */
#define RT_OK             0
#define RT_E_INVALID     (-1) /* bad argument (NULL, size 0 where not allowed, out of range) */
#define RT_E_DEVICE      (-2) /* HIP runtime error or no usable gfx950 device              */
#define RT_E_NOMEM       (-3) /* device allocation failed                                  */
#define RT_E_STATE       (-4) /* call order: e.g. trace before rt_build                    */
#define RT_E_UNSUPPORTED (-5) /* feature or size beyond what this build supports           */

#define RT_PTR_HOST   0
#define RT_PTR_DEVICE 1

#define RT_MISS 0xFFFFFFFFu   /* rt_hit.tri of a ray that hit nothing */

typedef struct rt_ctx rt_ctx;  /* opaque context */

/* ---- rays and hits ------------------------------------------------------- */
/* One ray, 32 bytes.  Replaces software_renderer/math/ray.h:6-16 (Ray: origin,
 * direction, cached 1/direction).  tmin/tmax are extensions: a hit is
 * accepted when t > 1e-7 (triangle.cpp:22,38), t > tmin and t < tmax.
 * tmin = 0, tmax = +inf gives the reference's behaviour exactly. */
typedef struct rt_ray {
    float ox, oy, oz, tmin;
    float dx, dy, dz, tmax;
} rt_ray;

/* Closest hit, 16 bytes.  tri is the caller's triangle index (RT_MISS on a
 * miss).  t, u, v are the Moller-Trumbore values of the winning triangle as
 * triangle.cpp:21-45 computes them.  The winner is the triangle whose hit
 * point o + t*d has the smallest squared distance to the ray origin
 * (graphics_object.cpp:75-83); exact ties go to the lowest triangle index. */
typedef struct rt_hit {
    uint32_t tri;
    float t, u, v;
} rt_hit;

/* ---- scene --------------------------------------------------------------- */
/* Per-material record.  Mirrors the GraphicsObject fields the path tracer
 * reads (graphic/graphics_object.h:28-52, thresholds from
 * graphics_object.cpp:48-51). */
typedef struct rt_material {
    float color[3];              /* GraphicsObject::color (albedo; a light's emitted colour) */
    float diffuse_threshold;     /* wd / (wd + wr + wt)                                     */
    float reflect_threshold;     /* diffuse_threshold + wr / (wd + wr + wt)                 */
    float transparent_threshold; /* reflect_threshold + wt / (wd + wr + wt)                 */
    float refraction_index;      /* GraphicsObject::refractionIndex                         */
    uint32_t light_source;       /* GraphicsObject::lightSource                             */
} rt_material;                   /* 32 bytes */

/* World-space triangle mesh, the flattening of every GraphicsObject after
 * GraphicsObject::init (graphics_object.cpp:12-41): positions and normals
 * already multiplied by the object matrix.  Replaces the Scene / GraphicsObject
 * / Mesh hierarchy (graphic/scene.h:9-23) as the geometry the tracer sees. */
typedef struct rt_scene_desc {
    const float*       pos;      /* 3 * num_vertices                          */
    const float*       nrm;      /* 3 * num_vertices (not renormalised)       */
    const uint32_t*    idx;      /* 3 * num_triangles vertex indices          */
    const uint32_t*    tri_mat;  /* num_triangles material indices            */
    const rt_material* mat;      /* num_materials                             */
    uint32_t num_vertices, num_triangles, num_materials;
} rt_scene_desc;

/* ---- acceleration structure --------------------------------------------- */
/* 4-wide BVH node as the build produces it, 128 bytes, structure-of-arrays
 * over the four children, float boxes.  child[k]: RT_CHILD_EMPTY, an inner
 * node index, or a leaf RT_LEAF_BIT | (count-1) << RT_LEAF_COUNT_SHIFT |
 * first (index into the leaf-ordered triangle records).  Node 0 is the root.
 * The traversal reads the 64-byte quantised form below. */
#define RT_CHILD_EMPTY      0xFFFFFFFFu
#define RT_LEAF_BIT         0x80000000u
#define RT_LEAF_COUNT_SHIFT 27
#define RT_LEAF_FIRST_MASK  0x07FFFFFFu
#define RT_MAX_LEAF_SIZE    16u
typedef struct rt_bvh4_node {
    float lo_x[4], hi_x[4];
    float lo_y[4], hi_y[4];
    float lo_z[4], hi_z[4];
    uint32_t child[4];
    uint32_t meta[4];            /* [0] = parent node index, rest reserved */
} rt_bvh4_node;

/* Traversal node, 64 bytes (two per 128-byte cache line): the same node
 * with its children's boxes quantised to 8 bits against the node's own box
 * (Ylitie, Karras and Laine, "Efficient incoherent ray traversal on GPUs
 * through compressed wide BVHs", HPG 2017).  On axis k the plane of byte q
 * is origin[k] + q * 2^exp[k]; low bounds are rounded down and high bounds
 * up, so every quantised box contains the float box it stands for.  An
 * empty child has low 255 and high 0.  child[] as in rt_bvh4_node. */
typedef struct rt_bvh4q_node {
    float origin[3];             /* low corner of the union of the children's boxes */
    int8_t exp[3];               /* per-axis scale 2^exp                              */
    uint8_t pad;
    uint8_t lo_x[4], hi_x[4];
    uint8_t lo_y[4], hi_y[4];
    uint8_t lo_z[4], hi_z[4];
    uint32_t parent, reserved;
    uint32_t child[4];
} rt_bvh4q_node;

/* Leaf-ordered triangle record, 48 bytes: v0, edge0 = v1 - v0,
 * edge1 = v2 - v0 (triangle.cpp:4-17) and the caller's triangle index. */
typedef struct rt_tri_rec {
    float v0[3]; uint32_t id;
    float e0[3]; uint32_t pad0;
    float e1[3]; uint32_t pad1;
} rt_tri_rec;

typedef struct rt_build_opts {
    uint32_t leaf_size;   /* max triangles per leaf, 1..16 (0 = default 4); a subtree this small becomes a leaf when its surface-area cost says so */
    uint32_t flags;       /* 0 or RT_BUILD_CLUSTER */
} rt_build_opts;
/* The default hierarchy is that of the Morton order of the triangles.  This
 * flag clusters them instead (agglomerative, as the reference's builder,
 * in parallel: PLOC): a lower surface-area cost, slower to build; fewer
 * nodes per ray on scenes like the 9.69M-triangle terrain, more on a dense
 * soup of mixed sizes.  Results do not depend on the flag. */
#define RT_BUILD_CLUSTER 1u

typedef struct rt_bvh_info {
    uint32_t num_nodes;       /* BVH4 nodes                                  */
    uint32_t num_leaves;      /* leaf children                               */
    uint32_t depth;           /* BVH4 levels (root = 1)                      */
    uint32_t num_triangles;
    float    build_ms;        /* device time of the last build (HIP events)  */
    uint32_t leaf_size;       /* leaf size the build used                    */
} rt_bvh_info;

/* Per-launch traversal counters (filled when requested; costs one extra
 * atomic per wave). */
typedef struct rt_trace_counters {
    uint64_t rays;
    uint64_t node_visits;     /* traversal nodes fetched (64 B each)         */
    uint64_t tri_tests;       /* triangle records fetched (48 B each)        */
    uint64_t hits;
} rt_trace_counters;

/* ---- camera and render settings ------------------------------------------ */
/* Inverse view and projection matrices and the eye point, as
 * GraphicsEngine::render computes them (graphics_engine.cpp:42-49). */
typedef struct rt_camera {
    float view_inv[16];
    float proj_inv[16];
    float origin[3];
    uint32_t width, height;
} rt_camera;

#define RT_RENDER_DETERMINISTIC 0x1u  /* one path slot per pixel, samples summed in order (bit-reproducible) */
#define RT_RENDER_SKY           0x2u  /* extension: a miss returns throughput * sky (reference: black)     */
#define RT_RENDER_SUN           0x4u  /* extension: plus sun_radiance inside the sun's cone              */

typedef struct rt_settings {
    uint32_t max_depth;       /* PathTracer visionJumpCount (path_tracer.cpp:14)  */
    uint32_t spp;             /* PathTracer raysPerPixel (path_tracer.cpp:15)     */
    uint64_t seed;            /* Philox key                                       */
    uint32_t flags;           /* RT_RENDER_*                                      */
    uint32_t sample_offset;   /* index of the first sample (progressive renders)  */
    float    sky[3];          /* RT_RENDER_SKY radiance                           */
    uint32_t num_views;       /* informational; rt_render_views takes the count   */
    float    sun_dir[3];      /* RT_RENDER_SUN: unit direction towards the sun    */
    float    sun_cos;         /* cosine of the sun's angular radius               */
    float    sun_radiance[3];
} rt_settings;

typedef struct rt_timing {
    float last_kernel_ms;     /* device time of the dominant kernel of the last call   *
                               * (traversal, post-filter; 0 after a render)            */
    float last_call_ms;       /* device time of the whole last call (first to last op) */
    uint32_t launches;        /* kernel launches the last call made                    */
    uint32_t iterations;      /* render: extend/shade iterations of the last call      */
    uint64_t rays_traced;     /* render: closest-hit queries the last call traced (a    *
                               * pixel's fixed primary-ray prefix once per work item)  */
} rt_timing;

/* ---- context ------------------------------------------------------------- */
/* Creates a context on HIP device `device`.  Replaces GraphicsEngine
 * construction (graphic/graphics_engine.cpp:7-10). */
int  rt_create(rt_ctx** out, int device);
/* Frees every device buffer the context owns (graphics_engine.cpp:12-14). */
void rt_destroy(rt_ctx* ctx);
const char* rt_last_error(const rt_ctx* ctx);
/* Enqueue subsequent work on this hipStream_t (NULL = the context's own).
 * Waits for the work already enqueued on the previous stream first. */
int  rt_set_stream(rt_ctx* ctx, void* hip_stream);
int  rt_get_timing(const rt_ctx* ctx, rt_timing* out);

/* Copies a world-space scene to the device.  Replaces GraphicsEngine::init's
 * per-object init + Scene::addObject/init (graphics_engine.cpp:20-30,
 * graphics_object.cpp:12-52, scene.cpp:9-20). */
int  rt_scene_upload(rt_ctx* ctx, const rt_scene_desc* scene);

/* Replaces the colours of all materials (albedo / emission) without a
 * re-upload; used by the differentiable renderer.  3 * num_materials floats. */
int  rt_set_material_colors(rt_ctx* ctx, const float* rgb, int ptr_kind);

/* Builds the 4-wide BVH on the device (Morton codes, radix sort, Karras
 * hierarchy, bottom-up bounds, SAH-guided collapse).  Replaces
 * BVH::init / BVH::rebuild (math/bounding_volume_hierachy.cpp:47-79) for
 * both levels (mesh.cpp:24-37, scene.cpp:13-20). */
int  rt_build(rt_ctx* ctx, const rt_build_opts* opts, rt_bvh_info* info);

/* Copies the built BVH out (for checking).  Either pointer may be NULL to
 * query sizes; *num_nodes / *num_tris are in-out capacities. */
int  rt_bvh_export(rt_ctx* ctx, rt_bvh4_node* nodes, uint32_t* num_nodes,
                   rt_tri_rec* tris, uint32_t* num_tris);

/* Copies the traversal's quantised nodes out (for checking and for
 * replaying the walk on the CPU): *num_nodes is the in-out capacity (nodes
 * NULL: query the count). */
int  rt_bvh_export_traversal(rt_ctx* ctx, rt_bvh4q_node* nodes, uint32_t* num_nodes);

/* Most rays (or segments) one trace call takes: each call hands its rays
 * to the waves in chunks of 128 counted in 32 bits; more returns
 * RT_E_UNSUPPORTED (split the batch). */
#define RT_MAX_RAYS 0xFFFFFF00u

/* Closest hit for n rays (n <= RT_MAX_RAYS).  Replaces Scene::traceRay's
 * geometric part (graphic/scene.cpp:22-41 -> graphics_object.cpp:64-87 ->
 * triangle.cpp:21-45).  counters may be NULL. */
int  rt_trace_closest(rt_ctx* ctx, const rt_ray* rays, uint32_t n,
                      rt_hit* hits, int ptr_kind, rt_trace_counters* counters);

/* Allocation-free closest hit.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same results as
 * rt_trace_closest; device pointers only (rays, hits and status on the
 * context's device), enqueued on the context stream, never waits.  It calls
 * no hipMalloc / hipFree: its traversal scratch (ray counters and per-lane
 * stack spill) is a fixed array of RT_NOALLOC_SCRATCH_BYTES in device memory,
 * one per GPU, held until the process ends; a BVH too deep for it returns
 * RT_E_UNSUPPORTED.  Calls that overlap on one GPU (from any context) share
 * that array and race: order them on one stream.  status (device, may be
 * NULL) receives 0, or 1 / 2 when the tree turned out corrupt (stack or step
 * bound exceeded). */
#define RT_NOALLOC_SCRATCH_BYTES (32u << 20)
int  rt_trace_closest_noalloc(rt_ctx* ctx, const rt_ray* rays, uint32_t n, rt_hit* hits, uint32_t* status);

/* Occlusion: occluded[i] = 1 when ray i hits any triangle with t in
 * (max(tmin, 1e-7), tmax), else 0 (the same acceptance as rt_trace_closest,
 * so occluded[i] == (closest hit != RT_MISS)).  The walk ends at the first
 * accepted triangle.  Used for the shadow rays of rt_render_nrc. */
int  rt_trace_any(rt_ctx* ctx, const rt_ray* rays, uint32_t n, uint8_t* occluded, int ptr_kind);

/* Scene::isOccluded (graphic/scene.cpp:58-83) for n segments, 6 floats each
 * (start, end): the ray from start along normalize(end - start) is occluded
 * when an accepted triangle's hit point lies within squared distance
 * |end - start|^2 of start (inclusive), over all triangles (the reference's
 * own BVH can miss triangles on tight box faces; this query cannot). */
int  rt_occluded(rt_ctx* ctx, const float* segments, uint32_t n, uint8_t* occluded, int ptr_kind);

/* Renders one frame with the diffuse path tracer.  Replaces the frame loop
 * GraphicsEngine::render (graphics_engine.cpp:42-97) calling
 * PathTracer::renderPixel (graphic/path_tracer.cpp:18-97) per pixel.
 * rgb: width*height*3 floats (the unclamped pixel values), may be NULL.
 * rgb8: width*height*3 bytes, clamp to [0,1] then (char)(v*255)
 * (graphics_engine.cpp:90-95), may be NULL. */
int  rt_render(rt_ctx* ctx, const rt_camera* cam, const rt_settings* s,
               float* rgb, uint8_t* rgb8, int ptr_kind);

/* Renders num_views frames of the same scene and settings (cams[v], all of
 * one size) in one launch sequence whose path slots move freely between the
 * views.  rgb / rgb8 hold the views one after another.  View v draws its
 * random numbers with counter (pixel, sample, depth, v), so view 0 equals
 * rt_render.  Extension: the reference renders one camera per process. */
int  rt_render_views(rt_ctx* ctx, const rt_camera* cams, uint32_t num_views, const rt_settings* s,
                     float* rgb, uint8_t* rgb8, int ptr_kind);

/* BidirectionalPathTracer (graphic/bidirectional_path_tracer.cpp:13-168):
 * per sample a vision path of up to vision_jumps vertices (as the path
 * tracer's, recording every vertex), and unless it ends on a light, a light
 * path from a random point of a random light object (object, then triangle
 * uniformly, sqrt-warped barycentrics) of up to light_jumps vertices, then
 * every vision/light vertex pair with vi + li <= max_depth joined by an
 * occlusion test between the points offset 0.01 along their normals (as
 * Scene::isOccluded).  Culled vision slots hold no vertex (the reference
 * reads whatever a previous sample left there).  Random numbers: Philox
 * with counter (pixel, sample, depth, 0) on the vision side, (pixel, sample,
 * k, 2^30) on the light side. */
typedef struct rt_bdpt_settings {
    uint32_t vision_jumps;         /* visionJumpCount, 1..16                    */
    uint32_t light_jumps;          /* lightJumpCount, 1..16                     */
    uint32_t max_depth;            /* maxDepth, 1..32                           */
    uint32_t spp;                  /* raysPerPixel                              */
    uint64_t seed;
    uint32_t sample_offset;
    uint32_t flags;                /* reserved, 0                               */
} rt_bdpt_settings;
int  rt_render_bdpt(rt_ctx* ctx, const rt_camera* cam, const rt_bdpt_settings* s, float* rgb, uint8_t* rgb8,
                    int ptr_kind);

/* Gradient of a scalar image loss with respect to every material colour,
 * by path replay.  d_image: dL/d(rgb) (width*height*3), d_color: output
 * 3*num_materials (overwritten).  Same camera/settings/seed as the forward
 * render give the gradient of that exact render.  No reference counterpart:
 * the reference has no differentiable renderer. */
int  rt_render_adjoint(rt_ctx* ctx, const rt_camera* cam, const rt_settings* s,
                       const float* d_image, float* d_color, int ptr_kind);

/* ---- learned radiance cache ---------------------------------------------
 * No counterpart in the software renderer: it stands in for the Vulkan
 * renderer's DDGI indirect term (res/shader/majercik2019_final_raygen.rgen:
 * 43-91,124).  A sample is 9 floats: hit position, shading normal, incoming
 * ray direction; the cache predicts the RGB expectation of the rest of the
 * path's contribution from that vertex.  Network: 78 encoded inputs (12
 * frequencies x sin/cos x 3 axes of the position in the normalisation box,
 * normal, direction) -> 5 hidden layers of 64 (ReLU) -> 3, BF16 MFMA with
 * FP32 accumulation, relative L2 loss, Adam. */
typedef struct rt_nrc rt_nrc;
typedef struct rt_nrc_config {
    float lr, beta1, beta2, eps;   /* Adam; 0 selects 1e-2, 0.9, 0.99, 1e-8        */
    uint64_t seed;                 /* weight initialisation                       */
    float bounds_lo[3];            /* position normalisation box; lo == hi: the   */
    float bounds_hi[3];            /* uploaded scene's bounds                      */
} rt_nrc_config;

#define RT_NRC_FEATURES 9          /* floats per sample: position, normal, direction */

int  rt_nrc_create(rt_ctx* ctx, const rt_nrc_config* cfg, rt_nrc** out);
/* rt_destroy of the context frees a live cache's device memory; its handle
 * then only accepts rt_nrc_destroy (other calls return RT_E_STATE). */
void rt_nrc_destroy(rt_nrc* nrc);
/* Gets (set = 0) or sets (set = 1, resets Adam) the FP32 parameters; *count
 * is the in-out capacity (params NULL: query the count). */
int  rt_nrc_params(rt_nrc* nrc, float* params, uint32_t* count, int set);
/* Predictions for n samples: feats n*9, out n*3. */
int  rt_nrc_infer(rt_nrc* nrc, const float* feats, uint32_t n, float* out, int ptr_kind);
/* Loss gradient for n samples with targets n*3, without an update: grad is
 * host memory of the parameter count; loss (may be NULL) the mean loss. */
int  rt_nrc_gradient(rt_nrc* nrc, const float* feats, const float* targets, uint32_t n, float* grad,
                     float* loss, int ptr_kind);
/* One Adam step on n samples; loss (may be NULL) receives the mean loss.
 * rt_nrc_infer, rt_nrc_gradient and rt_nrc_train are not timed (no event
 * markers around a short step): rt_get_timing after them reports zeros. */
int  rt_nrc_train(rt_nrc* nrc, const float* feats, const float* targets, uint32_t n, float* loss, int ptr_kind);

/* Path tracing with the cache (rt_render_nrc).  A sample's path runs as in
 * rt_render up to its query_vertex-th front-facing surface hit (1 = the hit
 * after the first bounce); there, unless that hit ends the path on a light,
 * the sample stops and adds throughput * cache(position, normal, direction)
 * instead of tracing the rest.  One sample in train_every (chosen by a hash
 * of pixel, sample and view) is a training sample: it traces its whole path
 * as rt_render does, adds that to the pixel, and yields one training record
 * at its query vertex whose target is the rest of its path's contribution.
 * After the frame the cache takes up to max_steps Adam steps of `batch`
 * records, each batch spread over the whole frame. */
#define RT_NRC_NO_TRAIN 0x1        /* records are made but no step is taken    */
#define RT_NRC_NO_QUERY 0x2        /* every sample traces its whole path        */
#define RT_NRC_NO_RECORDS 0x4      /* no training samples (train_every ignored) */
#define RT_NRC_NO_NEE 0x8          /* no light sampling: the reference estimator */
typedef struct rt_nrc_render {
    uint32_t query_vertex;         /* >= 1; 0 selects 1                         */
    uint32_t train_every;          /* 0 selects 16                              */
    uint32_t batch;                /* multiple of 64; 0 selects 16384           */
    uint32_t max_steps;            /* 0 selects 4                               */
    uint32_t flags;                /* RT_NRC_*                                  */
} rt_nrc_render;
typedef struct rt_nrc_stats {
    uint64_t queries;              /* samples that stopped at the cache         */
    uint64_t records;              /* training records made                     */
    uint32_t steps;                /* Adam steps taken                          */
    float loss;                    /* mean loss of the last step (0 if none)    */
    float render_ms, train_ms;     /* device time of the two phases             */
    uint64_t shadow_rays;          /* light-sample occlusion rays traced        */
} rt_nrc_stats;
/* Renders num_views frames like rt_render_views with the cache (which must
 * belong to ctx); stats may be NULL.  No counterpart in the software
 * renderer (see the cache section above). */
int  rt_render_nrc(rt_ctx* ctx, rt_nrc* nrc, const rt_camera* cams, uint32_t num_views, const rt_settings* s,
                   const rt_nrc_render* opts, float* rgb, uint8_t* rgb8, int ptr_kind, rt_nrc_stats* stats);
/* Copies the training records of the last rt_render_nrc to host memory, in
 * the order of their samples (view, pixel, sample: the order the cache
 * trains on them): feats 9 and targets 3 floats per record; *count is the
 * in-out capacity (both NULL: query the count). */
int  rt_render_nrc_records(rt_ctx* ctx, float* feats, float* targets, uint32_t* count);

/* ---- post-filters ----------------------------------------------------------
 * The Vulkan renderer's denoisers (src/graphic/denoiser/, chained after the
 * renderer by the extra entries of a .renderer file, src/main.cpp:163-168) on
 * float RGB images (width * height * 3, row-major), in and out distinct.
 * Gauss (GaussDenoiser, gauss_denoiser.cpp:31-47 and
 * res/shader/denoise_gauss_comp.comp:17-33): weights g[i] = exp(-(i - (k -
 * 1))^2 / (2 sigma^2)) normalised to sum 1 (centred on the last index, as the
 * reference computes them), out(p) = sum over x, y in [-k/2, k/2] of
 * in(p + (x, y)) g[x + k/2] g[y + k/2], taps outside the image skipped.
 * kernel_size odd, 1..19.  Median (MedianDenoiser, denoise_median_comp.comp:
 * 22-46): the median of the 3x3 neighbourhood per channel by the shader's
 * exchange network, taps outside the image 0. */
int  rt_denoise_gauss_weights(uint32_t kernel_size, float sigma, float* weights);
int  rt_denoise_gauss(rt_ctx* ctx, const float* in, float* out, uint32_t width, uint32_t height,
                      uint32_t kernel_size, float sigma, int ptr_kind);
int  rt_denoise_median(rt_ctx* ctx, const float* in, float* out, uint32_t width, uint32_t height, int ptr_kind);

/* ---- irradiance probes (DDGI) --------------------------------------------
 * The Vulkan renderer's Majercik2019 pipeline (src/graphic/renderer/
 * majercik2019_renderer.cpp:35-47; res/shader/majercik2019_probe_raygen.rgen,
 * majercik2019_shading_update_comp.comp, majercik2019_final_raygen.rgen)
 * restated as a cache beside rt_nrc, in the units of rt_render_nrc:
 *  - update: every probe of a regular grid (the scene file's Probes entry,
 *    mesh_manager.cpp:36-40) casts rays_per_probe rays in spherical
 *    Fibonacci directions (helper.glsl sphericalFibonacci), optionally under
 *    a random rotation per update; a ray's radiance is what a path leaving a
 *    diffuse vertex in that direction adds beyond the vertex's own light
 *    sample: sky without the sun on a miss, 0 on a light or a back face, and
 *    on a surface the mean of shadow_count light samples (plus, with
 *    RT_DDGI_MULTI_BOUNCE, the field's own diffuse term there);
 *  - atlases: per probe a (side + 2)^2 tile of octahedral texels with a
 *    one-texel wrapped border; irradiance = sum max(0, n_t . dir) radiance /
 *    sum of weights, depth moments (r, r^2) weighted by that cosine to the
 *    power depth_sharpness, r = min(hit distance, max_ray_distance)
 *    (majercik2019_shading_update_comp.comp:67-93); hysteresis blends the
 *    previous atlases in;
 *  - lookup (getIrradiance, majercik2019_final_raygen.rgen:44-91): the 8
 *    probes of the cell around the point, trilinear x wrap-shading weight x
 *    Chebyshev visibility from the depth moments, square-root blending
 *    unless RT_DDGI_LINEAR_BLENDING, times energy_preservation.
 * rt_render_ddgi runs rt_render_nrc's path loop with the field in place of
 * the network: at the query vertex (0 = the first surface hit) of a purely
 * diffuse material the sample takes a light sample and adds throughput *
 * albedo / 2 * lookup, the diffuse bounce's expectation in the path
 * tracer's units (it has no training records). */
typedef struct rt_ddgi rt_ddgi;
typedef struct rt_ddgi_config {
    uint32_t probe_count[3];       /* probeCount, each >= 1                     */
    float probe_start[3];          /* probeStartCorner                          */
    float probe_spacing[3];        /* betweenProbeDistance, each > 0            */
    uint32_t rays_per_probe;       /* perProbeRayCount, 1..1024                 */
    uint32_t side;                 /* probeSampleSideLength, 2..32              */
    float max_ray_distance;        /* maxProbeRayDistance, > 0                  */
    float depth_sharpness;         /* depthSharpness                            */
    float normal_bias;             /* normalBias                                */
    float energy_preservation;     /* energyPreservation; 0 selects 1           */
    float hysteresis;              /* previous field's weight in an update, 0..1 */
    uint32_t shadow_count;         /* shadowCountProbe, 1..64; 0 selects 1      */
    uint32_t flags;                /* RT_DDGI_*                                 */
} rt_ddgi_config;
#define RT_DDGI_LINEAR_BLENDING 0x1      /* linearBlending(1)                   */
#define RT_DDGI_MULTI_BOUNCE 0x2         /* hits add the field's diffuse term   */
#define RT_DDGI_ROTATE 0x4               /* random ray-set rotation per update  */
/* Chebyshev weight where the reference's final pass applies it (distance <=
 * mean, majercik2019_final_raygen.rgen:75); default: distance > mean, as in
 * Majercik et al. 2019 */
#define RT_DDGI_REFERENCE_VISIBILITY 0x8
#define RT_DDGI_SURFEL_FLOATS 8          /* direction, radiance, distance, hit */

int  rt_ddgi_create(rt_ctx* ctx, const rt_ddgi_config* cfg, rt_ddgi** out);
/* rt_destroy of the context frees a live field's device memory; its handle
 * then only accepts rt_ddgi_destroy (other calls return RT_E_STATE). */
void rt_ddgi_destroy(rt_ddgi* g);
/* atlas size in texels: width (side + 2) * count_x * count_y, height
 * (side + 2) * count_z; irradiance 4 floats (RGB, 0), depth 2 per texel */
int  rt_ddgi_atlas_size(const rt_ddgi* g, uint32_t* width, uint32_t* height);
/* One update with the scene of the field's context (built): probe rays,
 * light samples at their hits, atlases.  s supplies seed, sky and sun. */
int  rt_ddgi_update(rt_ddgi* g, const rt_settings* s);
/* Atlas update from caller surfels (probes * rays * 8 floats, probe-major). */
int  rt_ddgi_update_surfels(rt_ddgi* g, const float* surfels, int ptr_kind);
/* The last update's surfels and ray-set rotation (row-major 3x3, applied as
 * rot * direction); either may be NULL. */
int  rt_ddgi_surfels(rt_ddgi* g, float* surfels, float* rotation, int ptr_kind);
int  rt_ddgi_get_atlas(rt_ddgi* g, float* irradiance, float* depth, int ptr_kind);
/* Sets both atlases; later updates blend into them. */
int  rt_ddgi_set_atlas(rt_ddgi* g, const float* irradiance, const float* depth, int ptr_kind);
/* Lookups for n samples (9 floats each: position, normal, incoming ray
 * direction) into out (n * 3), energy_preservation included. */
int  rt_ddgi_eval(rt_ddgi* g, const float* samples, uint32_t n, float* out, int ptr_kind);
/* Frames of rt_render_views with the field (which must belong to ctx) at
 * query_vertex (0..63); stats may be NULL (records and steps stay 0). */
int  rt_render_ddgi(rt_ctx* ctx, rt_ddgi* g, const rt_camera* cams, uint32_t num_views, const rt_settings* s,
                    uint32_t query_vertex, float* rgb, uint8_t* rgb8, int ptr_kind, rt_nrc_stats* stats);

#ifdef __cplusplus
}
#endif
#endif /* MT_RT_H */
