/*
 * oracle.h — CPU restatement of the reference software renderer's hot path.
 *
 * TEST INFRASTRUCTURE.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load this library, and only as the checker (or the timed
 * CPU baseline).  The product path (libmtrt.so) never calls it.
 *
 * Every function states the reference file:line it follows (paths relative
 * to the reference's software_renderer/).  The arithmetic is plain float,
 * compiled with -ffp-contract=off, in the reference's operation order.
 *
 * Pinning (what checks each function independently of this restatement) is
 * listed per function; see tests/test_oracle.py.
 */
#ifndef MT_ORACLE_H
#define MT_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Ray / hit / material / BVH4 layouts are those of include/rt.h. */

/* Philox-4x32-10 (Salmon et al. SC'11).  Pinned by the published
 * known-answer vectors (tests/test_oracle.py). */
void or_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* sin/cos of 2*pi*x (x = k*2^-24).  Pinned against libm within 2 ulp and
 * by sin^2 + cos^2 = 1. */
void or_sincos_2pi(float x, float* s, float* c);
/* math/random.cpp:16-35.  Pinned by distribution moments. */
void or_hemisphere(float xi1, float xi2, const float n[3], float out[3]);

/* Triangle ctor + rayIntersects, graphic/triangle.cpp:4-17,21-45.
 * out = {t, u, v}; returns 1 on a hit.  Pinned by closed forms and the
 * compiled reference (oracle/_ref). */
int or_intersect(const float o[3], const float d[3], const float v0[3], const float v1[3], const float v2[3],
                 float out[3]);
/* Triangle::getBarycentricCoords, triangle.cpp:47-55. */
void or_barycentric(const float v0[3], const float v1[3], const float v2[3], const float p[3], float out[3]);

/* Hit point o + t d and interpolated, normalised shading normal of each hit
 * (graphic/scene.cpp:43-53); zeros for misses. */
void or_hit_shading(const float* pos, const float* nrm, const uint32_t* idx, const void* rays, const void* hits,
                    uint32_t n, float* p_out, float* n_out);

/* Closest hit by brute force over all triangles: the definition of the
 * result (graphics_object.cpp:64-87 + scene.cpp:22-41 with the lowest-index
 * tie rule).  rays: rt_ray[n], hits: rt_hit[n]. */
void or_trace_brute(const float* pos, const uint32_t* idx, uint32_t T, const void* rays, uint32_t n, void* hits,
                    int threads);

/* CPU BVH (binary, binned SAH) over the same triangles; traversal with the
 * same conservative culling as the GPU.  Used as the CPU baseline and by the
 * path-tracer restatement.  counters (may be NULL): rays, nodes, tris, hits. */
typedef struct or_bvh or_bvh;
or_bvh* or_bvh_build(const float* pos, const uint32_t* idx, uint32_t T, int leaf_size);
void or_bvh_free(or_bvh* b);
uint32_t or_bvh_num_nodes(const or_bvh* b);
void or_trace_bvh(const or_bvh* b, const void* rays, uint32_t n, void* hits, int threads, uint64_t counters[4]);

/* Traversal of an exported GPU BVH4 (its quantised traversal nodes,
 * rt_bvh4q_node[], and rt_tri_rec[]) with the GPU kernel's exact visiting
 * order and plane arithmetic: same hits and same node / triangle counts as
 * the GPU.  counters: rays, nodes, tris, hits. */
void or_trace_bvh4(const void* nodes, uint32_t num_nodes, const void* tris, uint32_t T, uint32_t depth,
                   const void* rays, uint32_t n, void* hits, int threads, uint64_t counters[4]);
/* The quantised traversal nodes against the build nodes (same children,
 * decoded boxes contain the float boxes, within one step).  0 if sound. */
int or_check_quantized(const void* qnodes, const void* nodes, uint32_t num_nodes);
/* Structural check of an exported BVH4: every triangle in exactly one leaf,
 * every child box contains its subtree.  Returns 0 if sound, else a code. */
int or_check_bvh4(const void* nodes, uint32_t num_nodes, const void* tris, uint32_t T, const float* pos,
                  const uint32_t* idx, uint32_t* out_depth);

/* Row-vector 4x4 product v * M, math/matrix.h:245-253. */
void or_vec4_mat(const float m[16], const float v[4], float out[4]);
/* Primary ray direction, graphic/path_tracer.cpp:19-23. */
void or_primary_dir(const float view_inv[16], const float proj_inv[16], uint32_t px, uint32_t py, uint32_t w,
                    uint32_t h, float out[3]);
/* reflect and customRefract, math/vector.h:319-345, with the reference's
 * float/double promotions; or_sin_acos is customRefract's sin(acos(x)). */
void or_reflect(const float I[3], const float N[3], float out[3]);
void or_custom_refract(const float I[3], const float N[3], float eta, float out[3]);
double or_sin_acos(double x);

/* The diffuse path tracer, graphic/path_tracer.cpp:18-97 driven per pixel as
 * graphic/graphics_engine.cpp:59-97 does (atomic pixel counter, `threads`
 * workers), with Philox (pixel, sample, depth, 0) for random numbers.
 * scene: rt_scene_desc fields.  cam: rt_camera.  s: rt_settings.
 * rgb: W*H*3 floats, rgb8 (may be NULL): W*H*3 bytes. */
int or_render(const float* pos, const float* nrm, const uint32_t* idx, const uint32_t* tri_mat, const void* mat,
              uint32_t V, uint32_t T, uint32_t M, const void* cam, const void* settings, float* rgb, uint8_t* rgb8,
              int threads);
/* The same for num_views cameras (rt_camera[num_views], one size); view v
 * draws with Philox counter (pixel, sample, depth, v); images view-major. */
int or_render_views(const float* pos, const float* nrm, const uint32_t* idx, const uint32_t* tri_mat, const void* mat,
                    uint32_t V, uint32_t T, uint32_t M, const void* cams, uint32_t num_views, const void* settings,
                    float* rgb, uint8_t* rgb8, int threads);
/* Wall time of the last or_render* pass in seconds, BVH build excluded. */
double or_last_render_seconds(void);
/* Path-replay gradient of sum(d_image * image) with respect to every
 * material colour (3*M doubles accumulated, returned as float). */
int or_render_adjoint(const float* pos, const float* nrm, const uint32_t* idx, const uint32_t* tri_mat,
                      const void* mat, uint32_t V, uint32_t T, uint32_t M, const void* cam, const void* settings,
                      const float* d_image, float* d_color, int threads);

/* Scene::isOccluded (graphic/scene.cpp:58-83) for n segments (start, end:
 * 6 floats each) by brute force over all T triangles; out[i] = 0 / 1. */
void or_occluded_brute(const float* pos, const uint32_t* idx, uint32_t T, const float* segs, uint32_t n,
                       uint8_t* out, int threads);

/* The cache mode of rt_render_nrc with a cache that predicts the constant
 * cache_rgb everywhere: a sample stops at its q_vertex-th front-facing hit
 * (unless that hit ends the path on a light, or is at the last depth) and
 * adds throughput * max(cache_rgb, 0), except training samples
 * (train_hash(pixel, sample, view, seed) % train_every == 0), which trace on
 * and record the rest of their path's contribution from that vertex.  nee:
 * light sampling at every vertex (emissive triangles by area, the sun's
 * cone) in place of the diffuse bounce's light hits and sun escapes.
 * records (may be NULL): 13 floats per (view-major pixel, sample) - 9
 * features, 3 targets, 1 if a record was made.  counts (may be NULL):
 * queries, training records. */
int or_render_cached(const float* pos, const float* nrm, const uint32_t* idx, const uint32_t* tri_mat,
                     const void* mat, uint32_t V, uint32_t T, uint32_t M, const void* cams, uint32_t num_views,
                     const void* settings, uint32_t q_vertex, uint32_t train_every, int query_on, int nee,
                     const float cache_rgb[3], float* rgb, float* records, uint64_t counts[2], int threads);

/* Irradiance probes (rt_ddgi_*, the Vulkan renderer's Majercik2019
 * pipeline in the path tracer's units); cfg is an rt_ddgi_config. */
void or_ddgi_oct_encode(const float v[3], float o[2]);
void or_ddgi_oct_decode(const float o[2], float v[3]);
void or_ddgi_fib(uint32_t i, uint32_t n, float out[3]);
/* atlas update from surfels (8 floats each); irr / dep: previous atlases in,
 * new ones out (blended by the hysteresis unless first) */
void or_ddgi_update(const void* cfg, const float* surfels, int first, float* irr, float* dep);
void or_ddgi_eval(const void* cfg, const float* irr, const float* dep, const float* samples, uint32_t n, float* out);
/* surfels of one update: ray directions rot * fib, radiance with the field
 * irr / dep before the update (multi-bounce), update index for Philox */
int or_ddgi_surfels(const float* pos, const float* nrm, const uint32_t* idx, const uint32_t* tri_mat, const void* mat,
                    uint32_t V, uint32_t T, uint32_t M, const void* cfg, const void* settings, const float rot[9],
                    uint32_t update, const float* irr, const float* dep, int multi, float* surfels, int threads);
/* rt_render_ddgi: the cache mode with the field answering the queries */
int or_render_ddgi(const float* pos, const float* nrm, const uint32_t* idx, const uint32_t* tri_mat, const void* mat,
                   uint32_t V, uint32_t T, uint32_t M, const void* cams, uint32_t num_views, const void* settings,
                   uint32_t q_vertex, const void* cfg, const float* irr, const float* dep, float* rgb,
                   uint64_t* queries, int threads);

/* Post-filters of the Vulkan renderer on float RGB images: GaussDenoiser
 * weights and filter, MedianDenoiser's exchange network. */
int or_gauss_weights(uint32_t k, float sigma, float* g);
int or_denoise_gauss(const float* in, float* out, uint32_t w, uint32_t h, uint32_t k, float sigma);
int or_denoise_median(const float* in, float* out, uint32_t w, uint32_t h);

/* BidirectionalPathTracer (graphic/bidirectional_path_tracer.cpp:20-168)
 * for one camera: settings supply spp, seed and sample_offset; Philox
 * counters (pixel, sample, depth, 0) on the vision side and (pixel, sample,
 * k, 2^30) on the light side (k = 0 light point, 1 its direction, i + 1 light
 * vertex i).  Culled vision slots hold no vertex. */
int or_render_bdpt(const float* pos, const float* nrm, const uint32_t* idx, const uint32_t* tri_mat, const void* mat,
                   uint32_t V, uint32_t T, uint32_t M, const void* cam, const void* settings, uint32_t vision_jumps,
                   uint32_t light_jumps, uint32_t max_depth, float* rgb, int threads);

#ifdef __cplusplus
}
#endif
#endif
