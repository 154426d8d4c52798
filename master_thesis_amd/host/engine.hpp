// engine.hpp — the software renderer's plugin interface and frame loop,
// with the GPU path tracer registered in the same factory.
//
//   Renderer        graphic/renderer.h:9-28: parseInput + renderPixel, plus an
//                   optional frame-level hook renderFrame(); GraphicsEngine
//                   calls the hook instead of its per-pixel thread pool when
//                   a renderer implements it (a GPU renders whole frames).
//   HipPathTracer   PathTracer (graphic/path_tracer.h:9-32) on libmtrt:
//                   same keys visionJumpCount / raysPerPixel, optional seed.
//   HipBidirectionalPathTracer  BidirectionalPathTracer
//                   (graphic/bidirectional_path_tracer.h:9-46): keys
//                   visionJumpCount / lightJumpCount / maxDepth /
//                   raysPerPixel, optional seed.
//   HipMajercik2019 the Vulkan renderer's Majercik2019 (renderer/
//                   majercik2019_renderer.cpp:75-91) on libmtrt's irradiance
//                   probes: same keys; each frame updates the field
//                   (fieldUpdates times, default 1) and renders with it at
//                   the first hit, shadowCountVision samples per pixel.
//   make_renderer   getRenderer (main.cpp:22-26): by the .renderer entry name.
//   make_denoiser   getDenoiser (src/main.cpp:73-77): GaussDenoiser /
//                   MedianDenoiser entries after the renderer's, run on the
//                   GPU (rt_denoise_*) in order after each frame.
//   GraphicsEngine  graphic/graphics_engine.h:17-63: objects -> scene upload
//                   and BVH build (init), render(), saveImage (clamp, 8-bit
//                   truncation, binary PPM).
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "../../include/rt.h"
#include "config.hpp"
#include "scene_io.hpp"

namespace mth {

struct PixelRenderData {
    uint32_t width = 0, height = 0;
    uint32_t px = 0, py = 0;
    const rt_camera* camera = nullptr;
};

struct FrameRenderData {
    rt_ctx* ctx = nullptr;
    const rt_camera* camera = nullptr;
    const ProbeGrid* probes = nullptr;  // the scene's Probes entry
};

class Renderer {
  public:
    virtual ~Renderer() = default;
    virtual void parseInput(const InputEntry& e) = 0;
    // per-pixel path (the reference's); GPU renderers use the frame hook
    virtual Vec3 renderPixel(const PixelRenderData& prd) const;
    // frame hook: renders all pixels into rgb (width*height*3 floats);
    // returns false when the renderer has no frame-level implementation
    virtual bool renderFrame(const FrameRenderData& frd, float* rgb) const;
};

class HipPathTracer : public Renderer {
  public:
    void parseInput(const InputEntry& e) override;
    bool renderFrame(const FrameRenderData& frd, float* rgb) const override;
    rt_settings settings{};
};

class HipBidirectionalPathTracer : public Renderer {
  public:
    void parseInput(const InputEntry& e) override;
    bool renderFrame(const FrameRenderData& frd, float* rgb) const override;
    rt_bdpt_settings settings{};
};

class HipMajercik2019 : public Renderer {
  public:
    ~HipMajercik2019() override;
    void parseInput(const InputEntry& e) override;
    bool renderFrame(const FrameRenderData& frd, float* rgb) const override;
    rt_ddgi_config config{};  // probe grid filled from the scene at the first frame
    rt_settings settings{};   // max_depth visionJumpCount, spp shadowCountVision, seed
    uint32_t field_updates = 1;

  private:
    mutable rt_ddgi* field_ = nullptr;
    mutable uint64_t frame_ = 0;
};

// A post-filter of the Vulkan renderer (src/main.cpp:73-77,163-168): the
// .renderer file's entries after the first, applied in file order to the
// float image before the clamp and 8-bit truncation.
struct DenoiserDesc {
    std::string name;          // "GaussDenoiser" (kernelSize, sigma) or "MedianDenoiser"
    uint32_t kernel_size = 0;
    float sigma = 0.0f;
};
// Throws InitError("getDenoiser not found", name) for unknown names.
DenoiserDesc make_denoiser(const InputEntry& e);

// Throws InitError("getRenderer not found", name) for unknown names.
std::unique_ptr<Renderer> make_renderer(const std::string& name);

class GraphicsEngine {
  public:
    explicit GraphicsEngine(int device = 0);
    ~GraphicsEngine();
    GraphicsEngine(const GraphicsEngine&) = delete;
    GraphicsEngine& operator=(const GraphicsEngine&) = delete;

    uint32_t width = 0, height = 0;
    CameraDesc camera;
    ProbeGrid probes;  // the scene's Probes entry, for renderers that use it
    std::vector<DenoiserDesc> denoisers;  // post-filters, in order

    void init(std::unique_ptr<Renderer> renderer, const FlatScene& scene, unsigned thread_count);
    void render();
    void save_image(const std::string& path) const;
    const std::vector<float>& pixels() const { return rgb_; }
    const std::vector<uint8_t>& image() const { return rgb8_; }
    double last_render_ms = 0.0;

  private:
    rt_ctx* ctx_ = nullptr;
    std::unique_ptr<Renderer> renderer_;
    unsigned threads_ = 1;
    std::vector<float> rgb_;
    std::vector<uint8_t> rgb8_;
};

}  // namespace mth
