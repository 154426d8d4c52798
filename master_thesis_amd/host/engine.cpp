// engine.cpp — Renderer factory, the GPU renderers and the frame loop.
#include "engine.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <thread>

namespace mth {

Vec3 Renderer::renderPixel(const PixelRenderData&) const {
    throw InitError("Renderer", "this renderer has no per-pixel implementation");
}

bool Renderer::renderFrame(const FrameRenderData&, float*) const { return false; }

void HipPathTracer::parseInput(const InputEntry& e) {
    settings = rt_settings{};
    settings.max_depth = e.get_uint("visionJumpCount");
    settings.spp = e.get_uint("raysPerPixel");
    settings.seed = e.has("seed") ? (uint64_t)std::stoull(e.get_string("seed")) : 1ull;
    settings.num_views = 1;
}

bool HipPathTracer::renderFrame(const FrameRenderData& frd, float* rgb) const {
    int rc = rt_render(frd.ctx, frd.camera, &settings, rgb, nullptr, RT_PTR_HOST);
    if (rc != RT_OK) throw InitError("HipPathTracer", rt_last_error(frd.ctx));
    return true;
}

void HipBidirectionalPathTracer::parseInput(const InputEntry& e) {
    settings = rt_bdpt_settings{};
    settings.vision_jumps = e.get_uint("visionJumpCount");
    settings.light_jumps = e.get_uint("lightJumpCount");
    settings.max_depth = e.get_uint("maxDepth");
    settings.spp = e.get_uint("raysPerPixel");
    settings.seed = e.has("seed") ? (uint64_t)std::stoull(e.get_string("seed")) : 1ull;
}

bool HipBidirectionalPathTracer::renderFrame(const FrameRenderData& frd, float* rgb) const {
    int rc = rt_render_bdpt(frd.ctx, frd.camera, &settings, rgb, nullptr, RT_PTR_HOST);
    if (rc != RT_OK) throw InitError("HipBidirectionalPathTracer", rt_last_error(frd.ctx));
    return true;
}

HipMajercik2019::~HipMajercik2019() { rt_ddgi_destroy(field_); }

void HipMajercik2019::parseInput(const InputEntry& e) {
    config = rt_ddgi_config{};
    config.rays_per_probe = e.get_uint("perProbeRayCount");
    config.max_ray_distance = e.get_float("maxProbeRayDistance");
    config.side = e.get_uint("probeSampleSideLength");
    config.depth_sharpness = e.get_float("depthSharpness");
    config.normal_bias = e.get_float("normalBias");
    config.flags = e.get_uint("linearBlending") ? RT_DDGI_LINEAR_BLENDING : 0u;
    config.energy_preservation = e.get_float("energyPreservation");
    config.shadow_count = e.get_uint("shadowCountProbe");
    if (e.has("hysteresis")) config.hysteresis = e.get_float("hysteresis");
    settings = rt_settings{};
    settings.max_depth = e.get_uint("visionJumpCount");
    settings.spp = e.get_uint("shadowCountVision");
    settings.seed = e.has("seed") ? (uint64_t)std::stoull(e.get_string("seed")) : 1ull;
    settings.num_views = 1;
    field_updates = e.has("fieldUpdates") ? e.get_uint("fieldUpdates") : 1u;
}

bool HipMajercik2019::renderFrame(const FrameRenderData& frd, float* rgb) const {
    if (!field_) {
        if (!frd.probes || !frd.probes->present) throw InitError("HipMajercik2019", "the scene has no Probes entry");
        rt_ddgi_config c = config;
        for (int k = 0; k < 3; ++k) {
            c.probe_count[k] = frd.probes->count[k];
            c.probe_start[k] = frd.probes->start[k];
            c.probe_spacing[k] = frd.probes->spacing[k];
        }
        if (rt_ddgi_create(frd.ctx, &c, &field_) != RT_OK) throw InitError("HipMajercik2019", rt_last_error(frd.ctx));
    }
    for (uint32_t k = 0; k < field_updates; ++k) {
        rt_settings u = settings;
        u.seed = settings.seed + frame_++;
        if (rt_ddgi_update(field_, &u) != RT_OK) throw InitError("HipMajercik2019", rt_last_error(frd.ctx));
    }
    if (rt_render_ddgi(frd.ctx, field_, frd.camera, 1, &settings, 0, rgb, nullptr, RT_PTR_HOST, nullptr) != RT_OK)
        throw InitError("HipMajercik2019", rt_last_error(frd.ctx));
    return true;
}

DenoiserDesc make_denoiser(const InputEntry& e) {
    DenoiserDesc d;
    d.name = e.name;
    if (e.name == "GaussDenoiser") {
        d.kernel_size = e.get_uint("kernelSize");
        d.sigma = e.get_float("sigma");
    } else if (e.name != "MedianDenoiser") {
        throw InitError("getDenoiser not found", e.name);
    }
    return d;
}

std::unique_ptr<Renderer> make_renderer(const std::string& name) {
    if (name == "PathTracer") return std::make_unique<HipPathTracer>();
    if (name == "BidirectionalPathTracer") return std::make_unique<HipBidirectionalPathTracer>();
    if (name == "Majercik2019") return std::make_unique<HipMajercik2019>();
    throw InitError("getRenderer not found", name);
}

GraphicsEngine::GraphicsEngine(int device) {
    int rc = rt_create(&ctx_, device);
    if (rc != RT_OK) throw InitError("GraphicsEngine", rt_last_error(nullptr));
}

GraphicsEngine::~GraphicsEngine() { rt_destroy(ctx_); }

void GraphicsEngine::init(std::unique_ptr<Renderer> renderer, const FlatScene& scene, unsigned thread_count) {
    renderer_ = std::move(renderer);
    threads_ = std::max(1u, thread_count);
    rt_scene_desc d = scene.desc();
    if (rt_scene_upload(ctx_, &d) != RT_OK) throw InitError("GraphicsEngine", rt_last_error(ctx_));
    rt_build_opts o{0u, 0u};  // library default leaf size
    if (rt_build(ctx_, &o, nullptr) != RT_OK) throw InitError("GraphicsEngine", rt_last_error(ctx_));
    rgb_.assign((size_t)width * height * 3, 0.0f);
    rgb8_.assign((size_t)width * height * 3, 0);
}

void GraphicsEngine::render() {
    if (!renderer_) throw InitError("GraphicsEngine", "render before init");
    const rt_camera cam = camera.rt(width, height);
    auto t0 = std::chrono::steady_clock::now();
    FrameRenderData frd{ctx_, &cam, &probes};
    if (!renderer_->renderFrame(frd, rgb_.data())) {
        // per-pixel path: workers pull pixels from an atomic counter
        std::atomic<uint32_t> next(0);
        const uint32_t total = width * height;
        auto work = [&]() {
            for (;;) {
                uint32_t i = next.fetch_add(1);
                if (i >= total) break;
                PixelRenderData prd{width, height, i % width, i / width, &cam};
                Vec3 c = renderer_->renderPixel(prd);
                for (int k = 0; k < 3; ++k) rgb_[3 * (size_t)i + k] = c[k];
            }
        };
        std::vector<std::thread> th;
        for (unsigned t = 1; t < threads_; ++t) th.emplace_back(work);
        work();
        for (auto& t : th) t.join();
    }
    if (!denoisers.empty()) {
        std::vector<float> tmp(rgb_.size());
        for (const DenoiserDesc& d : denoisers) {
            int rc = d.name == "GaussDenoiser"
                         ? rt_denoise_gauss(ctx_, rgb_.data(), tmp.data(), width, height, d.kernel_size, d.sigma,
                                            RT_PTR_HOST)
                         : rt_denoise_median(ctx_, rgb_.data(), tmp.data(), width, height, RT_PTR_HOST);
            if (rc != RT_OK) throw InitError(d.name, rt_last_error(ctx_));
            rgb_.swap(tmp);
        }
    }
    last_render_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    // clamp to [0, 1], then (char)(v * 255.0f) (graphics_engine.cpp:90-95)
    for (size_t i = 0; i < rgb_.size(); ++i) {
        float v = std::clamp(rgb_[i], 0.0f, 1.0f);
        rgb8_[i] = (uint8_t)(int)(v * 255.0f);
    }
    std::printf("100%% done\n");
}

void GraphicsEngine::save_image(const std::string& path) const {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) throw InitError("GraphicsEngine", "cannot write \"" + path + "\"");
    std::fprintf(f, "P6\n%u\n%u\n255\n", width, height);
    std::fwrite(rgb8_.data(), 1, rgb8_.size(), f);
    std::fclose(f);
}

}  // namespace mth
