// sw_render — command line compatible with the reference's SoftwareRenderer
// (software_renderer/main.cpp:28-87):
//
//   sw_render renderer scene image_width image_height thread_count camera resultimage
//
// The renderer is chosen by the first entry name of the .renderer file
// ("PathTracer" -> the GPU path tracer, "BidirectionalPathTracer" -> the GPU
// bidirectional tracer, "Majercik2019" -> the GPU irradiance probes with the
// scene's Probes grid); later entries (GaussDenoiser, MedianDenoiser) filter
// the image in order, as the Vulkan renderer chains them.  Meshes are read from
// $MTH_MESH_DIR, else from <scene dir>/../obj (the reference's res/ layout).
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <string>

#include "config.hpp"
#include "engine.hpp"
#include "scene_io.hpp"

int main(int argc, char* argv[]) {
    if (argc != 8) {
        std::cout << "Error: wrong paramter count!" << std::endl;
        std::cout << "Usage: SoftwareRenderer renderer scene image_width image_height thread_count camera resultimage"
                  << std::endl;
        return -1;
    }
    std::cout << argv[0];
    for (int i = 1; i < 7; ++i) std::cout << " " << argv[i];
    std::cout << std::endl;
    const std::string renderer_path = argv[1], scene_path = argv[2], camera_path = argv[6], out_path = argv[7];
    const unsigned width = (unsigned)std::atoi(argv[3]), height = (unsigned)std::atoi(argv[4]);
    const unsigned threads = (unsigned)std::atoi(argv[5]);
    std::string mesh_dir;
    if (const char* e = std::getenv("MTH_MESH_DIR")) {
        mesh_dir = e;
    } else {
        size_t slash = scene_path.find_last_of('/');
        mesh_dir = (slash == std::string::npos ? std::string(".") : scene_path.substr(0, slash)) + "/../obj";
    }
    try {
        mth::SceneDesc scene;
        scene.load(scene_path, mesh_dir);
        auto rentries = mth::parse_config_file(renderer_path);
        if (rentries.empty()) throw mth::InitError("InputParser", "empty renderer file \"" + renderer_path + "\"");
        auto renderer = mth::make_renderer(rentries[0].name);
        renderer->parseInput(rentries[0]);
        auto centries = mth::parse_config_file(camera_path);
        if (centries.empty()) throw mth::InitError("InputParser", "empty camera file \"" + camera_path + "\"");
        mth::GraphicsEngine engine(0);
        engine.width = width;
        engine.height = height;
        engine.camera.parse(centries[0]);
        engine.probes = scene.probes();
        for (size_t i = 1; i < rentries.size(); ++i) engine.denoisers.push_back(mth::make_denoiser(rentries[i]));
        engine.init(std::move(renderer), scene.flatten(), threads);
        engine.render();
        engine.save_image(out_path);
        std::printf("render time: %.3f ms\n", engine.last_render_ms);
    } catch (const std::exception& e) {
        std::cerr << e.what() << std::endl;
        return 1;
    }
    return 0;
}
