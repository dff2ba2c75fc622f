// `pt` command-line renderer: the reference's main.cpp (Part_2/src/main.cpp:
// 29-155) over the C++ facade.  Same arguments and flow:
//   pt [scene.json] [--spp N] [--sampler pcg|sobol] [--benchmark [ref.hdr]]
// plus --integrator path|wavefront (default wavefront), --tune (run the
// wavefront autotuner first) and --generate N [seed] (a procedural workload
// instead of a scene file).  The GL preview of the reference is not built.
#include <cstdlib>
#include <filesystem>
#include <iostream>
#include <string>
#include <vector>

#include "pt_amd.hpp"

namespace fs = std::filesystem;

static std::string find_project_root(const char* argv0) {
    for (auto start : {fs::path(argv0).parent_path(), fs::current_path()}) {
        for (auto dir = fs::absolute(start); dir.has_parent_path(); dir = dir.parent_path()) {
            if (fs::exists(dir / "scenes") && fs::exists(dir / "include" / "pt_amd.h")) return dir.string();
            if (dir == dir.parent_path()) break;
        }
    }
    return ".";
}

static bool load_hdr_rgba(const std::string& path, int w, int h, std::vector<float>& out) {
    int fw = 0, fh = 0;
    if (pt_image_read_hdr(path.c_str(), nullptr, &fw, &fh) != PT_OK) {
        std::cerr << pt_last_error() << std::endl;
        return false;
    }
    if (fw != w || fh != h) {
        std::cerr << path << " is " << fw << "x" << fh << ", the film is " << w << "x" << h << std::endl;
        return false;
    }
    out.resize(static_cast<size_t>(w) * h * 4);
    return pt_image_read_hdr(path.c_str(), out.data(), &fw, &fh) == PT_OK;
}

int main(int argc, char* argv[]) {
    std::string root = find_project_root(argv[0]);
    std::string scene_path;
    bool benchmark = false, preview = false, tune = false;
    std::string benchmark_ref;
    int spp_override = -1, sampler_override = -1, generate = -1;
    uint64_t gen_seed = 0;
    std::string integrator = "wavefront";
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        if (a == "--preview") { preview = true; continue; }
        if (a == "--tune") { tune = true; continue; }
        if (a == "--spp" && i + 1 < argc) { spp_override = std::atoi(argv[++i]); continue; }
        if (a == "--sampler" && i + 1 < argc) {
            std::string s = argv[++i];
            if (s == "pcg") sampler_override = 0;
            else if (s == "sobol") sampler_override = 1;
            continue;
        }
        if (a == "--integrator" && i + 1 < argc) { integrator = argv[++i]; continue; }
        if (a == "--generate" && i + 1 < argc) {
            generate = std::atoi(argv[++i]);
            if (i + 1 < argc && argv[i + 1][0] != '-') gen_seed = std::strtoull(argv[++i], nullptr, 10);
            continue;
        }
        if (a == "--benchmark") {
            benchmark = true;
            if (i + 1 < argc && argv[i + 1][0] != '-') benchmark_ref = argv[++i];
            continue;
        }
        if (a.rfind("--", 0) == 0) continue;
        if (scene_path.empty()) scene_path = a;
    }
    if (preview) std::cerr << "Preview not built (no GL in this renderer); rendering offline." << std::endl;
    if (pt_init(0) != PT_OK) {
        std::cerr << pt_last_error() << std::endl;
        return 1;
    }
    pt::SceneManager sm;
    if (generate >= 0) {
        if (!sm.generate(generate, gen_seed)) {
            std::cerr << pt_last_error() << std::endl;
            return 1;
        }
    } else {
        if (scene_path.empty()) {
            scene_path = (fs::path(root) / "scenes" / "cornell_box.json").string();
            std::cout << "Usage: pt [scene.json] [--spp N] [--sampler pcg|sobol] [--benchmark [ref.hdr]] "
                         "[--integrator path|wavefront] [--tune] [--generate N [seed]]" << std::endl;
            std::cout << "No scene specified, using default: " << scene_path << std::endl;
        } else if (!fs::exists(scene_path)) {
            auto p = fs::path(root) / scene_path;
            if (fs::exists(p)) scene_path = p.string();
        }
        if (!sm.load(scene_path)) return 1;
    }
    try {
        if (spp_override > 0) sm.setSpp(spp_override);
        if (sampler_override >= 0) sm.setSampler(sampler_override);
        sm.buildAccel();
        sm.uploadToDevice();
        std::unique_ptr<pt::Integrator> integ;
        if (integrator == "path") integ = std::make_unique<pt::PathIntegrator>();
        else integ = std::make_unique<pt::WavefrontIntegrator>();
        if (tune && integ->kind() == PT_INTEGRATOR_WAVEFRONT) {
            pt_wavefront_config cfg;
            float ms = 0.0f;
            pt::check(pt_scene_tune_wavefront(sm.handle(), 16, &cfg, &ms));
            std::printf("wavefront config: chunk %d, blocks %d/%d, waves %d, sort %d (%.3f ms per spp)\n",
                        cfg.chunk_spp, cfg.extend_block, cfg.shade_block, cfg.waves_per_simd, cfg.sort_by_material,
                        ms);
        }
        pt::Renderer renderer(sm, std::move(integ));
        if (benchmark) {
            // main.cpp:106-146: a 4096-spp Sobol reference, then PCG and Sobol
            // convergence at the checkpoints, as a table and convergence.csv
            std::vector<int> checkpoints = {16, 32, 64, 128, 256, 512, 1024};
            const std::string csv = "convergence.csv";
            std::vector<float> ref;
            if (!benchmark_ref.empty() && load_hdr_rgba(benchmark_ref, renderer.filmWidth(), renderer.filmHeight(), ref)) {
                std::cout << "=== Reference image: " << benchmark_ref << " ===" << std::endl;
            } else {
                std::cout << "=== Rendering reference image (4096 spp, sobol) ===" << std::endl;
                sm.setSampler(1);
                sm.setSpp(4096);
                sm.uploadToDevice();
                renderer.resetAccumulation();
                renderer.renderAll();
                renderer.downloadCurrentImage(ref);
                renderer.saveImage("benchmark_ref.hdr");
            }
            if (fs::exists(csv)) fs::remove(csv);
            for (int sampler : {0, 1}) {
                const char* label = sampler == 0 ? "pcg" : "sobol";
                std::cout << "\n=== Benchmark: " << (sampler == 0 ? "PCG" : "Sobol") << " ===" << std::endl;
                sm.setSampler(sampler);
                sm.setSpp(checkpoints.back());
                sm.uploadToDevice();
                auto results = renderer.runConvergenceBenchmark(ref.data(), checkpoints);
                pt::printConvergenceTable(results, label);
                pt::writeConvergenceCSV(csv, results, label);
            }
            return 0;
        }
        renderer.renderAll();
        renderer.saveImage(sm.outputPath());
    } catch (const std::exception& e) {
        std::cerr << "pt: " << e.what() << std::endl;
        return 1;
    }
    return 0;
}
