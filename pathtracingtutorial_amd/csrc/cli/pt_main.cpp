// `pt` command-line renderer over the C++ facade.  Its interface is the
// reference's (Part_2/src/main.cpp:29-155):
//   pt [scene.json|.gltf|.glb] [--spp N] [--sampler pcg|sobol] [--benchmark [ref.hdr]] [--preview]
// with these additions:
//   --integrator path|wavefront   the plain megakernel or the wavefront form (default wavefront)
//   --tune                        run the wavefront autotuner before rendering
//   --generate N [seed]           a procedural workload instead of a scene file
//   --checkpoints a,b,c           convergence checkpoints of --benchmark (default 16..1024)
//   --reference-spp N             samples of the --benchmark reference image (default 4096)
//   --csv PATH                    where --benchmark appends its table (default convergence.csv)
// The GL preview of the reference is not built: --preview renders offline.
#include <unistd.h>

#include <cstdlib>
#include <filesystem>
#include <functional>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "pt_amd.hpp"

namespace fs = std::filesystem;

namespace {

struct Options {
    std::string scene;
    std::string integrator = "wavefront";
    int spp = -1;
    int sampler = -1;
    int generate = -1;
    uint64_t generate_seed = 0;
    bool tune = false;
    bool preview = false;
    bool benchmark = false;
    std::string benchmark_ref;
    std::vector<int> checkpoints = {16, 32, 64, 128, 256, 512, 1024};
    int reference_spp = 4096;
    std::string csv = "convergence.csv";
};

// One command-line option: its name, whether it takes a value ("required",
// or "optional" when the next word is not itself an option) and what it sets.
enum class Arity { none, required, optional };
struct OptionSpec {
    const char* name;
    Arity arity;
    std::function<bool(Options&, const std::string&)> apply;  // false: the value is invalid
};

bool parse_int_list(const std::string& text, std::vector<int>& out) {
    out.clear();
    std::stringstream ss(text);
    for (std::string item; std::getline(ss, item, ',');) {
        if (item.empty()) return false;
        out.push_back(std::atoi(item.c_str()));
        if (out.back() <= 0) return false;
    }
    return !out.empty();
}

const std::vector<OptionSpec>& option_table() {
    static const std::vector<OptionSpec> table = {
        {"--spp", Arity::required, [](Options& o, const std::string& v) { return (o.spp = std::atoi(v.c_str())) > 0; }},
        {"--sampler", Arity::required,
         [](Options& o, const std::string& v) {
             o.sampler = v == "pcg" ? 0 : (v == "sobol" ? 1 : -1);
             return o.sampler >= 0;
         }},
        {"--integrator", Arity::required,
         [](Options& o, const std::string& v) {
             o.integrator = v;
             return v == "path" || v == "wavefront";
         }},
        {"--benchmark", Arity::optional,
         [](Options& o, const std::string& v) {
             o.benchmark = true;
             o.benchmark_ref = v;
             return true;
         }},
        {"--preview", Arity::none, [](Options& o, const std::string&) { return o.preview = true; }},
        {"--tune", Arity::none, [](Options& o, const std::string&) { return o.tune = true; }},
        {"--checkpoints", Arity::required,
         [](Options& o, const std::string& v) { return parse_int_list(v, o.checkpoints); }},
        {"--reference-spp", Arity::required,
         [](Options& o, const std::string& v) { return (o.reference_spp = std::atoi(v.c_str())) > 0; }},
        {"--csv", Arity::required, [](Options& o, const std::string& v) { return !(o.csv = v).empty(); }},
    };
    return table;
}

bool is_option(const char* word) { return word[0] == '-' && word[1] == '-'; }

// Options may come in any order around the scene path.  Unknown options are
// ignored, as the reference does; a malformed value is an error.
bool parse(int argc, char** argv, Options& o) {
    int i = 1;
    auto next_value = [&](Arity a, std::string& v) {
        v.clear();
        if (a == Arity::none) return true;
        const bool present = i + 1 < argc && !(a == Arity::optional && is_option(argv[i + 1]));
        if (present) v = argv[++i];
        return present || a == Arity::optional;
    };
    for (; i < argc; ++i) {
        const std::string word = argv[i];
        if (word == "--generate") {  // an index and an optional seed
            std::string v;
            if (!next_value(Arity::required, v)) return false;
            o.generate = std::atoi(v.c_str());
            if (i + 1 < argc && !is_option(argv[i + 1])) o.generate_seed = std::strtoull(argv[++i], nullptr, 10);
            continue;
        }
        const OptionSpec* spec = nullptr;
        for (const OptionSpec& s : option_table())
            if (word == s.name) spec = &s;
        if (spec) {
            std::string v;
            if (!next_value(spec->arity, v) || !spec->apply(o, v)) {
                std::cerr << "pt: bad value for " << word << std::endl;
                return false;
            }
            continue;
        }
        if (is_option(word.c_str())) continue;
        if (o.scene.empty()) o.scene = word;
    }
    return true;
}

// The directory holding scenes/ and include/: $PT_ROOT if set, else the
// parent of the directory the executable was built into (the package sits
// at the repository root), else the working directory.
fs::path repository_root() {
    if (const char* env = std::getenv("PT_ROOT")) return fs::path(env);
    std::error_code ec;
    const fs::path exe = fs::read_symlink("/proc/self/exe", ec);
    if (!ec) {
        const fs::path candidate = exe.parent_path().parent_path();
        if (fs::is_directory(candidate / "scenes")) return candidate;
    }
    return fs::current_path();
}

bool load_reference_image(const std::string& path, int w, int h, std::vector<float>& out) {
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

// The convergence study of main.cpp:106-146: a reference image (read from
// ref, or rendered with Sobol at reference_spp samples and saved as
// benchmark_ref.hdr), then PCG and Sobol at each checkpoint: RMSE, PSNR,
// time and relMSE as a table and appended to the CSV (replaced first).
int run_benchmark(const Options& o, pt::SceneManager& sm, pt::Renderer& renderer) {
    std::vector<float> ref;
    if (!o.benchmark_ref.empty() &&
        load_reference_image(o.benchmark_ref, renderer.filmWidth(), renderer.filmHeight(), ref)) {
        std::cout << "=== Reference image: " << o.benchmark_ref << " ===" << std::endl;
    } else {
        std::cout << "=== Rendering reference image (" << o.reference_spp << " spp, sobol) ===" << std::endl;
        sm.setSampler(1);
        sm.setSpp(o.reference_spp);
        sm.uploadToDevice();
        renderer.resetAccumulation();
        renderer.renderAll();
        renderer.downloadCurrentImage(ref);
        renderer.saveImage("benchmark_ref.hdr");
    }
    std::error_code ec;
    fs::remove(o.csv, ec);
    const struct { int sampler; const char* label; const char* title; } runs[] = {{0, "pcg", "PCG"},
                                                                                   {1, "sobol", "Sobol"}};
    for (const auto& run : runs) {
        std::cout << "\n=== Benchmark: " << run.title << " ===" << std::endl;
        sm.setSampler(run.sampler);
        sm.setSpp(o.checkpoints.back());
        sm.uploadToDevice();
        const auto results = renderer.runConvergenceBenchmark(ref.data(), o.checkpoints);
        pt::printConvergenceTable(results, run.label);
        pt::writeConvergenceCSV(o.csv, results, run.label);
    }
    return 0;
}

}  // namespace

int main(int argc, char* argv[]) {
    Options o;
    if (!parse(argc, argv, o)) return 2;
    const fs::path root = repository_root();
    if (o.preview) std::cerr << "Preview not built (no GL in this renderer); rendering offline." << std::endl;
    if (pt_init(0) != PT_OK) {
        std::cerr << pt_last_error() << std::endl;
        return 1;
    }
    pt::SceneManager sm;
    if (o.generate >= 0) {
        if (!sm.generate(o.generate, o.generate_seed)) {
            std::cerr << pt_last_error() << std::endl;
            return 1;
        }
    } else {
        if (o.scene.empty()) {
            o.scene = (root / "scenes" / "cornell_box.json").string();
            std::cout << "Usage: pt [scene.json|.gltf|.glb] [--preview] [--spp N] [--sampler pcg|sobol] "
                         "[--benchmark [ref.hdr]] [--integrator path|wavefront] [--tune] [--generate N [seed]] "
                         "[--checkpoints a,b,..] [--reference-spp N] [--csv path]" << std::endl;
            std::cout << "No scene specified, using default: " << o.scene << std::endl;
        } else if (!fs::exists(o.scene) && fs::exists(root / o.scene)) {
            o.scene = (root / o.scene).string();
        }
        if (!sm.load(o.scene)) return 1;
    }
    try {
        if (o.spp > 0) sm.setSpp(o.spp);
        if (o.sampler >= 0) sm.setSampler(o.sampler);
        sm.buildAccel();
        sm.uploadToDevice();
        std::unique_ptr<pt::Integrator> integ;
        if (o.integrator == "path") integ = std::make_unique<pt::PathIntegrator>();
        else integ = std::make_unique<pt::WavefrontIntegrator>();
        if (o.tune && integ->kind() == PT_INTEGRATOR_WAVEFRONT) {
            pt_wavefront_config cfg;
            float ms = 0.0f;
            pt::check(pt_scene_tune_wavefront(sm.handle(), 16, &cfg, &ms));
            std::printf("wavefront config: chunk %d, blocks %d/%d, waves %d, sort %d (%.3f ms per spp)\n",
                        cfg.chunk_spp, cfg.extend_block, cfg.shade_block, cfg.waves_per_simd, cfg.sort_by_material,
                        ms);
        }
        pt::Renderer renderer(sm, std::move(integ));
        if (o.benchmark) return run_benchmark(o, sm, renderer);
        renderer.renderAll();
        renderer.saveImage(sm.outputPath());
    } catch (const std::exception& e) {
        std::cerr << "pt: " << e.what() << std::endl;
        return 1;
    }
    return 0;
}
