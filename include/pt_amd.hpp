// C++ facade over the C ABI (pt_amd.h) with the reference renderer's class
// names and methods, so code written against them ports unchanged:
//   pt::SceneManager  <- Part_2/src/scene/scene_manager.h:11-45
//   pt::Renderer      <- Part_2/src/renderer.h:17-55
//   pt::Integrator, pt::PathIntegrator <- integrators/integrator.h:10-20,
//                        integrators/path_integrator.h:7-15
//   pt::QualityMetrics, computeRMSE, rmseToDb, writeConvergenceCSV,
//   printConvergenceTable <- film/quality_metrics.h
// plus pt::WavefrontIntegrator (the queue-based form of the same estimator)
// and computeRelMSE.  Errors throw std::runtime_error with pt_last_error().
#pragma once

#include <chrono>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "pt_amd.h"

namespace pt {

inline void check(pt_status s) {
    if (s != PT_OK) throw std::runtime_error(pt_last_error());
}

class SceneManager {
public:
    SceneManager() = default;
    SceneManager(const SceneManager&) = delete;
    SceneManager& operator=(const SceneManager&) = delete;
    ~SceneManager() { reset(); }

    // SceneManager::load: .json scenes (the reference also reads glTF).
    bool load(const std::string& path) {
        reset();
        if (pt_scene_load(path.c_str(), &s_) != PT_OK) {
            std::cerr << "SceneManager: " << pt_last_error() << std::endl;
            return false;
        }
        return true;
    }
    bool generate(int which, uint64_t seed = 0) {
        reset();
        return pt_scene_generate(which, seed, &s_) == PT_OK;
    }
    void buildAccel(const pt_bvh_options* opts = nullptr) { check(pt_scene_build_accel(s_, opts)); }
    void uploadToDevice() { check(pt_scene_upload(s_)); }

    void setSpp(int spp) { check(pt_scene_set_spp(s_, spp)); }
    void setSamplePerFrame(int spf) { check(pt_scene_set_sample_per_frame(s_, spf)); }
    void setSampler(int sampler) { check(pt_scene_set_sampler(s_, sampler)); }  // 0 PCG, 1 Sobol
    void setMaxDepth(int depth) { check(pt_scene_set_max_depth(s_, depth)); }
    void setSeed(uint32_t seed) { check(pt_scene_set_seed(s_, seed)); }
    void setOutput(const std::string& path) { output_ = path; }
    void updateCamera(const pt_camera& cam) { check(pt_scene_set_camera(s_, &cam)); }

    pt_camera camera() const {
        pt_camera c;
        check(pt_scene_get_camera(s_, &c));
        return c;
    }
    pt_scene_info info() const {
        pt_scene_info i;
        check(pt_scene_get_info(s_, &i));
        return i;
    }
    // The scene file's "render.output" unless setOutput() chose another path.
    std::string outputPath() const {
        if (!output_.empty()) return output_;
        char buf[4096];
        check(pt_scene_output_path(s_, buf, sizeof(buf)));
        return buf;
    }
    pt_scene* handle() const { return s_; }

private:
    void reset() {
        if (s_) pt_scene_destroy(s_);
        s_ = nullptr;
    }
    pt_scene* s_ = nullptr;
    std::string output_;
};

class Integrator {
public:
    virtual ~Integrator() = default;
    virtual int kind() const = 0;
    virtual const char* name() const = 0;
};
class PathIntegrator : public Integrator {
public:
    int kind() const override { return PT_INTEGRATOR_PATH; }
    const char* name() const override { return "Path"; }
};
class WavefrontIntegrator : public Integrator {
public:
    int kind() const override { return PT_INTEGRATOR_WAVEFRONT; }
    const char* name() const override { return "Wavefront"; }
};

struct QualityMetrics {
    float rmse = 0.0f;
    float psnr = 0.0f;
    int spp = 0;
    float timeMs = 0.0f;
    float relMSE = 0.0f;
};

// film/quality_metrics.h: RMSE over RGB of two RGBA images.
inline float computeRMSE(const float* img, const float* ref, int width, int height, int channels = 3) {
    double sum = 0.0;
    long long count = 0;
    for (long long i = 0; i < static_cast<long long>(width) * height; ++i)
        for (int c = 0; c < channels; ++c) {
            double d = static_cast<double>(img[i * 4 + c]) - static_cast<double>(ref[i * 4 + c]);
            sum += d * d;
            ++count;
        }
    return count > 0 ? static_cast<float>(std::sqrt(sum / count)) : 0.0f;
}
// mean over pixels and channels of (I - R)^2 / (R^2 + 1e-2)
inline float computeRelMSE(const float* img, const float* ref, int width, int height) {
    double sum = 0.0;
    long long count = 0;
    for (long long i = 0; i < static_cast<long long>(width) * height; ++i)
        for (int c = 0; c < 3; ++c) {
            double r = ref[i * 4 + c], d = static_cast<double>(img[i * 4 + c]) - r;
            sum += d * d / (r * r + 1e-2);
            ++count;
        }
    return count > 0 ? static_cast<float>(sum / count) : 0.0f;
}
inline float rmseToDb(float rmse) { return rmse <= 0.0f ? 100.0f : -20.0f * std::log10(rmse); }

// The convergence table of film/quality_metrics.h:41-84, one row per
// checkpoint: sampler,spp,rmse,psnr_db,time_ms plus relmse as a sixth column.
inline std::string convergenceRow(const std::string& label, const QualityMetrics& m) {
    char row[160];
    std::snprintf(row, sizeof(row), "%s,%d,%.6f,%.2f,%.1f,%.8f", label.c_str(), m.spp, m.rmse, m.psnr, m.timeMs,
                  m.relMSE);
    return row;
}
inline void writeConvergenceCSV(const std::string& path, const std::vector<QualityMetrics>& data,
                                const std::string& label) {
    const bool fresh = !std::ifstream(path).good();
    std::ofstream f(path, std::ios::app);
    if (!f) {
        std::cerr << "Cannot open " << path << " for writing" << std::endl;
        return;
    }
    if (fresh) f << "sampler,spp,rmse,psnr_db,time_ms,relmse\n";
    for (const auto& m : data) f << convergenceRow(label, m) << "\n";
    std::cout << "Convergence data appended to " << path << " (" << label << ", " << data.size() << " checkpoints)"
              << std::endl;
}
inline void printConvergenceTable(const std::vector<QualityMetrics>& data, const std::string& label) {
    std::printf("\n=== Convergence: %s ===\n%8s %12s %12s %12s %14s\n", label.c_str(), "SPP", "RMSE", "PSNR(dB)",
                "Time(ms)", "relMSE");
    for (const auto& m : data)
        std::printf("%8d %12.6f %12.2f %12.1f %14.8f\n", m.spp, m.rmse, m.psnr, m.timeMs, m.relMSE);
}

class Renderer {
public:
    Renderer(SceneManager& scene, std::unique_ptr<Integrator> integrator)
        : scene_(scene), integrator_(std::move(integrator)) {
        pt_camera c = scene_.camera();
        w_ = c.width;
        h_ = c.height;
        check(pt_film_create(w_, h_, &film_));
    }
    Renderer(const Renderer&) = delete;
    Renderer& operator=(const Renderer&) = delete;
    ~Renderer() {
        if (film_) pt_film_destroy(film_);
    }

    // Renderer::renderBatch: one batch of samplePerFrame samples; returns the spp so far.
    uint32_t renderBatch() {
        float ms = 0.0f;
        check(pt_render_batch(scene_.handle(), film_, integrator_->kind(), &ms));
        lastMs_ = ms;
        totalMs_ += ms;
        return static_cast<uint32_t>(currentSpp());
    }
    void renderAll() {
        while (!isComplete()) renderBatch();
        std::printf("[%s] %d spp in %.1f ms (GPU)\n", integrator_->name(), currentSpp(), totalMs_);
    }
    int currentSpp() const {
        pt_scene_info i = scene_.info();
        return static_cast<int>(i.frame_count) * i.sample_per_frame;
    }
    int targetSpp() const { return scene_.info().target_samples; }
    bool isComplete() const { return currentSpp() >= targetSpp(); }
    int samplePerFrame() const { return scene_.info().sample_per_frame; }
    float lastBatchMs() const { return lastMs_; }
    float totalTraceMs() const { return totalMs_; }
    void resetAccumulation() {
        check(pt_film_reset(film_, scene_.handle()));
        totalMs_ = lastMs_ = 0.0f;
    }
    void resize(int w, int h) {
        if (w == w_ && h == h_) return;
        pt_camera c = scene_.camera();
        c.width = w;
        c.height = h;
        scene_.updateCamera(c);
        pt_film_destroy(film_);
        film_ = nullptr;
        w_ = w;
        h_ = h;
        check(pt_film_create(w_, h_, &film_));
        resetAccumulation();
    }
    void downloadCurrentImage(std::vector<float>& buf) const {
        buf.resize(static_cast<size_t>(w_) * h_ * 4);
        check(pt_film_download(film_, 0, buf.data()));
    }
    void saveImage(const std::string& path) const {
        std::vector<float> buf;
        downloadCurrentImage(buf);
        if (pt_image_write(path.c_str(), buf.data(), w_, h_) == PT_OK) std::cout << "Image saved: " << path << std::endl;
        else std::cerr << "Failed to save: " << path << " (" << pt_last_error() << ")" << std::endl;
    }
    // Renderer::runConvergenceBenchmark (renderer.cu:138-173), with relMSE.
    std::vector<QualityMetrics> runConvergenceBenchmark(const float* ref, const std::vector<int>& checkpoints) {
        resetAccumulation();
        std::vector<QualityMetrics> out;
        size_t next = 0;
        auto start = std::chrono::high_resolution_clock::now();
        while (next < checkpoints.size()) {
            uint32_t spp = renderBatch();
            if (static_cast<int>(spp) >= checkpoints[next]) {
                auto now = std::chrono::high_resolution_clock::now();
                std::vector<float> cur;
                downloadCurrentImage(cur);
                QualityMetrics m;
                m.spp = static_cast<int>(spp);
                m.rmse = computeRMSE(cur.data(), ref, w_, h_);
                m.psnr = rmseToDb(m.rmse);
                m.relMSE = computeRelMSE(cur.data(), ref, w_, h_);
                m.timeMs = std::chrono::duration<float, std::milli>(now - start).count();
                out.push_back(m);
                ++next;
            }
        }
        return out;
    }
    int filmWidth() const { return w_; }
    int filmHeight() const { return h_; }
    const Integrator& integrator() const { return *integrator_; }

private:
    SceneManager& scene_;
    std::unique_ptr<Integrator> integrator_;
    pt_film* film_ = nullptr;
    int w_ = 0, h_ = 0;
    float lastMs_ = 0.0f, totalMs_ = 0.0f;
};

}  // namespace pt
