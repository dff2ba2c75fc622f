// C ABI of libptamd (include/pt_amd.h): scene lifecycle, device upload,
// rendering and probes.  Host stages live in csrc/host (plain C++), kernels in
// csrc/device.  Errors are returned as status codes with a thread-local
// message instead of the reference's print-and-exit (cuda/check.h:6-24).
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstring>
#include <memory>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "device/adjoint.h"
#include "device/bvh_build.h"
#include "device/bvh_wide_build.h"
#include "device/dev_scene.h"
#include "device/light_build.h"
#include "device/render.h"
#include "device/wavefront.h"
#include "host/bvh_wide.h"
#include "host/host_scene.h"
#include "pt_amd.h"

using ptamd::HostScene;

namespace {

thread_local std::string g_err;

struct Error : std::runtime_error {
    pt_status code;
    Error(pt_status c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw Error(PT_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

template <class F>
pt_status guarded(F&& f) {
    try {
        f();
        return PT_OK;
    } catch (const Error& e) {
        g_err = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        g_err = "out of host memory";
        return PT_ERR_OUT_OF_MEMORY;
    } catch (const std::exception& e) {
        g_err = e.what();
        return PT_ERR_INVALID_ARG;
    }
}

int g_device = -1;

void ensure_device() {
    if (g_device >= 0) return;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0)
        throw Error(PT_ERR_DEVICE, std::string("no HIP device available: ") + hipGetErrorString(e));
    hip_check(hipSetDevice(0), "hipSetDevice");
    g_device = 0;
}

// Owning device allocation.
template <class T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf() { reset(); }
    void reset() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void alloc(size_t count) {
        reset();
        if (count == 0) return;
        hip_check(hipMalloc(reinterpret_cast<void**>(&p), sizeof(T) * count), "hipMalloc");
        n = count;
    }
    void upload(const T* src, size_t count, hipStream_t s) {
        alloc(count);
        if (count) hip_check(hipMemcpyAsync(p, src, sizeof(T) * count, hipMemcpyHostToDevice, s), "hipMemcpy H2D");
    }
    void adopt(T* ptr, size_t count) {
        reset();
        p = ptr;
        n = count;
    }
};

int wrap_mode(int gl) {
    if (gl == 33071) return 1;  // GL_CLAMP_TO_EDGE
    if (gl == 33648) return 2;  // GL_MIRRORED_REPEAT
    return 0;                   // GL_REPEAT and anything else (cuda_texture.cu:30-37)
}

}  // namespace

struct pt_scene {
    HostScene host;
    ptamd::LightTables lights;
    hipStream_t stream = nullptr;
    bool accel_built = false, uploaded = false;
    // geometry (set by build_accel)
    DBuf<float4> vpos;
    DBuf<uint4> faces_orig, faces_bvh;
    DBuf<float4> tris, nodes;
    DBuf<float4> shrec;  // per-triangle shading records (DevScene::shade_rec), built at upload
    DBuf<float4> wnodes;  // 4-wide BVH (empty when the tree could overflow its traversal stack)
    uint32_t wide_max_stack = 0;
    DBuf<uint32_t> prim_order;
    std::vector<uint32_t> prim_order_host;
    uint32_t depth = 0, max_leaf = 0;
    float build_ms = 0.0f;
    // shading data (set by upload)
    DBuf<float4> vnrm, vtan, texels, env;
    DBuf<float2> vuv, tri_alias, env_alias;
    DBuf<pt_material> materials;
    DBuf<ptd::TexDesc> texdesc;
    DBuf<uint4> emissive;
    DBuf<float> tri_pmf, env_pmf;
    DBuf<pt_punctual_light> plights;
    int env_w = 0, env_h = 0, env_wrap_s = 0, env_wrap_t = 0;
    // render state
    DBuf<ptd::DevCounters> counters;
    DBuf<uint32_t> plain_tiles;               // the megakernel's tile-queue counters
    DBuf<unsigned long long> plain_spill;     // its traversal stack entries beyond LDS
    bool counting = false;
    uint32_t frame_count = 0;
    // wavefront integrator state (created at its first batch)
    ptd::Wavefront* wf = nullptr;
    ptd::WavefrontConfig wf_cfg;
    int adjoint_integrator = PT_INTEGRATOR_WAVEFRONT;
    bool env_tables_on_device = false;  // the device env tables were built on the GPU (host copy stale)
    bool emissive_tables_on_device = false;  // likewise for the emissive-triangle tables
    uint64_t generation = 0;  // bumped by every upload; keys the captured graph
    // bumped by every call that changes the scene's device data (geometry,
    // materials, textures, environment, light tables): keys the radiance a
    // wavefront forward keeps for the adjoint of the same batch
    uint64_t data_version = 1;
    bool keep_radiance = false;

    ~pt_scene() {
        if (wf) ptd::wavefront_destroy(wf);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

struct pt_film {
    int w = 0, h = 0;
    DBuf<float4> film_new, film_acc;
};

namespace {

void validate_scene(const HostScene& s) {
    const uint32_t nv = static_cast<uint32_t>(s.vertices.size());
    const int nm = static_cast<int>(s.materials.size());
    if (nm == 0) throw Error(PT_ERR_INVALID_ARG, "scene has no materials");
    for (size_t i = 0; i < s.faces.size(); ++i) {
        const pt_face& f = s.faces[i];
        if (f.v0 >= nv || f.v1 >= nv || f.v2 >= nv)
            throw Error(PT_ERR_INVALID_ARG, "face " + std::to_string(i) + " references a vertex out of range");
        if (f.material_id < 0 || f.material_id >= nm)
            throw Error(PT_ERR_INVALID_ARG, "face " + std::to_string(i) + " has material id " +
                                                std::to_string(f.material_id) + " outside [0, " +
                                                std::to_string(nm) + ")");
    }
    for (size_t i = 0; i < s.vertices.size(); ++i) {
        const pt_vec3p& p = s.vertices[i].position;
        if (!std::isfinite(p.x) || !std::isfinite(p.y) || !std::isfinite(p.z))
            throw Error(PT_ERR_INVALID_ARG, "vertex " + std::to_string(i) + " has a non-finite position");
    }
    for (size_t i = 0; i < s.textures.size(); ++i)
        if (s.textures[i].empty() ||
            s.textures[i].rgba.size() != static_cast<size_t>(s.textures[i].width) * s.textures[i].height * 4)
            throw Error(PT_ERR_INVALID_ARG, "texture " + std::to_string(i) + " is empty or malformed");
    if (s.camera.width <= 0 || s.camera.height <= 0 || s.camera.width > 65536 || s.camera.height > 65536)
        throw Error(PT_ERR_INVALID_ARG, "camera resolution out of range");
    if (s.settings.sample_per_frame <= 0) throw Error(PT_ERR_INVALID_ARG, "samplePerFrame must be positive");
    if (s.settings.max_depth < 0) throw Error(PT_ERR_INVALID_ARG, "maxDepth must be >= 0");
}

// Host-only until a call needs the device (parsing and queries run without a GPU).
pt_scene* new_scene() {
    std::unique_ptr<pt_scene> sc(new pt_scene());
    ptamd::init_scene_defaults(sc->host);
    return sc.release();
}

void scene_device(pt_scene* sc) {
    if (sc->stream) return;
    ensure_device();
    hip_check(hipStreamCreate(&sc->stream), "hipStreamCreate");
    sc->counters.alloc(1);
    hip_check(hipMemsetAsync(sc->counters.p, 0, sizeof(ptd::DevCounters), sc->stream), "hipMemset");
    hip_check(hipStreamSynchronize(sc->stream), "init");
}

ptd::DevScene make_view(const pt_scene* s) {
    ptd::DevScene v;
    std::memset(&v, 0, sizeof(v));
    v.nodes = s->nodes.p;
    v.tris = s->tris.p;
    v.faces = s->faces_bvh.p;
    v.vpos = s->vpos.p;
    v.vnrm = s->vnrm.p;
    v.vuv = s->vuv.p;
    v.vtan = s->vtan.p;
    v.materials = s->materials.p;
    v.textures = s->texdesc.p;
    v.texels = s->texels.p;
    v.emissive = s->emissive.p;
    v.tri_pmf = s->tri_pmf.p;
    v.tri_alias = s->tri_alias.p;
    v.env_pmf = s->env_pmf.p;
    v.env_alias = s->env_alias.p;
    v.env = s->env.p;
    v.lights = s->plights.p;
    v.shade_rec = s->shrec.p;
    v.num_nodes = s->nodes.p ? static_cast<uint32_t>(s->nodes.n / 2) : 0;
    v.wnodes = s->wnodes.p;
    v.num_wnodes = s->wnodes.p ? static_cast<uint32_t>(s->wnodes.n / 4) : 0;
    v.num_faces = static_cast<uint32_t>(s->host.faces.size());
    v.num_materials = static_cast<uint32_t>(s->host.materials.size());
    v.num_textures = static_cast<uint32_t>(s->texdesc.n);
    v.num_emissive = static_cast<uint32_t>(s->emissive.n);
    v.num_env_texels = static_cast<uint32_t>(s->env_pmf.n);
    v.num_lights = static_cast<uint32_t>(s->plights.n);
    v.env_w = s->env_w;
    v.env_h = s->env_h;
    v.env_wrap_s = s->env_wrap_s;
    v.env_wrap_t = s->env_wrap_t;
    return v;
}

ptd::RenderParams make_params_cam(const pt_scene* s, const pt_camera& cam) {
    ptd::RenderParams p;
    std::memset(&p, 0, sizeof(p));
    pt_camera_frame f = ptamd::camera_frame(cam);
    for (int i = 0; i < 3; ++i) {
        p.cam_pos[i] = f.position[i];
        p.cam_fwd[i] = f.forward[i];
        p.cam_right[i] = f.right[i];
        p.cam_up[i] = f.up[i];
    }
    p.tan_half_fov = f.tan_half_fov;
    p.aspect = f.aspect;
    p.aperture = f.aperture;
    p.focal_distance = f.focal_distance;
    p.width = cam.width;
    p.height = cam.height;
    p.spf = s->host.settings.sample_per_frame;
    p.max_depth = s->host.settings.max_depth;
    p.sampler = s->host.settings.sampler;
    p.seed = s->host.settings.seed;
    p.frame_count = s->frame_count;
    return p;
}

ptd::RenderParams make_params(const pt_scene* s) { return make_params_cam(s, s->host.camera); }

// Depth-first (reference) node order from the device-order nodes (2 float4
// per node, children adjacent): each inner node is followed by its left
// subtree, then its right one; inner words become depth-first slots.
void dfs_layout(const std::vector<float>& dev, std::vector<float>& out) {
    out.clear();
    if (dev.empty()) return;
    struct Item {
        uint32_t node, parent_slot;
        int which;
    };
    std::vector<Item> st;
    st.push_back({0, 0xFFFFFFFFu, 0});
    while (!st.empty()) {
        const Item it = st.back();
        st.pop_back();
        const uint32_t slot = static_cast<uint32_t>(out.size() / 8);
        if (it.parent_slot != 0xFFFFFFFFu) std::memcpy(&out[it.parent_slot * 8 + (it.which == 0 ? 3 : 7)], &slot, 4);
        out.resize(out.size() + 8);
        const float* src = &dev[8 * static_cast<size_t>(it.node)];
        std::memcpy(&out[8 * static_cast<size_t>(slot)], src, 8 * sizeof(float));
        uint32_t d0, d1;
        std::memcpy(&d0, src + 3, 4);
        std::memcpy(&d1, src + 7, 4);
        if (!(d0 & 0x80000000u)) {
            // push right first so the left subtree is laid out next (left = slot + 1)
            st.push_back({d1, slot, 1});
            st.push_back({d0, slot, 0});
        }
    }
}

}  // namespace

// ---------------------------------------------------------------------------

namespace {

void check_renderable(const pt_scene* s, const pt_film* f) {
    if (!s || !f) throw Error(PT_ERR_INVALID_ARG, "null argument");
    if (!s->uploaded) throw Error(PT_ERR_STATE, "scene not uploaded (pt_scene_upload)");
    if (f->w != s->host.camera.width || f->h != s->host.camera.height)
        throw Error(PT_ERR_INVALID_ARG, "film size does not match the camera resolution");
}

ptd::WavefrontJob make_job(const pt_scene* s, const pt_film* f, uint32_t frame_count) {
    ptd::WavefrontJob j;
    j.scene = make_view(s);
    j.params = make_params(s);
    j.params.frame_count = frame_count;
    j.film_new = f->film_new.p;
    j.film_acc = f->film_acc.p;
    return j;
}

// Renders the jobs with the wavefront engine of `owner` (its buffers, stream,
// counters and graph cache).
float render_wavefront_jobs(pt_scene* owner, const std::vector<ptd::WavefrontJob>& jobs,
                            const ptd::WavefrontConfig& cfg, uint64_t radiance_key = 0) {
    if (!owner->wf) owner->wf = ptd::wavefront_create();
    float ms = 0.0f;
    std::string err;
    hipError_t e =
        ptd::wavefront_render(owner->wf, jobs, cfg, owner->counters.p, owner->counting, owner->stream, &ms, &err,
                              radiance_key);
    if (e == hipErrorOutOfMemory) throw Error(PT_ERR_OUT_OF_MEMORY, err.empty() ? "wavefront: out of memory" : err);
    if (e != hipSuccess && !err.empty()) throw Error(e == hipErrorInvalidValue ? PT_ERR_INVALID_ARG : PT_ERR_DEVICE, err);
    hip_check(e, "wavefront render");
    return ms;
}

// The key under which a wavefront forward of s keeps its paths' radiance for
// the adjoint of the same batch (0: keep nothing).
uint64_t radiance_key(const pt_scene* s) { return s->keep_radiance ? s->data_version : 0; }

float render_wavefront(pt_scene* s, pt_film* f, const ptd::WavefrontConfig& cfg, uint32_t frame_count,
                       uint64_t key = 0) {
    return render_wavefront_jobs(s, std::vector<ptd::WavefrontJob>{make_job(s, f, frame_count)}, cfg, key);
}

void check_wavefront_config(const pt_wavefront_config& c) {
    auto block_ok = [](int b) { return b == 64 || b == 128 || b == 256; };
    if (c.chunk_spp < 1 || c.chunk_spp > 4096) throw Error(PT_ERR_INVALID_ARG, "chunk_spp must be in [1, 4096]");
    if (!block_ok(c.extend_block) || !block_ok(c.shade_block))
        throw Error(PT_ERR_INVALID_ARG, "block sizes must be 64, 128 or 256");
    if (c.waves_per_simd < 0 || c.waves_per_simd > 8) throw Error(PT_ERR_INVALID_ARG, "waves_per_simd must be in [0, 8]");
    if (c.shade_waves != 0 && c.shade_waves != 3 && c.shade_waves != 4)
        throw Error(PT_ERR_INVALID_ARG, "shade_waves must be 0, 3 or 4");
    if (c.stack_lds != 16 && c.stack_lds != 32) throw Error(PT_ERR_INVALID_ARG, "stack_lds must be 16 or 32");
    auto grab_ok = [](int g) { return g >= 64 && g <= 1024 && (g & (g - 1)) == 0; };
    if (!grab_ok(c.grab_extend) || !grab_ok(c.grab_shade) || !grab_ok(c.grab_shadow))
        throw Error(PT_ERR_INVALID_ARG, "grab sizes must be 64, 128, 256, 512 or 1024");
    if (c.bvh_width != 2 && c.bvh_width != 4) throw Error(PT_ERR_INVALID_ARG, "bvh_width must be 2 or 4");
}

ptd::WavefrontConfig to_device_config(const pt_wavefront_config& c) {
    ptd::WavefrontConfig d;
    d.chunk_spp = c.chunk_spp;
    d.extend_block = c.extend_block;
    d.shade_block = c.shade_block;
    d.waves_per_simd = c.waves_per_simd;
    d.sort_by_material = c.sort_by_material ? 1 : 0;
    d.use_graph = c.use_graph ? 1 : 0;
    d.time_stages = c.time_stages ? 1 : 0;
    d.shade_waves = c.shade_waves;
    d.stack_lds = c.stack_lds;
    d.grab_extend = c.grab_extend;
    d.grab_shade = c.grab_shade;
    d.grab_shadow = c.grab_shadow;
    d.bvh_width = c.bvh_width;
    return d;
}

pt_wavefront_config to_abi_config(const ptd::WavefrontConfig& d) {
    pt_wavefront_config c;
    c.chunk_spp = d.chunk_spp;
    c.extend_block = d.extend_block;
    c.shade_block = d.shade_block;
    c.waves_per_simd = d.waves_per_simd;
    c.sort_by_material = d.sort_by_material;
    c.use_graph = d.use_graph;
    c.time_stages = d.time_stages;
    c.shade_waves = d.shade_waves;
    c.stack_lds = d.stack_lds;
    c.grab_extend = d.grab_extend;
    c.grab_shade = d.grab_shade;
    c.grab_shadow = d.grab_shadow;
    c.bvh_width = d.bvh_width;
    return c;
}

}  // namespace

extern "C" {

int pt_abi_version(void) { return PT_ABI_VERSION; }

const char* pt_last_error(void) { return g_err.c_str(); }

pt_status pt_init(int device) {
    return guarded([&] {
        int n = 0;
        hipError_t e = hipGetDeviceCount(&n);
        if (e != hipSuccess || n <= 0)
            throw Error(PT_ERR_DEVICE, std::string("no HIP device available: ") + hipGetErrorString(e));
        if (device < 0 || device >= n) throw Error(PT_ERR_INVALID_ARG, "device index out of range");
        hip_check(hipSetDevice(device), "hipSetDevice");
        g_device = device;
    });
}

pt_status pt_struct_sizes(uint32_t* out, int n) {
    return guarded([&] {
        if (!out || n < 8) throw Error(PT_ERR_INVALID_ARG, "need room for 8 sizes");
        out[0] = sizeof(pt_vertex);
        out[1] = sizeof(pt_face);
        out[2] = sizeof(pt_material);
        out[3] = sizeof(pt_punctual_light);
        out[4] = sizeof(pt_camera);
        out[5] = sizeof(pt_camera_frame);
        out[6] = sizeof(pt_host_scene_desc);
        out[7] = sizeof(pt_counters);
    });
}

pt_status pt_scene_load(const char* path, pt_scene** out) {
    return guarded([&] {
        if (!path || !out) throw Error(PT_ERR_INVALID_ARG, "null argument");
        *out = nullptr;
        std::string p(path);
        std::string ext = p.substr(std::min(p.size(), p.find_last_of('.')));
        for (auto& c : ext) c = static_cast<char>(std::tolower(static_cast<unsigned char>(c)));
        if (ext != ".json" && ext != ".gltf" && ext != ".glb")
            throw Error(PT_ERR_UNSUPPORTED, "unknown scene format: " + p);
        std::unique_ptr<pt_scene> sc(new_scene());
        std::string err;
        if (!ptamd::load_scene(p, sc->host, &err)) {
            pt_status code = PT_ERR_PARSE;
            if (err.rfind("cannot open", 0) == 0) code = PT_ERR_IO;
            throw Error(code, err);
        }
        validate_scene(sc->host);
        *out = sc.release();
    });
}

pt_status pt_scene_create(const pt_host_scene_desc* d, pt_scene** out) {
    return guarded([&] {
        if (!d || !out) throw Error(PT_ERR_INVALID_ARG, "null argument");
        *out = nullptr;
        std::unique_ptr<pt_scene> sc(new_scene());
        HostScene& h = sc->host;
        if (d->num_vertices && !d->vertices) throw Error(PT_ERR_INVALID_ARG, "vertices is null");
        if (d->num_faces && !d->faces) throw Error(PT_ERR_INVALID_ARG, "faces is null");
        h.vertices.assign(d->vertices, d->vertices + d->num_vertices);
        h.faces.assign(d->faces, d->faces + d->num_faces);
        if (d->num_materials && d->materials) h.materials.assign(d->materials, d->materials + d->num_materials);
        if (h.materials.empty()) h.materials.push_back(ptamd::default_material());
        for (uint32_t i = 0; i < d->num_textures; ++i) {
            const pt_texture_desc& t = d->textures[i];
            if (!t.rgba || t.width <= 0 || t.height <= 0) throw Error(PT_ERR_INVALID_ARG, "bad texture descriptor");
            ptamd::Image img;
            img.width = t.width;
            img.height = t.height;
            img.wrap_s = t.wrap_s;
            img.wrap_t = t.wrap_t;
            img.linear = t.linear;
            img.rgba.assign(t.rgba, t.rgba + static_cast<size_t>(t.width) * t.height * 4);
            h.textures.push_back(std::move(img));
        }
        if (d->num_lights && d->lights) h.lights.assign(d->lights, d->lights + d->num_lights);
        if (d->camera) h.camera = *d->camera;
        else ptamd::auto_fit_camera(h.vertices, h.camera);
        if (d->env_rgba && d->env_width > 0 && d->env_height > 0) {
            h.env.width = d->env_width;
            h.env.height = d->env_height;
            h.env.rgba.assign(d->env_rgba, d->env_rgba + static_cast<size_t>(d->env_width) * d->env_height * 4);
        }
        if (d->settings) h.settings = *d->settings;
        validate_scene(h);
        *out = sc.release();
    });
}

pt_status pt_scene_generate(int which, uint64_t seed, pt_scene** out) {
    return guarded([&] {
        if (!out) throw Error(PT_ERR_INVALID_ARG, "null argument");
        *out = nullptr;
        std::unique_ptr<pt_scene> sc(new_scene());
        if (which == 0) ptamd::generate_cornell(sc->host);
        else if (which == 1) ptamd::generate_spheres(sc->host, seed, 64, 5, 2048, 1024);
        else if (which == 2) ptamd::generate_spheres(sc->host, seed, 8, 3, 256, 128);  // small variant for tests
        else if (which == 3) ptamd::generate_diff_cornell(sc->host, seed);
        else if (which == 4) ptamd::generate_batch_job(sc->host, seed);
        else if (which == 5) ptamd::generate_diff_views(sc->host, seed);
        else throw Error(PT_ERR_INVALID_ARG, "unknown workload id " + std::to_string(which));
        validate_scene(sc->host);
        *out = sc.release();
    });
}

void pt_scene_destroy(pt_scene* scene) { delete scene; }

pt_status pt_scene_get_info(const pt_scene* s, pt_scene_info* info) {
    return guarded([&] {
        if (!s || !info) throw Error(PT_ERR_INVALID_ARG, "null argument");
        std::memset(info, 0, sizeof(*info));
        info->num_vertices = static_cast<uint32_t>(s->host.vertices.size());
        info->num_faces = static_cast<uint32_t>(s->host.faces.size());
        info->num_materials = static_cast<uint32_t>(s->host.materials.size());
        info->num_textures = static_cast<uint32_t>(s->host.textures.size());
        info->num_lights = static_cast<uint32_t>(s->host.lights.size());
        info->num_emissive_triangles = static_cast<uint32_t>(s->emissive_tables_on_device ? s->emissive.n
                                                                                          : s->lights.emissive.size());
        info->num_bvh_nodes = s->nodes.p ? static_cast<uint32_t>(s->nodes.n / 2) : 0;
        info->bvh_depth = s->depth;
        info->max_leaf_size = s->max_leaf;
        info->env_width = s->host.env.width;
        info->env_height = s->host.env.height;
        info->width = s->host.camera.width;
        info->height = s->host.camera.height;
        info->sample_per_frame = s->host.settings.sample_per_frame;
        info->target_samples = s->host.settings.target_samples;
        info->max_depth = s->host.settings.max_depth;
        info->sampler = s->host.settings.sampler;
        info->seed = s->host.settings.seed;
        info->frame_count = s->frame_count;
        info->bvh_build_ms = s->build_ms;
    });
}

pt_status pt_scene_build_accel(pt_scene* s, const pt_bvh_options* opts) {
    return guarded([&] {
        if (!s) throw Error(PT_ERR_INVALID_ARG, "null scene");
        s->data_version++;
        scene_device(s);
        pt_bvh_options o = opts ? *opts : (s->host.has_bvh_opts ? s->host.bvh_opts : pt_bvh_options{24, 4, 0, 0});
        if (o.bins == 0) o.bins = 24;
        if (o.leaf_threshold == 0) o.leaf_threshold = 4;
        if (o.builder != 0) throw Error(PT_ERR_UNSUPPORTED, "only the GPU builder (0) is available");
        if (o.bins < 2 || o.bins > 32) throw Error(PT_ERR_INVALID_ARG, "bins must be in [2, 32]");
        if (o.leaf_threshold < 1 || o.max_leaf < 0) throw Error(PT_ERR_INVALID_ARG, "bad leaf options");
        HostScene& h = s->host;
        const uint32_t nv = static_cast<uint32_t>(h.vertices.size()), nf = static_cast<uint32_t>(h.faces.size());
        std::vector<float4> pos(nv);
        for (uint32_t i = 0; i < nv; ++i)
            pos[i] = make_float4(h.vertices[i].position.x, h.vertices[i].position.y, h.vertices[i].position.z, 0.0f);
        std::vector<uint4> fo(nf);
        for (uint32_t i = 0; i < nf; ++i)
            fo[i] = make_uint4(h.faces[i].v0, h.faces[i].v1, h.faces[i].v2, static_cast<uint32_t>(h.faces[i].material_id));
        s->accel_built = s->uploaded = false;
        s->vpos.upload(pos.data(), nv, s->stream);
        s->faces_orig.upload(fo.data(), nf, s->stream);
        s->nodes.reset();
        s->wnodes.reset();
        s->wide_max_stack = 0;
        s->tris.reset();
        s->faces_bvh.reset();
        s->prim_order.reset();
        s->depth = s->max_leaf = 0;
        s->build_ms = 0.0f;
        if (nf > 0) {
            ptd::BvhBuildOptions bo;
            bo.bins = o.bins;
            bo.leaf_threshold = o.leaf_threshold;
            bo.max_leaf = o.max_leaf;
            ptd::BvhDeviceResult r;
            std::string err;
            if (!ptd::gpu_build_bvh(s->vpos.p, s->faces_orig.p, nf, bo, s->stream, &r, nullptr, &err))
                throw Error(err.find("supports at most") != std::string::npos ? PT_ERR_LIMIT : PT_ERR_DEVICE, err);
            s->nodes.adopt(r.nodes, static_cast<size_t>(r.num_nodes) * 2);
            s->prim_order.adopt(r.prim_order, nf);
            s->depth = r.depth;
            s->max_leaf = r.max_leaf;
            s->build_ms = r.build_ms;
            s->tris.alloc(static_cast<size_t>(nf) * 3);
            s->faces_bvh.alloc(nf);
            hip_check(ptd::launch_emit_tris(s->vpos.p, s->faces_orig.p, s->prim_order.p, s->nodes.p, r.num_nodes, nf,
                                            s->tris.p, s->faces_bvh.p, s->stream),
                      "emit triangles");
            s->prim_order_host.resize(nf);
            hip_check(hipMemcpyAsync(s->prim_order_host.data(), s->prim_order.p, sizeof(uint32_t) * nf,
                                     hipMemcpyDeviceToHost, s->stream),
                      "download face order");
            // the 4-wide BVH, collapsed from the binary tree on the GPU (a tree
            // with non-finite bounds, or one deeper than the traversal stack,
            // keeps the binary traversal only)
            ptd::WideDeviceResult wr;
            std::string werr;
            if (!ptd::gpu_collapse_wide(s->nodes.p, r.num_nodes, r.level_begin, r.level_count, s->stream, &wr, &werr))
                throw Error(PT_ERR_DEVICE, werr);
            s->wide_max_stack = wr.max_stack;
            if (wr.nodes && wr.num_nodes > 0 && wr.max_stack <= static_cast<uint32_t>(ptd::kWideStack))
                s->wnodes.adopt(wr.nodes, static_cast<size_t>(wr.num_nodes) * 4);
            else if (wr.nodes)
                (void)hipFree(wr.nodes);
        } else {
            s->prim_order_host.clear();
        }
        hip_check(hipStreamSynchronize(s->stream), "build accel");
        // light tables over the BVH-ordered faces (scene_manager.cpp:31-34)
        std::vector<pt_face> faces_bvh(nf);
        for (uint32_t i = 0; i < nf; ++i) faces_bvh[i] = h.faces[s->prim_order_host[i]];
        ptamd::build_emissive_tables(h.vertices, faces_bvh, h.materials, s->lights);
        ptamd::build_env_tables(h.env, s->lights);
        s->accel_built = true;
    });
}

namespace {
void upload_emissive_tables(pt_scene* s) {
    const ptamd::LightTables& L = s->lights;
    std::vector<uint4> em(L.emissive.size());
    for (size_t i = 0; i < em.size(); ++i) {
        uint32_t pmf, cdf;
        std::memcpy(&pmf, &L.emissive[i].pmf, 4);
        std::memcpy(&cdf, &L.emissive[i].cdf, 4);
        em[i] = make_uint4(L.emissive[i].tri, pmf, cdf, 0u);
    }
    s->emissive.upload(em.data(), em.size(), s->stream);
    s->tri_pmf.upload(L.tri_pmf.data(), L.tri_pmf.size(), s->stream);
    s->tri_alias.upload(reinterpret_cast<const float2*>(L.tri_alias.data()), L.tri_alias.size(), s->stream);
    s->emissive_tables_on_device = false;
}
}  // namespace

namespace {
// The shading records of the current BVH order and vertex data (after an
// upload, and after a rebuild of the tree of an uploaded scene).
void build_shade_records(pt_scene* s) {
    const size_t nf = s->faces_bvh.n;
    if (nf == 0 || !s->vnrm.p || !s->vuv.p) {
        s->shrec.reset();
        return;
    }
    if (s->shrec.n != 4 * nf) s->shrec.alloc(4 * nf);
    hip_check(ptd::launch_shade_records(s->faces_bvh.p, s->vnrm.p, s->vuv.p, static_cast<uint32_t>(nf), s->shrec.p,
                                        s->stream),
              "shading records");
}
}  // namespace

pt_status pt_scene_upload(pt_scene* s) {
    return guarded([&] {
        if (!s) throw Error(PT_ERR_INVALID_ARG, "null scene");
        s->data_version++;
        if (!s->accel_built) throw Error(PT_ERR_STATE, "pt_scene_build_accel must run before pt_scene_upload");
        const HostScene& h = s->host;
        const size_t nv = h.vertices.size();
        std::vector<float4> nrm(nv), tan;
        std::vector<float2> uv(nv);
        for (size_t i = 0; i < nv; ++i) {
            const pt_vertex& v = h.vertices[i];
            nrm[i] = make_float4(v.normal.x, v.normal.y, v.normal.z, 0.0f);
            uv[i] = make_float2(v.u, v.v);
        }
        bool need_tan = false;
        for (const auto& m : h.materials) need_tan = need_tan || (m.normal_tex_id >= 0);
        s->vnrm.upload(nrm.data(), nv, s->stream);
        s->vuv.upload(uv.data(), nv, s->stream);
        if (need_tan) {
            tan.resize(nv);
            for (size_t i = 0; i < nv; ++i)
                tan[i] = make_float4(h.vertices[i].tangent[0], h.vertices[i].tangent[1], h.vertices[i].tangent[2],
                                     h.vertices[i].tangent[3]);
            s->vtan.upload(tan.data(), nv, s->stream);
        } else {
            s->vtan.reset();
        }
        s->materials.upload(h.materials.data(), h.materials.size(), s->stream);
        // textures: one texel pool + descriptors
        std::vector<ptd::TexDesc> td;
        size_t total = 0;
        for (const auto& t : h.textures) total += static_cast<size_t>(t.width) * t.height;
        if (total > 0xFFFFFFFFull) throw Error(PT_ERR_LIMIT, "texture pool larger than 2^32 texels");
        std::vector<float4> pool(total);
        size_t off = 0;
        for (const auto& t : h.textures) {
            ptd::TexDesc d;
            std::memset(&d, 0, sizeof(d));
            d.offset = static_cast<uint32_t>(off);
            d.w = t.width;
            d.h = t.height;
            d.wrap_s = wrap_mode(t.wrap_s);
            d.wrap_t = wrap_mode(t.wrap_t);
            d.linear = t.linear;
            td.push_back(d);
            std::memcpy(&pool[off], t.rgba.data(), sizeof(float4) * static_cast<size_t>(t.width) * t.height);
            off += static_cast<size_t>(t.width) * t.height;
        }
        s->texdesc.upload(td.data(), td.size(), s->stream);
        s->texels.upload(pool.data(), pool.size(), s->stream);
        // environment (a 4-channel image, device_scene.cu:19-27)
        if (!h.env.empty()) {
            s->env.upload(reinterpret_cast<const float4*>(h.env.rgba.data()),
                          static_cast<size_t>(h.env.width) * h.env.height, s->stream);
            s->env_w = h.env.width;
            s->env_h = h.env.height;
            s->env_wrap_s = wrap_mode(h.env.wrap_s);
            s->env_wrap_t = wrap_mode(h.env.wrap_t);
        } else {
            s->env.reset();
            s->env_w = s->env_h = 0;
        }
        const ptamd::LightTables& L = s->lights;
        upload_emissive_tables(s);
        s->env_pmf.upload(L.env_pmf.data(), L.env_pmf.size(), s->stream);
        s->env_alias.upload(reinterpret_cast<const float2*>(L.env_alias.data()), L.env_alias.size(), s->stream);
        if (L.env_pmf.empty()) s->env_pmf.reset(), s->env_alias.reset();
        s->env_tables_on_device = false;
        s->plights.upload(h.lights.data(), h.lights.size(), s->stream);
        build_shade_records(s);
        hip_check(hipStreamSynchronize(s->stream), "upload");
        s->frame_count = 0;
        s->uploaded = true;
        s->generation++;
    });
}

pt_status pt_scene_set_spp(pt_scene* s, int spp) {
    return guarded([&] {
        if (!s || spp <= 0) throw Error(PT_ERR_INVALID_ARG, "bad argument");
        s->host.settings.target_samples = spp;
    });
}
pt_status pt_scene_set_sample_per_frame(pt_scene* s, int spf) {
    return guarded([&] {
        if (!s || spf <= 0) throw Error(PT_ERR_INVALID_ARG, "bad argument");
        s->host.settings.sample_per_frame = spf;
    });
}
pt_status pt_scene_set_max_depth(pt_scene* s, int depth) {
    return guarded([&] {
        if (!s || depth < 0) throw Error(PT_ERR_INVALID_ARG, "bad argument");
        s->host.settings.max_depth = depth;
    });
}
pt_status pt_scene_set_sampler(pt_scene* s, int sampler) {
    return guarded([&] {
        if (!s || (sampler != 0 && sampler != 1)) throw Error(PT_ERR_INVALID_ARG, "sampler must be 0 (pcg) or 1 (sobol)");
        s->host.settings.sampler = sampler;
    });
}
pt_status pt_scene_set_seed(pt_scene* s, uint32_t seed) {
    return guarded([&] {
        if (!s) throw Error(PT_ERR_INVALID_ARG, "null scene");
        s->host.settings.seed = seed;
    });
}
pt_status pt_scene_set_camera(pt_scene* s, const pt_camera* cam) {
    return guarded([&] {
        if (!s || !cam) throw Error(PT_ERR_INVALID_ARG, "null argument");
        if (cam->width <= 0 || cam->height <= 0) throw Error(PT_ERR_INVALID_ARG, "camera resolution must be positive");
        s->host.camera = *cam;
    });
}
pt_status pt_scene_get_camera(const pt_scene* s, pt_camera* cam) {
    return guarded([&] {
        if (!s || !cam) throw Error(PT_ERR_INVALID_ARG, "null argument");
        *cam = s->host.camera;
    });
}
pt_status pt_scene_camera_frame(const pt_scene* s, pt_camera_frame* f) {
    return guarded([&] {
        if (!s || !f) throw Error(PT_ERR_INVALID_ARG, "null argument");
        *f = ptamd::camera_frame(s->host.camera);
    });
}

pt_status pt_scene_host_view(const pt_scene* s, pt_host_scene_view* o) {
    return guarded([&] {
        if (!s || !o) throw Error(PT_ERR_INVALID_ARG, "null argument");
        std::memset(o, 0, sizeof(*o));
        const HostScene& h = s->host;
        o->vertices = h.vertices.data();
        o->num_vertices = static_cast<uint32_t>(h.vertices.size());
        o->faces = h.faces.data();
        o->num_faces = static_cast<uint32_t>(h.faces.size());
        o->materials = h.materials.data();
        o->num_materials = static_cast<uint32_t>(h.materials.size());
        o->lights = h.lights.data();
        o->num_lights = static_cast<uint32_t>(h.lights.size());
        o->env_rgba = h.env.empty() ? nullptr : h.env.rgba.data();
        o->env_width = h.env.width;
        o->env_height = h.env.height;
        o->num_textures = static_cast<uint32_t>(h.textures.size());
        o->camera = h.camera;
        o->settings = h.settings;
    });
}

pt_status pt_scene_texture(const pt_scene* s, uint32_t i, pt_texture_desc* o) {
    return guarded([&] {
        if (!s || !o) throw Error(PT_ERR_INVALID_ARG, "null argument");
        if (i >= s->host.textures.size()) throw Error(PT_ERR_INVALID_ARG, "texture index out of range");
        const ptamd::Image& t = s->host.textures[i];
        std::memset(o, 0, sizeof(*o));
        o->rgba = t.rgba.data();
        o->width = t.width;
        o->height = t.height;
        o->wrap_s = t.wrap_s;
        o->wrap_t = t.wrap_t;
        o->linear = t.linear;
    });
}

pt_status pt_scene_set_texture_data(pt_scene* s, uint32_t i, const float* rgba, size_t count) {
    return guarded([&] {
        if (!s || !rgba) throw Error(PT_ERR_INVALID_ARG, "null argument");
        ptamd::Image* t = nullptr;
        if (i == 0xFFFFFFFFu) t = &s->host.env;
        else if (i < s->host.textures.size()) t = &s->host.textures[i];
        else throw Error(PT_ERR_INVALID_ARG, "texture index out of range");
        if (t->empty() || count != t->rgba.size()) throw Error(PT_ERR_INVALID_ARG, "texel count does not match");
        std::memcpy(t->rgba.data(), rgba, count * sizeof(float));
        s->uploaded = false;
    });
}

pt_status pt_scene_get_bvh(const pt_scene* s, int layout, float* nodes8, uint32_t* num_nodes, uint32_t* face_order,
                           uint32_t* num_faces) {
    return guarded([&] {
        if (!s) throw Error(PT_ERR_INVALID_ARG, "null scene");
        if (!s->accel_built) throw Error(PT_ERR_STATE, "acceleration structure not built");
        if (layout != 0 && layout != 1) throw Error(PT_ERR_INVALID_ARG, "layout must be 0 or 1");
        std::vector<float> out(s->nodes.n * 4);
        if (s->nodes.n)
            hip_check(hipMemcpy(out.data(), s->nodes.p, sizeof(float4) * s->nodes.n, hipMemcpyDeviceToHost),
                      "download nodes");
        if (layout == 0) {
            std::vector<float> dev;
            dev.swap(out);
            dfs_layout(dev, out);
        }
        uint32_t nn = static_cast<uint32_t>(out.size() / 8);
        if (nodes8) {
            if (!num_nodes || *num_nodes < nn) throw Error(PT_ERR_INVALID_ARG, "node buffer too small");
            std::memcpy(nodes8, out.data(), out.size() * sizeof(float));
        }
        if (num_nodes) *num_nodes = nn;
        uint32_t nf = static_cast<uint32_t>(s->prim_order_host.size());
        if (face_order) {
            if (!num_faces || *num_faces < nf) throw Error(PT_ERR_INVALID_ARG, "face buffer too small");
            std::memcpy(face_order, s->prim_order_host.data(), nf * sizeof(uint32_t));
        }
        if (num_faces) *num_faces = nf;
    });
}

pt_status pt_scene_get_wide_bvh(const pt_scene* s, float* wide16, uint32_t* num_wide, uint32_t* max_stack) {
    return guarded([&] {
        if (!s) throw Error(PT_ERR_INVALID_ARG, "null scene");
        if (!s->accel_built) throw Error(PT_ERR_STATE, "acceleration structure not built");
        const uint32_t nw = static_cast<uint32_t>(s->wnodes.n / 4);
        if (wide16 && nw) {
            if (!num_wide || *num_wide < nw) throw Error(PT_ERR_INVALID_ARG, "wide node buffer too small");
            hip_check(hipMemcpy(wide16, s->wnodes.p, sizeof(float4) * s->wnodes.n, hipMemcpyDeviceToHost),
                      "download wide nodes");
        }
        if (num_wide) *num_wide = nw;
        if (max_stack) *max_stack = s->wide_max_stack;
    });
}

pt_status pt_bvh_collapse_wide(const float* nodes8, uint32_t num_nodes, float* wide16, uint32_t* num_wide,
                               uint32_t* max_stack) {
    return guarded([&] {
        if (num_nodes > 0 && !nodes8) throw Error(PT_ERR_INVALID_ARG, "null nodes");
        ptamd::WideBvh wb;
        std::string err;
        if (!ptamd::collapse_wide_bvh(nodes8, num_nodes, &wb, &err)) throw Error(PT_ERR_INVALID_ARG, err);
        if (wide16) {
            if (!num_wide || *num_wide < wb.num_nodes) throw Error(PT_ERR_INVALID_ARG, "wide node buffer too small");
            std::memcpy(wide16, wb.words.data(), wb.words.size() * sizeof(float));
        }
        if (num_wide) *num_wide = wb.num_nodes;
        if (max_stack) *max_stack = wb.max_stack;
    });
}

pt_status pt_scene_get_light_tables(const pt_scene* s, float* tri_pmf, uint32_t* n_tri_pmf, float* tri_alias,
                                    uint32_t* n_tri_alias, float* env_pmf, float* env_alias, uint32_t* n_env) {
    return guarded([&] {
        if (!s) throw Error(PT_ERR_INVALID_ARG, "null scene");
        if (!s->accel_built) throw Error(PT_ERR_STATE, "light tables not built");
        const ptamd::LightTables& L = s->lights;
        const bool dev = s->emissive_tables_on_device;  // built on the GPU: read the device tables
        const size_t n_alias = dev ? s->tri_alias.n : L.tri_alias.size();
        if (tri_pmf) {
            if (!n_tri_pmf || *n_tri_pmf < L.tri_pmf.size()) throw Error(PT_ERR_INVALID_ARG, "tri_pmf too small");
            if (dev)
                hip_check(hipMemcpy(tri_pmf, s->tri_pmf.p, L.tri_pmf.size() * 4, hipMemcpyDeviceToHost), "light tables");
            else
                std::memcpy(tri_pmf, L.tri_pmf.data(), L.tri_pmf.size() * 4);
        }
        if (n_tri_pmf) *n_tri_pmf = static_cast<uint32_t>(L.tri_pmf.size());
        if (tri_alias) {
            if (!n_tri_alias || *n_tri_alias < n_alias) throw Error(PT_ERR_INVALID_ARG, "tri_alias too small");
            if (dev && n_alias)
                hip_check(hipMemcpy(tri_alias, s->tri_alias.p, n_alias * 8, hipMemcpyDeviceToHost), "light tables");
            else if (!dev)
                std::memcpy(tri_alias, L.tri_alias.data(), n_alias * 8);
        }
        if (n_tri_alias) *n_tri_alias = static_cast<uint32_t>(n_alias);
        if (env_pmf || env_alias) {
            if (!n_env || *n_env < L.env_pmf.size()) throw Error(PT_ERR_INVALID_ARG, "env buffers too small");
            if (s->env_tables_on_device) {  // built on the GPU: read the device tables
                if (env_pmf)
                    hip_check(hipMemcpy(env_pmf, s->env_pmf.p, L.env_pmf.size() * 4, hipMemcpyDeviceToHost), "env tables");
                if (env_alias)
                    hip_check(hipMemcpy(env_alias, s->env_alias.p, L.env_alias.size() * 8, hipMemcpyDeviceToHost),
                              "env tables");
            } else {
                if (env_pmf) std::memcpy(env_pmf, L.env_pmf.data(), L.env_pmf.size() * 4);
                if (env_alias) std::memcpy(env_alias, L.env_alias.data(), L.env_alias.size() * 8);
            }
        }
        if (n_env) *n_env = static_cast<uint32_t>(L.env_pmf.size());
    });
}

pt_status pt_film_create(int width, int height, pt_film** out) {
    return guarded([&] {
        if (!out || width <= 0 || height <= 0 || width > 65536 || height > 65536)
            throw Error(PT_ERR_INVALID_ARG, "bad film size");
        ensure_device();
        std::unique_ptr<pt_film> f(new pt_film());
        f->w = width;
        f->h = height;
        size_t n = static_cast<size_t>(width) * height;
        f->film_new.alloc(n);
        f->film_acc.alloc(n);
        hip_check(hipMemset(f->film_new.p, 0, n * sizeof(float4)), "hipMemset");
        hip_check(hipMemset(f->film_acc.p, 0, n * sizeof(float4)), "hipMemset");
        *out = f.release();
    });
}

void pt_film_destroy(pt_film* film) { delete film; }

pt_status pt_film_reset(pt_film* f, pt_scene* s) {
    return guarded([&] {
        if (!f) throw Error(PT_ERR_INVALID_ARG, "null film");
        size_t n = static_cast<size_t>(f->w) * f->h;
        hip_check(hipMemset(f->film_new.p, 0, n * sizeof(float4)), "hipMemset");
        hip_check(hipMemset(f->film_acc.p, 0, n * sizeof(float4)), "hipMemset");
        hip_check(hipDeviceSynchronize(), "film reset");
        if (s) s->frame_count = 0;
    });
}

pt_status pt_film_download(const pt_film* f, int which, float* rgba) {
    return guarded([&] {
        if (!f || !rgba) throw Error(PT_ERR_INVALID_ARG, "null argument");
        const float4* src = which == 1 ? f->film_new.p : f->film_acc.p;
        hip_check(hipMemcpy(rgba, src, sizeof(float4) * static_cast<size_t>(f->w) * f->h, hipMemcpyDeviceToHost),
                  "film download");
    });
}

pt_status pt_film_device_ptr(const pt_film* f, int which, void** p) {
    return guarded([&] {
        if (!f || !p) throw Error(PT_ERR_INVALID_ARG, "null argument");
        *p = which == 1 ? static_cast<void*>(f->film_new.p) : static_cast<void*>(f->film_acc.p);
    });
}


pt_status pt_scene_output_path(const pt_scene* s, char* buf, size_t cap) {
    return guarded([&] {
        if (!s || !buf || cap == 0) throw Error(PT_ERR_INVALID_ARG, "null argument");
        const std::string& o = s->host.output_file.empty() ? std::string("output.hdr") : s->host.output_file;
        if (o.size() + 1 > cap) throw Error(PT_ERR_INVALID_ARG, "buffer too small for the output path");
        std::memcpy(buf, o.c_str(), o.size() + 1);
    });
}

pt_status pt_image_read(const char* path, float* rgba, int* width, int* height) {
    return guarded([&] {
        if (!path || !width || !height) throw Error(PT_ERR_INVALID_ARG, "null argument");
        ptamd::Image img;
        std::string err;
        if (!ptamd::load_image(path, img, &err))
            throw Error(PT_ERR_IO, err.empty() ? std::string("cannot read ") + path : err);
        if (rgba) {
            if (*width != img.width || *height != img.height) throw Error(PT_ERR_INVALID_ARG, "buffer size mismatch");
            // the loader returns the file's top row first; films keep row 0 at the bottom
            const size_t row = static_cast<size_t>(img.width) * 4;
            for (int y = 0; y < img.height; ++y)
                std::memcpy(rgba + static_cast<size_t>(img.height - 1 - y) * row, img.rgba.data() + y * row,
                            row * sizeof(float));
        }
        *width = img.width;
        *height = img.height;
    });
}

pt_status pt_image_read_hdr(const char* path, float* rgba, int* width, int* height) {
    return pt_image_read(path, rgba, width, height);
}

pt_status pt_image_write(const char* path, const float* rgba, int w, int h) {
    return guarded([&] {
        if (!path || !rgba || w <= 0 || h <= 0) throw Error(PT_ERR_INVALID_ARG, "bad image arguments");
        std::string p(path), err;
        std::string ext = p.substr(std::min(p.size(), p.find_last_of('.')));
        for (auto& ch : ext) ch = static_cast<char>(std::tolower(static_cast<unsigned char>(ch)));
        static const char* known[] = {".hdr", ".pfm", ".png", ".jpg", ".jpeg", ".bmp", ".tga"};
        if (std::none_of(std::begin(known), std::end(known), [&](const char* k) { return ext == k; }))
            throw Error(PT_ERR_UNSUPPORTED, "image format not supported (.hdr .pfm .png .jpg .jpeg .bmp .tga): " + p);
        if (!ptamd::write_image(p, w, h, rgba, true, &err))
            throw Error(PT_ERR_IO, err.empty() ? "cannot write " + p : err);
    });
}

pt_status pt_film_copy_device(const pt_film* f, int which, float* dst) {
    return guarded([&] {
        if (!f || !dst) throw Error(PT_ERR_INVALID_ARG, "null argument");
        const float4* src = which == 1 ? f->film_new.p : f->film_acc.p;
        hip_check(hipMemcpy(dst, src, sizeof(float4) * static_cast<size_t>(f->w) * f->h, hipMemcpyDeviceToDevice),
                  "film copy");
    });
}

pt_status pt_render_batch(pt_scene* s, pt_film* f, int integrator, float* ms_out) {
    return guarded([&] {
        check_renderable(s, f);
        float ms = 0.0f;
        if (integrator == PT_INTEGRATOR_WAVEFRONT) {
            ms = render_wavefront(s, f, s->wf_cfg, s->frame_count, radiance_key(s));
        } else if (integrator == PT_INTEGRATOR_PATH) {
            ptd::DevScene v = make_view(s);
            ptd::RenderParams p = make_params(s);
            const bool wide = s->wf_cfg.bvh_width == 4 && v.num_wnodes > 0;
            size_t spill = 0;
            int ntiles = 0;
            hip_check(ptd::render_plain_scratch(wide, &spill, &ntiles), "megakernel occupancy");
            if (s->plain_spill.n < spill) s->plain_spill.alloc(spill);
            if (s->plain_tiles.n < static_cast<size_t>(ntiles)) s->plain_tiles.alloc(static_cast<size_t>(ntiles));
            const ptd::PlainScratch scratch{s->plain_tiles.p, s->plain_spill.p, s->plain_spill.n};
            hipEvent_t e0, e1;
            hip_check(hipEventCreate(&e0), "hipEventCreate");
            hip_check(hipEventCreate(&e1), "hipEventCreate");
            hip_check(hipEventRecord(e0, s->stream), "hipEventRecord");
            hipError_t le = ptd::launch_render_plain(v, p, f->film_new.p, f->film_acc.p, s->counters.p, s->counting,
                                                     wide, scratch, s->stream);
            hip_check(hipEventRecord(e1, s->stream), "hipEventRecord");
            hipError_t se = hipEventSynchronize(e1);
            (void)hipEventElapsedTime(&ms, e0, e1);
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            hip_check(le, "render launch");
            hip_check(se, "render");
        } else {
            throw Error(PT_ERR_INVALID_ARG, "unknown integrator " + std::to_string(integrator));
        }
        if (ms_out) *ms_out = ms;
        s->frame_count++;
    });
}

pt_status pt_render_jobs(const pt_job* jobs, int n, float* ms_out) {
    return guarded([&] {
        if (!jobs || n <= 0) throw Error(PT_ERR_INVALID_ARG, "no jobs");
        std::vector<ptd::WavefrontJob> js;
        js.reserve(static_cast<size_t>(n));
        for (int i = 0; i < n; ++i) {
            check_renderable(jobs[i].scene, jobs[i].film);
            for (int k = 0; k < i; ++k)
                if (jobs[k].film == jobs[i].film) throw Error(PT_ERR_INVALID_ARG, "two jobs share a film");
            js.push_back(make_job(jobs[i].scene, jobs[i].film, jobs[i].scene->frame_count));
        }
        pt_scene* owner = jobs[0].scene;
        float ms = render_wavefront_jobs(owner, js, owner->wf_cfg);
        // a scene that appears in several jobs advances once per job
        for (int i = 0; i < n; ++i) jobs[i].scene->frame_count++;
        if (ms_out) *ms_out = ms;
    });
}

pt_status pt_scene_set_wavefront_config(pt_scene* s, const pt_wavefront_config* cfg) {
    return guarded([&] {
        if (!s || !cfg) throw Error(PT_ERR_INVALID_ARG, "null argument");
        check_wavefront_config(*cfg);
        s->wf_cfg = to_device_config(*cfg);
    });
}

pt_status pt_scene_get_wavefront_config(const pt_scene* s, pt_wavefront_config* cfg) {
    return guarded([&] {
        if (!s || !cfg) throw Error(PT_ERR_INVALID_ARG, "null argument");
        *cfg = to_abi_config(s->wf_cfg);
    });
}

namespace {

ptd::AdjointGrads to_grads(const pt_adjoint_grads* g) {
    ptd::AdjointGrads G;
    G.base = g->base_color;
    G.rough = g->roughness;
    G.metal = g->metallic;
    G.tex = g->textures;
    G.env = g->env;
    return G;
}

float run_adjoint(pt_scene* s, const std::vector<ptd::AdjointView>& views, const pt_adjoint_grads* g) {
    ptd::DevScene v = make_view(s);
    if (s->adjoint_integrator == PT_INTEGRATOR_WAVEFRONT) {
        std::vector<ptd::WavefrontJob> jobs(views.size());
        for (size_t i = 0; i < views.size(); ++i) {
            jobs[i].scene = v;
            jobs[i].params = views[i].params;
            jobs[i].film_new = jobs[i].film_acc = nullptr;
            jobs[i].img_grad = views[i].img_grad;
            jobs[i].replay = views[i].replay;
        }
        if (!s->wf) s->wf = ptd::wavefront_create();
        float ms = 0.0f;
        std::string err;
        hipError_t e = ptd::wavefront_adjoint(s->wf, jobs, to_grads(g), s->wf_cfg, s->counters.p, s->counting,
                                              s->stream, &ms, &err, radiance_key(s));
        if (e == hipErrorOutOfMemory) throw Error(PT_ERR_OUT_OF_MEMORY, err.empty() ? "adjoint: out of memory" : err);
        if (e != hipSuccess && !err.empty())
            throw Error(e == hipErrorInvalidValue ? PT_ERR_INVALID_ARG : PT_ERR_DEVICE, err);
        hip_check(e, "wavefront adjoint");
        return ms;
    }
    hipEvent_t e0, e1;
    hip_check(hipEventCreate(&e0), "hipEventCreate");
    hip_check(hipEventCreate(&e1), "hipEventCreate");
    hip_check(hipEventRecord(e0, s->stream), "hipEventRecord");
    hipError_t le = ptd::launch_adjoint(v, views, to_grads(g), s->counters.p, s->counting, s->stream);
    hip_check(hipEventRecord(e1, s->stream), "hipEventRecord");
    hipError_t se = hipEventSynchronize(e1);
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    hip_check(le, "adjoint launch");
    hip_check(se, "adjoint");
    return ms;
}

void check_camera(const pt_camera& c) {
    if (c.width <= 0 || c.height <= 0 || c.width > 65536 || c.height > 65536)
        throw Error(PT_ERR_INVALID_ARG, "camera resolution out of range");
}

}  // namespace

pt_status pt_render_adjoint(pt_scene* s, uint32_t frame, const float* image_grad, const pt_adjoint_grads* g,
                            float* replay_image, float* ms_out) {
    return guarded([&] {
        if (!s || !image_grad || !g) throw Error(PT_ERR_INVALID_ARG, "null argument");
        if (!s->uploaded) throw Error(PT_ERR_STATE, "scene not uploaded (pt_scene_upload)");
        ptd::AdjointView view;
        view.params = make_params(s);
        view.params.frame_count = frame;
        view.img_grad = reinterpret_cast<const float4*>(image_grad);
        view.replay = reinterpret_cast<float4*>(replay_image);
        float ms = run_adjoint(s, std::vector<ptd::AdjointView>{view}, g);
        if (ms_out) *ms_out = ms;
    });
}

pt_status pt_render_adjoint_views(pt_scene* s, int n, const pt_camera* cameras, const uint32_t* frames,
                                  const float* const* image_grads, const pt_adjoint_grads* g,
                                  float* const* replay_images, float* ms_out) {
    return guarded([&] {
        if (!s || n <= 0 || !cameras || !frames || !image_grads || !g) throw Error(PT_ERR_INVALID_ARG, "null argument");
        if (!s->uploaded) throw Error(PT_ERR_STATE, "scene not uploaded (pt_scene_upload)");
        std::vector<ptd::AdjointView> views(static_cast<size_t>(n));
        for (int i = 0; i < n; ++i) {
            check_camera(cameras[i]);
            if (!image_grads[i]) throw Error(PT_ERR_INVALID_ARG, "null image gradient");
            views[i].params = make_params_cam(s, cameras[i]);
            views[i].params.frame_count = frames[i];
            views[i].img_grad = reinterpret_cast<const float4*>(image_grads[i]);
            views[i].replay = replay_images ? reinterpret_cast<float4*>(replay_images[i]) : nullptr;
        }
        float ms = run_adjoint(s, views, g);
        if (ms_out) *ms_out = ms;
    });
}

pt_status pt_render_views(pt_scene* s, int n, const pt_camera* cameras, pt_film* const* films,
                          const uint32_t* frames, float* ms_out) {
    return guarded([&] {
        if (!s || n <= 0 || !cameras || !films) throw Error(PT_ERR_INVALID_ARG, "null argument");
        if (!s->uploaded) throw Error(PT_ERR_STATE, "scene not uploaded (pt_scene_upload)");
        std::vector<ptd::WavefrontJob> jobs(static_cast<size_t>(n));
        for (int i = 0; i < n; ++i) {
            check_camera(cameras[i]);
            const pt_film* f = films[i];
            if (!f || f->w != cameras[i].width || f->h != cameras[i].height)
                throw Error(PT_ERR_INVALID_ARG, "film size does not match the view's camera");
            for (int k = 0; k < i; ++k)
                if (films[k] == f) throw Error(PT_ERR_INVALID_ARG, "two views share a film");
            jobs[i].scene = make_view(s);
            jobs[i].params = make_params_cam(s, cameras[i]);
            jobs[i].params.frame_count = frames ? frames[i] : s->frame_count;
            jobs[i].film_new = f->film_new.p;
            jobs[i].film_acc = f->film_acc.p;
        }
        float ms = render_wavefront_jobs(s, jobs, s->wf_cfg, radiance_key(s));
        if (ms_out) *ms_out = ms;
    });
}

pt_status pt_scene_set_adjoint_integrator(pt_scene* s, int integrator) {
    return guarded([&] {
        if (!s) throw Error(PT_ERR_INVALID_ARG, "null scene");
        if (integrator != PT_INTEGRATOR_PATH && integrator != PT_INTEGRATOR_WAVEFRONT)
            throw Error(PT_ERR_INVALID_ARG, "unknown integrator");
        s->adjoint_integrator = integrator;
    });
}

pt_status pt_scene_get_adjoint_integrator(const pt_scene* s, int* integrator) {
    return guarded([&] {
        if (!s || !integrator) throw Error(PT_ERR_INVALID_ARG, "null argument");
        *integrator = s->adjoint_integrator;
    });
}

pt_status pt_scene_set_keep_radiance(pt_scene* s, int on) {
    return guarded([&] {
        if (!s) throw Error(PT_ERR_INVALID_ARG, "null scene");
        s->keep_radiance = on != 0;
        if (!s->keep_radiance && s->wf) ptd::wavefront_release_kept(s->wf);
    });
}

pt_status pt_scene_get_keep_radiance(const pt_scene* s, int* on, int* reused) {
    return guarded([&] {
        if (!s) throw Error(PT_ERR_INVALID_ARG, "null scene");
        if (on) *on = s->keep_radiance ? 1 : 0;
        if (reused) *reused = ptd::wavefront_adjoint_reused_radiance(s->wf) ? 1 : 0;
    });
}

pt_status pt_scene_set_frame(pt_scene* s, uint32_t frame_count) {
    return guarded([&] {
        if (!s) throw Error(PT_ERR_INVALID_ARG, "null scene");
        s->frame_count = frame_count;
    });
}

pt_status pt_scene_texture_offset(const pt_scene* s, uint32_t tex, uint64_t* texel_offset, uint64_t* texel_count) {
    return guarded([&] {
        if (!s) throw Error(PT_ERR_INVALID_ARG, "null scene");
        if (tex >= s->host.textures.size()) throw Error(PT_ERR_INVALID_ARG, "texture index out of range");
        uint64_t off = 0;
        for (uint32_t i = 0; i < tex; ++i)
            off += static_cast<uint64_t>(s->host.textures[i].width) * s->host.textures[i].height;
        if (texel_offset) *texel_offset = off;
        if (texel_count)
            *texel_count = static_cast<uint64_t>(s->host.textures[tex].width) * s->host.textures[tex].height;
    });
}

pt_status pt_scene_update_materials(pt_scene* s, const float* base_color, const float* roughness,
                                    const float* metallic) {
    return guarded([&] {
        if (!s) throw Error(PT_ERR_INVALID_ARG, "null scene");
        s->data_version++;
        auto& M = s->host.materials;
        for (size_t i = 0; i < M.size(); ++i) {
            if (base_color) {
                M[i].base_color.x = base_color[3 * i + 0];
                M[i].base_color.y = base_color[3 * i + 1];
                M[i].base_color.z = base_color[3 * i + 2];
            }
            if (roughness) M[i].roughness = roughness[i];
            if (metallic) M[i].metallic = metallic[i];
        }
        if (s->uploaded && s->materials.p && s->materials.n == M.size())
            hip_check(hipMemcpy(s->materials.p, M.data(), sizeof(pt_material) * M.size(), hipMemcpyHostToDevice),
                      "material update");
    });
}

pt_status pt_scene_update_texture(pt_scene* s, uint32_t tex, const float* rgba, size_t count, int on_device) {
    return guarded([&] {
        if (!s || !rgba) throw Error(PT_ERR_INVALID_ARG, "null argument");
        s->data_version++;
        if (!s->uploaded) throw Error(PT_ERR_STATE, "scene not uploaded (pt_scene_upload)");
        ptamd::Image* img = nullptr;
        float4* dst = nullptr;
        if (tex == 0xFFFFFFFFu) {
            img = &s->host.env;
            dst = s->env.p;
        } else {
            if (tex >= s->host.textures.size()) throw Error(PT_ERR_INVALID_ARG, "texture index out of range");
            img = &s->host.textures[tex];
            uint64_t off = 0;
            for (uint32_t i = 0; i < tex; ++i)
                off += static_cast<uint64_t>(s->host.textures[i].width) * s->host.textures[i].height;
            dst = s->texels.p + off;
        }
        if (img->empty() || dst == nullptr || count != img->rgba.size())
            throw Error(PT_ERR_INVALID_ARG, "texel count does not match the texture");
        const size_t bytes = count * sizeof(float);
        if (on_device) {
            hip_check(hipMemcpy(dst, rgba, bytes, hipMemcpyDeviceToDevice), "texture update");
            hip_check(hipMemcpy(img->rgba.data(), rgba, bytes, hipMemcpyDeviceToHost), "texture update");
        } else {
            std::memcpy(img->rgba.data(), rgba, bytes);
            hip_check(hipMemcpy(dst, rgba, bytes, hipMemcpyHostToDevice), "texture update");
        }
    });
}

pt_status pt_scene_rebuild_env_tables(pt_scene* s) {
    return guarded([&] {
        if (!s) throw Error(PT_ERR_INVALID_ARG, "null scene");
        s->data_version++;
        if (!s->uploaded) throw Error(PT_ERR_STATE, "scene not uploaded (pt_scene_upload)");
        ptamd::LightTables t;
        ptamd::build_env_tables(s->host.env, t);
        if (t.env_pmf.size() != s->env_pmf.n || t.env_alias.size() != s->env_alias.n)
            throw Error(PT_ERR_STATE, "environment sampling tables changed size (an all-zero environment?)");
        s->lights.env_pmf = t.env_pmf;
        s->lights.env_alias = t.env_alias;
        s->env_tables_on_device = false;
        if (!t.env_pmf.empty()) {
            hip_check(hipMemcpy(s->env_pmf.p, t.env_pmf.data(), t.env_pmf.size() * 4, hipMemcpyHostToDevice),
                      "env table update");
            hip_check(hipMemcpy(s->env_alias.p, t.env_alias.data(), t.env_alias.size() * 8, hipMemcpyHostToDevice),
                      "env table update");
        }
    });
}

pt_status pt_scene_build_env_tables_device(pt_scene* s) {
    return guarded([&] {
        if (!s) throw Error(PT_ERR_INVALID_ARG, "null scene");
        s->data_version++;
        if (!s->uploaded) throw Error(PT_ERR_STATE, "scene not uploaded (pt_scene_upload)");
        const size_t n = static_cast<size_t>(s->env_w) * s->env_h;
        if (n == 0 || s->env.p == nullptr) return;
        if (s->env_pmf.n != n || s->env_alias.n != n)
            throw Error(PT_ERR_STATE, "the uploaded environment has no sampling tables (an all-zero environment?)");
        std::string err;
        hipError_t e = ptd::build_env_tables_device(s->env.p, s->env_w, s->env_h, s->env_pmf.p, s->env_alias.p,
                                                    s->stream, &err);
        if (e == hipErrorInvalidValue) throw Error(PT_ERR_INVALID_ARG, err.empty() ? "environment is all zero" : err);
        hip_check(e, "environment tables on the GPU");
        s->env_tables_on_device = true;
    });
}

pt_status pt_scene_update_emission(pt_scene* s, const float* emissive, const float* strength) {
    return guarded([&] {
        if (!s) throw Error(PT_ERR_INVALID_ARG, "null scene");
        s->data_version++;
        auto& M = s->host.materials;
        for (size_t i = 0; i < M.size(); ++i) {
            if (emissive) {
                M[i].emissive.x = emissive[3 * i + 0];
                M[i].emissive.y = emissive[3 * i + 1];
                M[i].emissive.z = emissive[3 * i + 2];
            }
            if (strength) M[i].emissive_strength = strength[i];
        }
        if (s->uploaded && s->materials.p && s->materials.n == M.size())
            hip_check(hipMemcpy(s->materials.p, M.data(), sizeof(pt_material) * M.size(), hipMemcpyHostToDevice),
                      "material update");
    });
}


pt_status pt_scene_rebuild_emissive_tables(pt_scene* s) {
    return guarded([&] {
        if (!s) throw Error(PT_ERR_INVALID_ARG, "null scene");
        s->data_version++;
        if (!s->uploaded) throw Error(PT_ERR_STATE, "scene not uploaded (pt_scene_upload)");
        const HostScene& h = s->host;
        const uint32_t nf = static_cast<uint32_t>(s->prim_order_host.size());
        std::vector<pt_face> faces_bvh(nf);
        for (uint32_t i = 0; i < nf; ++i) faces_bvh[i] = h.faces[s->prim_order_host[i]];
        ptamd::build_emissive_tables(h.vertices, faces_bvh, h.materials, s->lights);
        upload_emissive_tables(s);
        hip_check(hipStreamSynchronize(s->stream), "emissive table update");
    });
}

pt_status pt_scene_build_emissive_tables_device(pt_scene* s) {
    return guarded([&] {
        if (!s) throw Error(PT_ERR_INVALID_ARG, "null scene");
        s->data_version++;
        if (!s->uploaded) throw Error(PT_ERR_STATE, "scene not uploaded (pt_scene_upload)");
        const uint32_t nf = static_cast<uint32_t>(s->faces_bvh.n);
        if (nf == 0) return;
        if (s->tri_pmf.n != nf) throw Error(PT_ERR_STATE, "per-face light pmf not uploaded");
        uint32_t m = 0;
        uint4* em = nullptr;
        float2* table = nullptr;
        std::string err;
        hipError_t e = ptd::build_emissive_tables_device(s->vpos.p, s->faces_bvh.p, nf, s->materials.p,
                                                         static_cast<uint32_t>(s->materials.n), s->tri_pmf.p, &m, &em,
                                                         &table, s->stream, &err);
        if (e != hipSuccess) {
            hip_check(e, err.empty() ? "emissive tables on the GPU" : err.c_str());
        }
        s->emissive.adopt(em, m);
        s->tri_alias.adopt(table, m);
        s->emissive_tables_on_device = true;
    });
}

pt_status pt_alias_build_device(const float* weights, uint32_t n, float* pmf, float* alias) {
    return guarded([&] {
        if (n == 0) return;
        if (!weights || !pmf || !alias) throw Error(PT_ERR_INVALID_ARG, "null argument");
        ensure_device();
        std::string err;
        hipError_t e = ptd::build_alias_device(weights, n, pmf, reinterpret_cast<float2*>(alias), nullptr, &err);
        if (e == hipErrorInvalidValue) throw Error(PT_ERR_INVALID_ARG, err.empty() ? "weights sum to zero" : err);
        hip_check(e, "alias table on the GPU");
    });
}

pt_status pt_alias_build_device_noalloc(const float* weights, uint32_t n, float* pmf, float* alias) {
    return guarded([&] {
        if (n == 0) return;
        if (!weights || !pmf || !alias) throw Error(PT_ERR_INVALID_ARG, "null argument");
        if (n > PT_ALIAS_NOALLOC_MAX_WEIGHTS)
            throw Error(PT_ERR_INVALID_ARG, "pt_alias_build_device_noalloc: " + std::to_string(n) +
                                                " weights exceed the maximum of " +
                                                std::to_string(PT_ALIAS_NOALLOC_MAX_WEIGHTS));
        ensure_device();
        std::string err;
        hipError_t e =
            ptd::build_alias_device_noalloc(weights, n, pmf, reinterpret_cast<float2*>(alias), nullptr, &err);
        if (e == hipErrorInvalidValue) throw Error(PT_ERR_INVALID_ARG, err.empty() ? "weights sum to zero" : err);
        hip_check(e, "alias table on the GPU without allocation");
    });
}

pt_status pt_scene_wavefront_stats(const pt_scene* s, pt_wavefront_stats* out) {
    return guarded([&] {
        if (!s || !out) throw Error(PT_ERR_INVALID_ARG, "null argument");
        ptd::wavefront_stage_times(s->wf, out->stage_ms, out->launches);
    });
}

pt_status pt_scene_tune_wavefront(pt_scene* s, int max_trials, pt_wavefront_config* best, float* ms_per_spp) {
    return guarded([&] {
        if (!s) throw Error(PT_ERR_INVALID_ARG, "null scene");
        if (!s->uploaded) throw Error(PT_ERR_STATE, "scene not uploaded (pt_scene_upload)");
        if (max_trials < 1) throw Error(PT_ERR_INVALID_ARG, "max_trials must be positive");
        pt_film tf;
        tf.w = s->host.camera.width;
        tf.h = s->host.camera.height;
        size_t n = static_cast<size_t>(tf.w) * tf.h;
        tf.film_new.alloc(n);
        tf.film_acc.alloc(n);
        const bool counting = s->counting;
        const int spf = s->host.settings.sample_per_frame;
        s->counting = false;
        int trials = 0;
        // time per sample per pixel of one configuration: the batch is cut to
        // two chunks (warm-up run first: graph capture is not timed)
        auto measure = [&](const ptd::WavefrontConfig& c) -> float {
            int S = std::max(1, std::min(c.chunk_spp, spf));
            while (spf % S) --S;
            const int saved = s->host.settings.sample_per_frame;
            s->host.settings.sample_per_frame = std::min(spf, 2 * S);
            float ms = 0.0f;
            try {
                render_wavefront(s, &tf, c, 0);
                ms = render_wavefront(s, &tf, c, 1);
            } catch (...) {
                s->host.settings.sample_per_frame = saved;
                throw;
            }
            const float per = ms / static_cast<float>(s->host.settings.sample_per_frame);
            s->host.settings.sample_per_frame = saved;
            ++trials;
            return per;
        };
        ptd::WavefrontConfig cur = s->wf_cfg;
        cur.time_stages = 0;
        float cur_ms;
        try {
            cur_ms = measure(cur);
            auto try_axis = [&](auto set, std::initializer_list<int> values) {
                for (int v : values) {
                    if (trials >= max_trials) return;
                    ptd::WavefrontConfig c = cur;
                    set(c, v);
                    if (std::memcmp(&c, &cur, sizeof(c)) == 0) continue;
                    float t;
                    try {
                        t = measure(c);
                    } catch (const Error& e) {
                        if (e.code == PT_ERR_OUT_OF_MEMORY) continue;  // chunk too large for device memory
                        throw;
                    }
                    if (t < cur_ms) {
                        cur_ms = t;
                        cur = c;
                    }
                }
            };
            try_axis([](ptd::WavefrontConfig& c, int v) { c.bvh_width = v; }, {4, 2});
            try_axis([](ptd::WavefrontConfig& c, int v) { c.grab_extend = v; }, {64, 128, 256});
            try_axis([](ptd::WavefrontConfig& c, int v) { c.grab_shade = v; }, {128, 256, 512, 1024});
            try_axis([](ptd::WavefrontConfig& c, int v) { c.grab_shadow = v; }, {64, 128});
            try_axis([](ptd::WavefrontConfig& c, int v) { c.chunk_spp = v; }, {4, 8, 16, 32, 64});
            try_axis([](ptd::WavefrontConfig& c, int v) { c.extend_block = v; }, {64, 128, 256});
            try_axis([](ptd::WavefrontConfig& c, int v) { c.shade_block = v; }, {64, 128, 256});
            try_axis([](ptd::WavefrontConfig& c, int v) { c.sort_by_material = v; }, {0, 1});
            try_axis([](ptd::WavefrontConfig& c, int v) { c.waves_per_simd = v; }, {0, 1, 2, 4});
            try_axis([](ptd::WavefrontConfig& c, int v) { c.shade_waves = v; }, {0, 3, 4});
            try_axis([](ptd::WavefrontConfig& c, int v) { c.stack_lds = v; }, {16, 32});
        } catch (...) {
            s->counting = counting;
            throw;
        }
        s->counting = counting;
        s->wf_cfg = cur;
        if (best) *best = to_abi_config(cur);
        if (ms_per_spp) *ms_per_spp = cur_ms;
    });
}

pt_status pt_render_all(pt_scene* s, pt_film* f, int integrator, float* total_ms) {
    return guarded([&] {
        if (!s) throw Error(PT_ERR_INVALID_ARG, "null scene");
        float total = 0.0f;
        while (static_cast<int64_t>(s->frame_count) * s->host.settings.sample_per_frame <
               s->host.settings.target_samples) {
            float ms = 0.0f;
            pt_status st = pt_render_batch(s, f, integrator, &ms);
            if (st != PT_OK) throw Error(st, g_err);
            total += ms;
        }
        if (total_ms) *total_ms = total;
    });
}

pt_status pt_counters_enable(pt_scene* s, int enable) {
    return guarded([&] {
        if (!s) throw Error(PT_ERR_INVALID_ARG, "null scene");
        s->counting = enable != 0;
    });
}

pt_status pt_counters_get(pt_scene* s, pt_counters* out) {
    return guarded([&] {
        if (!s || !out) throw Error(PT_ERR_INVALID_ARG, "null argument");
        scene_device(s);
        ptd::DevCounters c;
        hip_check(hipMemcpy(&c, s->counters.p, sizeof(c), hipMemcpyDeviceToHost), "counter download");
        out->closest_rays = c.closest_rays;
        out->shadow_rays = c.shadow_rays;
        out->node_visits = c.node_visits;
        out->tri_tests = c.tri_tests;
        out->paths = c.paths;
        out->stack_overflows = c.stack_overflows;
        out->shadow_node_visits = c.shadow_node_visits;
        out->shadow_tri_tests = c.shadow_tri_tests;
        out->shade_items = c.shade_items;
        out->shade_env_lookups = c.shade_env_lookups;
        out->shade_hits = c.shade_hits;
        out->shade_light_samples = c.shade_light_samples;
        out->shade_shadow_entries = c.shade_shadow_entries;
        out->shade_continuations = c.shade_continuations;
        out->shade_radiance_adds = c.shade_radiance_adds;
    });
}

pt_status pt_counters_reset(pt_scene* s) {
    return guarded([&] {
        if (!s) throw Error(PT_ERR_INVALID_ARG, "null scene");
        scene_device(s);
        hip_check(hipMemset(s->counters.p, 0, sizeof(ptd::DevCounters)), "hipMemset");
        hip_check(hipDeviceSynchronize(), "counter reset");
    });
}

pt_status pt_probe_sampler(int sampler, uint32_t seed, const int32_t* xyi, int n, int dims, uint32_t* out_u,
                           float* out_f) {
    return guarded([&] {
        if (n < 0 || dims <= 0 || (n > 0 && (!xyi || !out_u || !out_f)) || (sampler != 0 && sampler != 1))
            throw Error(PT_ERR_INVALID_ARG, "bad argument");
        if (n == 0) return;
        ensure_device();
        DBuf<int32_t> d_xyi;
        DBuf<uint32_t> d_u;
        DBuf<float> d_f;
        d_xyi.upload(xyi, static_cast<size_t>(n) * 3, nullptr);
        d_u.alloc(static_cast<size_t>(n) * dims);
        d_f.alloc(static_cast<size_t>(n) * dims);
        hip_check(ptd::launch_probe_sampler(sampler, seed, d_xyi.p, n, dims, d_u.p, d_f.p, nullptr), "probe launch");
        hip_check(hipMemcpy(out_u, d_u.p, 4 * d_u.n, hipMemcpyDeviceToHost), "probe download");
        hip_check(hipMemcpy(out_f, d_f.p, 4 * d_f.n, hipMemcpyDeviceToHost), "probe download");
    });
}

pt_status pt_trace_rays(pt_scene* s, const float* rays8, int n, int any_hit, float* t_out, int32_t* face_out,
                        float* bary_out) {
    return guarded([&] {
        if (!s || n < 0 || (n > 0 && (!rays8 || !t_out || !face_out))) throw Error(PT_ERR_INVALID_ARG, "bad argument");
        if (!s->accel_built) throw Error(PT_ERR_STATE, "acceleration structure not built");
        if (n == 0) return;
        ptd::DevScene v = make_view(s);
        DBuf<float> d_r, d_t, d_b;
        DBuf<int32_t> d_face;
        d_r.upload(rays8, static_cast<size_t>(n) * 8, s->stream);
        d_t.alloc(n);
        d_face.alloc(n);
        if (bary_out) d_b.alloc(static_cast<size_t>(n) * 2);
        const bool wide = (any_hit & 2) != 0;
        if (wide && !s->wnodes.p) throw Error(PT_ERR_STATE, "the scene has no 4-wide BVH");
        hip_check(ptd::launch_trace_rays(v, d_r.p, n, any_hit & 1, wide, d_t.p, d_face.p, d_b.p, s->stream),
                  "trace launch");
        hip_check(hipStreamSynchronize(s->stream), "trace");
        hip_check(hipMemcpy(t_out, d_t.p, 4 * static_cast<size_t>(n), hipMemcpyDeviceToHost), "download");
        hip_check(hipMemcpy(face_out, d_face.p, 4 * static_cast<size_t>(n), hipMemcpyDeviceToHost), "download");
        if (bary_out) hip_check(hipMemcpy(bary_out, d_b.p, 8 * static_cast<size_t>(n), hipMemcpyDeviceToHost), "download");
    });
}

pt_status pt_probe_bsdf(const pt_material* mats, int n, const float* dirs9, const int32_t* entering, const float* u3,
                        float* out12) {
    return guarded([&] {
        if (n < 0 || (n > 0 && (!mats || !dirs9 || !entering || !u3 || !out12)))
            throw Error(PT_ERR_INVALID_ARG, "bad argument");
        if (n == 0) return;
        ensure_device();
        DBuf<pt_material> dm;
        DBuf<float> dd, du, dout;
        DBuf<int32_t> de;
        dm.upload(mats, n, nullptr);
        dd.upload(dirs9, static_cast<size_t>(n) * 9, nullptr);
        du.upload(u3, static_cast<size_t>(n) * 3, nullptr);
        de.upload(entering, n, nullptr);
        dout.alloc(static_cast<size_t>(n) * 12);
        hip_check(ptd::launch_probe_bsdf(dm.p, n, dd.p, de.p, du.p, dout.p, nullptr), "bsdf probe launch");
        hip_check(hipDeviceSynchronize(), "bsdf probe");
        hip_check(hipMemcpy(out12, dout.p, sizeof(float) * 12 * static_cast<size_t>(n), hipMemcpyDeviceToHost),
                  "download");
    });
}

pt_status pt_probe_copy(size_t bytes, int iters, float* gbps) {
    return guarded([&] {
        if (bytes < 16 || iters <= 0 || !gbps) throw Error(PT_ERR_INVALID_ARG, "bad argument");
        ensure_device();
        size_t n = bytes / 16;
        DBuf<float4> a, b;
        a.alloc(n);
        b.alloc(n);
        hip_check(hipMemset(a.p, 0, n * 16), "hipMemset");
        hipEvent_t e0, e1;
        hip_check(hipEventCreate(&e0), "hipEventCreate");
        hip_check(hipEventCreate(&e1), "hipEventCreate");
        hip_check(ptd::launch_copy_probe(a.p, b.p, n, nullptr), "copy warmup");
        hip_check(hipEventRecord(e0, nullptr), "record");
        for (int i = 0; i < iters; ++i) hip_check(ptd::launch_copy_probe(a.p, b.p, n, nullptr), "copy");
        hip_check(hipEventRecord(e1, nullptr), "record");
        hip_check(hipEventSynchronize(e1), "sync");
        float ms = 0.0f;
        hip_check(hipEventElapsedTime(&ms, e0, e1), "elapsed");
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        *gbps = static_cast<float>(2.0 * n * 16.0 * iters / (ms * 1e-3) / 1e9);
    });
}

pt_status pt_probe_reciprocal(uint64_t* mismatches, uint64_t* compared) {
    return guarded([&] {
        if (!mismatches || !compared) throw Error(PT_ERR_INVALID_ARG, "bad argument");
        ensure_device();
        DBuf<unsigned long long> d;
        d.alloc(1);
        hip_check(hipMemset(d.p, 0, sizeof(unsigned long long)), "hipMemset");
        hip_check(ptd::launch_probe_reciprocal(d.p, nullptr), "reciprocal probe launch");
        unsigned long long bad = 0;
        hip_check(hipMemcpy(&bad, d.p, sizeof(bad), hipMemcpyDeviceToHost), "reciprocal probe download");
        *mismatches = bad;
        *compared = 2ull * 252ull * (1ull << 23);
    });
}

}  // extern "C"
