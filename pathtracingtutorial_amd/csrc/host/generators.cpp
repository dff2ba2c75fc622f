// Seeded procedural workloads.
//
//  * Cornell box, 36 triangles: five 1x1 walls of the box [-0.5,0.5]^3 with
//    the +z side open (white floor/ceiling/back, red left, green right), a
//    tall mirror box (0.3x0.6x0.3, metallic 1, roughness 0) turned +18 deg
//    about y, a diffuse 0.3^3 box turned -18 deg, and a 0.25x0.25 ceiling
//    light of radiance (17,12,4).  Pinhole camera at (0,0,1.87), 40 deg,
//    512x512, 16 spp in one batch, depth 8, PCG sampler.
//  * Displaced icospheres: N instances of a level-L subdivided icosahedron
//    (L=5: 20,480 triangles, the tessellation of the reference asset
//    models/sphere2.obj), radius U[1,2], vertices pushed along the normal by
//    seeded value noise of amplitude 0.1, centres U[-10,10]^3, one random
//    Disney material each, a 40x40 ground quad at y=-11, and a synthetic sky
//    + Gaussian sun environment (peak 1e4).  1920x1080, 64 spp, depth 16, Sobol.
//
// Every number comes from a splitmix64 stream of the seed, so the scene is
// identical on every machine.
#include <cmath>
#include <cstring>
#include <map>

#include "host_scene.h"

namespace ptamd {

namespace {

struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    // uniform in [0,1) with 24 bits
    float uniform() { return static_cast<float>(next() >> 40) * (1.0f / 16777216.0f); }
    float uniform(float a, float b) { return a + (b - a) * uniform(); }
};

struct P3 {
    double x, y, z;
};

pt_vec3p v3p(float x, float y, float z) { return {x, y, z, 0.0f}; }

void add_quad_cuv(const float c[3], const float u[3], const float v[3], int mat, HostScene& s) {
    float a[3], b[3], cc[3], d[3];
    for (int i = 0; i < 3; ++i) {
        a[i] = c[i] - u[i] - v[i];
        b[i] = c[i] + u[i] - v[i];
        cc[i] = c[i] + u[i] + v[i];
        d[i] = c[i] - u[i] + v[i];
    }
    build_quad(a, b, cc, d, mat, s.vertices, s.faces);
}

// Axis-aligned box of half-extents h centred at c, turned by deg about y,
// six outward-facing quads.
void add_box(const float c[3], const float h[3], float deg, int mat, HostScene& s) {
    const double rad = deg * 3.14159265358979323846 / 180.0;
    const float cs = static_cast<float>(std::cos(rad)), sn = static_cast<float>(std::sin(rad));
    auto rot = [&](const float v[3], float o[3]) {  // rotation about +y
        o[0] = cs * v[0] + sn * v[2];
        o[1] = v[1];
        o[2] = -sn * v[0] + cs * v[2];
    };
    // (centre offset, u edge, v edge) in box space with u x v pointing outward
    const float faces[6][3][3] = {
        {{0, h[1], 0}, {h[0], 0, 0}, {0, 0, -h[2]}},   // top +y
        {{0, -h[1], 0}, {h[0], 0, 0}, {0, 0, h[2]}},   // bottom -y
        {{h[0], 0, 0}, {0, h[1], 0}, {0, 0, h[2]}},    // +x
        {{-h[0], 0, 0}, {0, 0, h[2]}, {0, h[1], 0}},   // -x
        {{0, 0, h[2]}, {h[0], 0, 0}, {0, h[1], 0}},    // +z
        {{0, 0, -h[2]}, {0, h[1], 0}, {h[0], 0, 0}},   // -z
    };
    for (const auto& f : faces) {
        float oc[3], ou[3], ov[3];
        rot(f[0], oc);
        rot(f[1], ou);
        rot(f[2], ov);
        for (int i = 0; i < 3; ++i) oc[i] += c[i];
        add_quad_cuv(oc, ou, ov, mat, s);
    }
}

// Lattice value noise in [-1,1] with smoothstep trilinear interpolation.
float hash3(int x, int y, int z, uint32_t seed) {
    uint32_t h = static_cast<uint32_t>(x) * 0x8da6b343u ^ static_cast<uint32_t>(y) * 0xd8163841u ^
                 static_cast<uint32_t>(z) * 0xcb1ab31fu ^ seed * 0x9E3779B1u;
    h ^= h >> 16;
    h *= 0x7feb352du;
    h ^= h >> 15;
    h *= 0x846ca68bu;
    h ^= h >> 16;
    return static_cast<float>(h >> 8) * (2.0f / 16777216.0f) - 1.0f;
}

float value_noise(float x, float y, float z, uint32_t seed) {
    float fx = std::floor(x), fy = std::floor(y), fz = std::floor(z);
    int ix = static_cast<int>(fx), iy = static_cast<int>(fy), iz = static_cast<int>(fz);
    float tx = x - fx, ty = y - fy, tz = z - fz;
    tx = tx * tx * (3 - 2 * tx);
    ty = ty * ty * (3 - 2 * ty);
    tz = tz * tz * (3 - 2 * tz);
    float c[2][2][2];
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int d = 0; d < 2; ++d) c[a][b][d] = hash3(ix + a, iy + b, iz + d, seed);
    float x00 = c[0][0][0] + (c[1][0][0] - c[0][0][0]) * tx;
    float x10 = c[0][1][0] + (c[1][1][0] - c[0][1][0]) * tx;
    float x01 = c[0][0][1] + (c[1][0][1] - c[0][0][1]) * tx;
    float x11 = c[0][1][1] + (c[1][1][1] - c[0][1][1]) * tx;
    float y0 = x00 + (x10 - x00) * ty;
    float y1 = x01 + (x11 - x01) * ty;
    return y0 + (y1 - y0) * tz;
}

// Unit icosphere subdivided `level` times (midpoints re-projected to the sphere).
void icosphere(int level, std::vector<P3>& verts, std::vector<uint32_t>& tris) {
    const double t = (1.0 + std::sqrt(5.0)) / 2.0;
    verts = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
             {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
    for (auto& v : verts) {
        double l = std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z);
        v = {v.x / l, v.y / l, v.z / l};
    }
    tris = {0, 11, 5, 0, 5, 1, 0, 1, 7, 0, 7, 10, 0, 10, 11, 1, 5, 9, 5, 11, 4, 11, 10, 2, 10, 7, 6, 7, 1, 8,
            3, 9, 4, 3, 4, 2, 3, 2, 6, 3, 6, 8, 3, 8, 9, 4, 9, 5, 2, 4, 11, 6, 2, 10, 8, 6, 7, 9, 8, 1};
    for (int l = 0; l < level; ++l) {
        std::map<uint64_t, uint32_t> mid;
        auto midpoint = [&](uint32_t a, uint32_t b) -> uint32_t {
            uint64_t key = a < b ? (static_cast<uint64_t>(a) << 32 | b) : (static_cast<uint64_t>(b) << 32 | a);
            auto it = mid.find(key);
            if (it != mid.end()) return it->second;
            P3 m{(verts[a].x + verts[b].x) * 0.5, (verts[a].y + verts[b].y) * 0.5, (verts[a].z + verts[b].z) * 0.5};
            double len = std::sqrt(m.x * m.x + m.y * m.y + m.z * m.z);
            verts.push_back({m.x / len, m.y / len, m.z / len});
            uint32_t id = static_cast<uint32_t>(verts.size() - 1);
            mid[key] = id;
            return id;
        };
        std::vector<uint32_t> nt;
        nt.reserve(tris.size() * 4);
        for (size_t i = 0; i < tris.size(); i += 3) {
            uint32_t a = tris[i], b = tris[i + 1], c = tris[i + 2];
            uint32_t ab = midpoint(a, b), bc = midpoint(b, c), ca = midpoint(c, a);
            uint32_t q[12] = {a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca};
            nt.insert(nt.end(), q, q + 12);
        }
        tris.swap(nt);
    }
}

}  // namespace

// Sky gradient plus a Gaussian sun lobe.  Texel (x, y) looks along the
// direction the reference's equirect mapping assigns to its centre
// (lights/envmap.h:27-32): phi = (u-0.5)*2pi, theta = v*pi, dir =
// (cos phi sin theta, cos theta, sin phi sin theta).
void generate_sky_env(Image& env, int w, int h, uint64_t seed) {
    Rng rng(seed ^ 0x5EEDull);
    const double pi = 3.14159265358979323846;
    double az = rng.uniform() * 2.0 * pi;
    double el = (20.0 + 40.0 * rng.uniform()) * pi / 180.0;
    double sx = std::cos(el) * std::cos(az), sy = std::sin(el), sz = std::cos(el) * std::sin(az);
    const double sigma = 1.0 * pi / 180.0;  // 1 degree
    env.width = w;
    env.height = h;
    env.wrap_s = env.wrap_t = 10497;
    env.linear = 1;
    env.rgba.assign(static_cast<size_t>(w) * h * 4, 0.0f);
    for (int y = 0; y < h; ++y) {
        double theta = (y + 0.5) / h * pi;
        for (int x = 0; x < w; ++x) {
            double phi = ((x + 0.5) / w - 0.5) * 2.0 * pi;
            double dx = std::cos(phi) * std::sin(theta), dy = std::cos(theta), dz = std::sin(phi) * std::sin(theta);
            double r, g, b;
            if (dy >= 0.0) {
                double t = std::sqrt(dy);
                r = 1.0 + (0.25 - 1.0) * t;
                g = 0.95 + (0.45 - 0.95) * t;
                b = 0.9 + (0.95 - 0.9) * t;
            } else {
                double t = std::fmin(1.0, -dy * 4.0);
                r = 0.35 + (0.12 - 0.35) * t;
                g = 0.33 + (0.11 - 0.33) * t;
                b = 0.3 + (0.1 - 0.3) * t;
            }
            double c = std::fmax(-1.0, std::fmin(1.0, dx * sx + dy * sy + dz * sz));
            double ang = std::acos(c);
            double sun = 1e4 * std::exp(-ang * ang / (2.0 * sigma * sigma));
            float* o = &env.rgba[(static_cast<size_t>(y) * w + x) * 4];
            // the renderer scales env radiance by 2 pi (envmap.h:43-45); the sky is set
            // so that the scaled sky stays near 1.5 while the sun peak is 1e4
            const double sky = 0.25;
            o[0] = static_cast<float>(sky * r + sun * 1.0);
            o[1] = static_cast<float>(sky * g + sun * 0.93);
            o[2] = static_cast<float>(sky * b + sun * 0.82);
            o[3] = 1.0f;
        }
    }
}

bool generate_cornell(HostScene& s) {
    init_scene_defaults(s);
    s.path = "<cornell>";
    auto mk = [](float r, float g, float b) {
        pt_material m = default_material();
        m.base_color = v3p(r, g, b);
        m.roughness = 1.0f;
        return m;
    };
    s.materials.push_back(mk(0.73f, 0.73f, 0.73f));  // 0 white
    s.materials.push_back(mk(0.65f, 0.05f, 0.05f));  // 1 red
    s.materials.push_back(mk(0.12f, 0.45f, 0.15f));  // 2 green
    pt_material light = default_material();
    light.emissive = v3p(17.0f, 12.0f, 4.0f);
    light.emissive_strength = 1.0f;
    s.materials.push_back(light);                    // 3 light
    pt_material mirror = default_material();
    mirror.base_color = v3p(0.95f, 0.95f, 0.95f);
    mirror.metallic = 1.0f;
    mirror.roughness = 0.0f;
    s.materials.push_back(mirror);                   // 4 mirror

    const float H = 0.5f;
    {  // walls: (centre, u, v) with u x v pointing into the box
        float c[3], u[3], v[3];
        auto wall = [&](float cx, float cy, float cz, float ux, float uy, float uz, float vx, float vy, float vz,
                        int m) {
            c[0] = cx; c[1] = cy; c[2] = cz;
            u[0] = ux; u[1] = uy; u[2] = uz;
            v[0] = vx; v[1] = vy; v[2] = vz;
            add_quad_cuv(c, u, v, m, s);
        };
        wall(0, -H, 0, H, 0, 0, 0, 0, -H, 0);  // floor, normal +y
        wall(0, H, 0, H, 0, 0, 0, 0, H, 0);    // ceiling, normal -y
        wall(0, 0, -H, H, 0, 0, 0, H, 0, 0);   // back, normal +z
        wall(-H, 0, 0, 0, H, 0, 0, 0, H, 1);   // left (red), normal +x
        wall(H, 0, 0, 0, 0, H, 0, H, 0, 2);    // right (green), normal -x
        wall(0, 0.499f, 0, 0.125f, 0, 0, 0, 0, 0.125f, 3);  // light, normal -y
    }
    {
        float c[3] = {-0.17f, -0.2f, -0.12f}, h[3] = {0.15f, 0.3f, 0.15f};
        add_box(c, h, 18.0f, 4, s);
    }
    {
        float c[3] = {0.17f, -0.35f, 0.12f}, h[3] = {0.15f, 0.15f, 0.15f};
        add_box(c, h, -18.0f, 0, s);
    }
    s.camera = default_camera();
    s.camera.position = v3p(0.0f, 0.0f, 1.87f);
    s.camera.look_at = v3p(0.0f, 0.0f, 0.0f);
    s.camera.up = v3p(0.0f, 1.0f, 0.0f);
    s.camera.fov = 40.0f;
    s.camera.width = 512;
    s.camera.height = 512;
    s.settings.sample_per_frame = 16;
    s.settings.target_samples = 16;
    s.settings.max_depth = 8;
    s.settings.sampler = 0;
    s.settings.seed = 0;
    s.output_file = "cornell.hdr";
    s.bvh_opts = {24, 4, 0, 0};
    s.has_bvh_opts = true;
    return true;
}

// Differentiable Cornell box: the Cornell geometry with eight materials
// (floor, ceiling, textured back wall, red, green, light, tall box, short
// box), a linear 256x256 RGB albedo texture on the back wall, a 512x256
// environment seen around the open +z side, 256x256, 64 spp, depth 6, PCG.
// The seeded values are the ground truth of an inverse-rendering problem:
// value-noise albedo in [0.1, 0.9], roughness U[0.2, 0.9] and metallic
// U[0, 0.6] for the six reflective materials, and a smooth two-colour sky
// gradient of magnitude about 1.
bool generate_diff_cornell(HostScene& s, uint64_t seed) {
    generate_cornell(s);
    s.path = "<diff-cornell>";
    Rng rng(seed ^ 0xD1FFull);
    std::vector<pt_material> old = s.materials;
    s.materials.clear();
    auto mk = [](float r, float g, float b) {
        pt_material m = default_material();
        m.base_color = v3p(r, g, b);
        m.roughness = 1.0f;
        return m;
    };
    s.materials.push_back(mk(0.73f, 0.73f, 0.73f));  // 0 floor
    s.materials.push_back(mk(0.73f, 0.73f, 0.73f));  // 1 ceiling
    s.materials.push_back(mk(1.0f, 1.0f, 1.0f));     // 2 back wall (albedo texture)
    s.materials.push_back(mk(0.65f, 0.05f, 0.05f));  // 3 left
    s.materials.push_back(mk(0.12f, 0.45f, 0.15f));  // 4 right
    s.materials.push_back(old[3]);                   // 5 light
    s.materials.push_back(mk(0.95f, 0.95f, 0.95f));  // 6 tall box
    s.materials.push_back(mk(0.73f, 0.73f, 0.73f));  // 7 short box
    for (int m : {0, 1, 2, 3, 4, 6, 7}) {
        s.materials[m].roughness = rng.uniform(0.2f, 0.9f);
        s.materials[m].metallic = rng.uniform(0.0f, 0.6f);
    }
    // faces in generate_cornell order: floor 2, ceiling 2, back 2, left 2, right 2, light 2, tall box 12, short box 12
    const int remap[] = {0, 0, 1, 1, 2, 2, 3, 3, 4, 4, 5, 5};
    for (size_t i = 0; i < s.faces.size(); ++i) {
        if (i < 12) s.faces[i].material_id = remap[i];
        else if (i < 24) s.faces[i].material_id = 6;
        else s.faces[i].material_id = 7;
    }
    Image tex;
    tex.width = tex.height = 256;
    tex.wrap_s = tex.wrap_t = 10497;
    tex.linear = 1;
    tex.rgba.resize(256 * 256 * 4);
    const uint32_t ts = static_cast<uint32_t>(rng.next());
    for (int y = 0; y < 256; ++y)
        for (int x = 0; x < 256; ++x) {
            float* o = &tex.rgba[(static_cast<size_t>(y) * 256 + x) * 4];
            for (int c = 0; c < 3; ++c) {
                float n = 0.6f * value_noise(x / 32.0f, y / 32.0f, c * 7.0f, ts) +
                          0.4f * value_noise(x / 9.0f, y / 9.0f, c * 7.0f + 3.0f, ts);
                o[c] = std::fmin(0.9f, std::fmax(0.1f, 0.5f + 0.45f * n));
            }
            o[3] = 1.0f;
        }
    s.textures.clear();
    s.textures.push_back(std::move(tex));
    s.materials[2].base_color_tex_id = 0;
    Image& env = s.env;
    env.width = 512;
    env.height = 256;
    env.wrap_s = env.wrap_t = 10497;
    env.linear = 1;
    env.rgba.assign(512 * 256 * 4, 0.0f);
    float top[3], bot[3];
    for (int c = 0; c < 3; ++c) {
        top[c] = rng.uniform(0.4f, 1.6f);
        bot[c] = rng.uniform(0.2f, 1.0f);
    }
    for (int y = 0; y < 256; ++y)
        for (int x = 0; x < 512; ++x) {
            const float t = (y + 0.5f) / 256.0f;
            const float wob = 1.0f + 0.15f * std::sin(6.2831853f * (x + 0.5f) / 512.0f);
            float* o = &env.rgba[(static_cast<size_t>(y) * 512 + x) * 4];
            for (int c = 0; c < 3; ++c) o[c] = (top[c] + (bot[c] - top[c]) * t) * wob;
            o[3] = 1.0f;
        }
    s.camera.width = 256;
    s.camera.height = 256;
    s.settings.sample_per_frame = 64;
    s.settings.target_samples = 64;
    s.settings.max_depth = 6;
    s.output_file = "diff_cornell.hdr";
    return true;
}

bool generate_spheres(HostScene& s, uint64_t seed, int instances, int level, int env_w, int env_h) {
    init_scene_defaults(s);
    s.path = "<spheres>";
    Rng rng(seed);
    std::vector<P3> unit;
    std::vector<uint32_t> tris;
    icosphere(level, unit, tris);
    const size_t nv = unit.size(), nt = tris.size() / 3;
    s.vertices.reserve(nv * instances + 4);
    s.faces.reserve(nt * instances + 2);
    for (int i = 0; i < instances; ++i) {
        pt_material m = default_material();
        m.base_color = v3p(rng.uniform(), rng.uniform(), rng.uniform());
        m.metallic = rng.uniform();
        m.roughness = std::fmax(rng.uniform(), 0.05f);
        m.specular_tint = rng.uniform();
        m.anisotropic = rng.uniform();
        m.sheen = rng.uniform();
        m.sheen_tint = rng.uniform();
        m.clearcoat = rng.uniform();
        m.clearcoat_roughness = rng.uniform();
        m.spec_trans = rng.uniform();
        m.subsurface = rng.uniform();
        m.ior = 1.5f;
        s.materials.push_back(m);
    }
    pt_material ground = default_material();
    ground.base_color = v3p(0.5f, 0.5f, 0.5f);
    ground.roughness = 0.8f;
    s.materials.push_back(ground);

    for (int i = 0; i < instances; ++i) {
        float cx = rng.uniform(-10.0f, 10.0f), cy = rng.uniform(-10.0f, 10.0f), cz = rng.uniform(-10.0f, 10.0f);
        float radius = rng.uniform(1.0f, 2.0f);
        float ox = rng.uniform(0.0f, 100.0f), oy = rng.uniform(0.0f, 100.0f), oz = rng.uniform(0.0f, 100.0f);
        uint32_t nseed = static_cast<uint32_t>(rng.next());
        uint32_t base = static_cast<uint32_t>(s.vertices.size());
        std::vector<float> px(nv * 3);
        for (size_t k = 0; k < nv; ++k) {
            const P3& u = unit[k];
            float n = value_noise(static_cast<float>(u.x) * 2.5f + ox, static_cast<float>(u.y) * 2.5f + oy,
                                  static_cast<float>(u.z) * 2.5f + oz, nseed);
            float r = radius * (1.0f + 0.1f * n);
            px[k * 3 + 0] = cx + r * static_cast<float>(u.x);
            px[k * 3 + 1] = cy + r * static_cast<float>(u.y);
            px[k * 3 + 2] = cz + r * static_cast<float>(u.z);
        }
        // area-weighted smooth normals of the displaced surface
        std::vector<double> acc(nv * 3, 0.0);
        for (size_t t = 0; t < nt; ++t) {
            uint32_t a = tris[3 * t], b = tris[3 * t + 1], c = tris[3 * t + 2];
            double e1[3], e2[3];
            for (int j = 0; j < 3; ++j) { e1[j] = px[b * 3 + j] - px[a * 3 + j]; e2[j] = px[c * 3 + j] - px[a * 3 + j]; }
            double n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
            for (uint32_t id : {a, b, c})
                for (int j = 0; j < 3; ++j) acc[id * 3 + j] += n[j];
        }
        for (size_t k = 0; k < nv; ++k) {
            pt_vertex v;
            std::memset(&v, 0, sizeof(v));
            v.position = v3p(px[k * 3], px[k * 3 + 1], px[k * 3 + 2]);
            double l = std::sqrt(acc[k * 3] * acc[k * 3] + acc[k * 3 + 1] * acc[k * 3 + 1] + acc[k * 3 + 2] * acc[k * 3 + 2]);
            if (l > 0) v.normal = v3p(static_cast<float>(acc[k * 3] / l), static_cast<float>(acc[k * 3 + 1] / l),
                                      static_cast<float>(acc[k * 3 + 2] / l));
            const P3& u = unit[k];
            v.u = static_cast<float>(0.5 + std::atan2(u.z, u.x) / (2.0 * 3.14159265358979323846));
            v.v = static_cast<float>(0.5 - std::asin(std::fmax(-1.0, std::fmin(1.0, u.y))) / 3.14159265358979323846);
            s.vertices.push_back(v);
        }
        for (size_t t = 0; t < nt; ++t)
            s.faces.push_back({base + tris[3 * t], base + tris[3 * t + 1], base + tris[3 * t + 2], i});
    }
    {
        float a[3] = {-20.0f, -11.0f, 20.0f}, b[3] = {20.0f, -11.0f, 20.0f}, c[3] = {20.0f, -11.0f, -20.0f},
              d[3] = {-20.0f, -11.0f, -20.0f};
        build_quad(a, b, c, d, instances, s.vertices, s.faces);
    }
    if (env_w > 0 && env_h > 0) generate_sky_env(s.env, env_w, env_h, seed);
    s.camera = default_camera();
    s.camera.position = v3p(0.0f, 4.0f, 32.0f);
    s.camera.look_at = v3p(0.0f, 0.0f, 0.0f);
    s.camera.up = v3p(0.0f, 1.0f, 0.0f);
    s.camera.fov = 40.0f;
    s.camera.width = 1920;
    s.camera.height = 1080;
    s.settings.sample_per_frame = 64;
    s.settings.target_samples = 64;
    s.settings.max_depth = 16;
    s.settings.sampler = 1;
    s.settings.seed = 0;  // reference sampler sequences; the seed only shapes the scene
    s.output_file = "spheres.hdr";
    s.bvh_opts = {32, 5, 4, 0};
    s.has_bvh_opts = true;
    return true;
}

// One job of the batched workload: the icosphere scene with a triangle count
// log-uniform in [1e3, 5e5] (sphere level 2..5, up to 64 instances), a pixel
// count log-uniform in [128^2, 1024^2] at an aspect ratio in [1, 16/9], spp
// from {4, 8, 16, 32, 64}, depth 12 and its own 512x256 sky.
bool generate_batch_job(HostScene& s, uint64_t seed) {
    Rng rng(seed ^ 0xBA7Cull);
    const double tris = std::exp(std::log(1e3) + rng.uniform() * (std::log(5e5) - std::log(1e3)));
    int level = 2;
    for (int l = 5; l >= 2; --l)
        if (tris >= 2.0 * 20.0 * std::pow(4.0, l)) {
            level = l;
            break;
        }
    const double per = 20.0 * std::pow(4.0, level);
    const int instances = static_cast<int>(std::max(1.0, std::min(64.0, std::round(tris / per))));
    generate_spheres(s, seed, instances, level, 512, 256);
    const double pix = std::exp(std::log(128.0 * 128.0) + rng.uniform() * (std::log(1024.0 * 1024.0) - std::log(128.0 * 128.0)));
    const double aspect = 1.0 + rng.uniform() * (16.0 / 9.0 - 1.0);
    s.camera.width = std::max(16, static_cast<int>(std::lround(std::sqrt(pix * aspect))));
    s.camera.height = std::max(16, static_cast<int>(std::lround(pix / s.camera.width)));
    const int spps[5] = {4, 8, 16, 32, 64};
    const int spp = spps[std::min<int>(4, static_cast<int>(rng.uniform() * 5.0f))];
    s.settings.sample_per_frame = spp;
    s.settings.target_samples = spp;
    s.settings.max_depth = 12;
    s.path = "<batch-job>";
    s.output_file = "batch_job.hdr";
    return true;
}

// Differentiable multi-view scene: 100 displaced level-5 icospheres
// (2,048,000 triangles) of radius U[0.25, 0.55] with centres in [-4, 4]^3
// sharing 16 random Disney materials, a 20x20 ground quad at y = -4.6 with a
// linear 1024x1024 RGB albedo texture (value noise), and a 1024x512 sky
// (smooth gradient, no sun).  512x512, 32 spp, depth 10, PCG.  Cameras for
// the 16-view ring are set by the caller.
bool generate_diff_views(HostScene& s, uint64_t seed) {
    init_scene_defaults(s);
    s.path = "<diff-views>";
    Rng rng(seed ^ 0x1637ull);
    for (int m = 0; m < 16; ++m) {
        pt_material mat = default_material();
        mat.base_color = v3p(rng.uniform(0.1f, 0.9f), rng.uniform(0.1f, 0.9f), rng.uniform(0.1f, 0.9f));
        mat.roughness = rng.uniform(0.2f, 0.9f);
        mat.metallic = rng.uniform(0.0f, 0.7f);
        s.materials.push_back(mat);
    }
    pt_material ground = default_material();
    ground.base_color = v3p(1.0f, 1.0f, 1.0f);
    ground.roughness = 0.8f;
    ground.base_color_tex_id = 0;
    s.materials.push_back(ground);  // 16
    std::vector<P3> unit;
    std::vector<uint32_t> tris;
    icosphere(5, unit, tris);
    const uint32_t ns = 100;
    const uint32_t noise_seed = static_cast<uint32_t>(rng.next());
    for (uint32_t i = 0; i < ns; ++i) {
        const float cx = rng.uniform(-4.0f, 4.0f), cy = rng.uniform(-3.5f, 3.5f), cz = rng.uniform(-4.0f, 4.0f);
        const float r = rng.uniform(0.25f, 0.55f);
        const uint32_t base = static_cast<uint32_t>(s.vertices.size());
        for (const P3& u : unit) {
            const float disp = 0.1f * r * value_noise(static_cast<float>(u.x) * 3.0f + i, static_cast<float>(u.y) * 3.0f,
                                                      static_cast<float>(u.z) * 3.0f, noise_seed);
            const float rr = r + disp;
            pt_vertex v;
            std::memset(&v, 0, sizeof(v));
            v.position = v3p(cx + rr * static_cast<float>(u.x), cy + rr * static_cast<float>(u.y),
                             cz + rr * static_cast<float>(u.z));
            v.normal = v3p(static_cast<float>(u.x), static_cast<float>(u.y), static_cast<float>(u.z));
            s.vertices.push_back(v);
        }
        for (size_t t = 0; t + 2 < tris.size(); t += 3)
            s.faces.push_back({base + tris[t], base + tris[t + 1], base + tris[t + 2], static_cast<int32_t>(i % 16)});
    }
    {
        float a[3] = {-10.0f, -4.6f, 10.0f}, b[3] = {10.0f, -4.6f, 10.0f}, c[3] = {10.0f, -4.6f, -10.0f},
              d[3] = {-10.0f, -4.6f, -10.0f};
        build_quad(a, b, c, d, 16, s.vertices, s.faces);
    }
    Image tex;
    tex.width = tex.height = 1024;
    tex.wrap_s = tex.wrap_t = 10497;
    tex.linear = 1;
    tex.rgba.resize(static_cast<size_t>(1024) * 1024 * 4);
    for (int y = 0; y < 1024; ++y)
        for (int x = 0; x < 1024; ++x) {
            float* o = &tex.rgba[(static_cast<size_t>(y) * 1024 + x) * 4];
            for (int ch = 0; ch < 3; ++ch) {
                float n = 0.6f * value_noise(x / 96.0f, y / 96.0f, ch * 5.0f, noise_seed + 1) +
                          0.4f * value_noise(x / 23.0f, y / 23.0f, ch * 5.0f + 2.0f, noise_seed + 1);
                o[ch] = std::fmin(0.9f, std::fmax(0.1f, 0.5f + 0.45f * n));
            }
            o[3] = 1.0f;
        }
    s.textures.push_back(std::move(tex));
    Image& env = s.env;
    env.width = 1024;
    env.height = 512;
    env.wrap_s = env.wrap_t = 10497;
    env.linear = 1;
    env.rgba.assign(static_cast<size_t>(1024) * 512 * 4, 0.0f);
    float top[3], bot[3];
    for (int ch = 0; ch < 3; ++ch) {
        top[ch] = rng.uniform(0.5f, 1.5f);
        bot[ch] = rng.uniform(0.1f, 0.6f);
    }
    for (int y = 0; y < 512; ++y)
        for (int x = 0; x < 1024; ++x) {
            const float t = (y + 0.5f) / 512.0f;
            const float wob = 1.0f + 0.2f * std::sin(6.2831853f * (x + 0.5f) / 1024.0f);
            float* o = &env.rgba[(static_cast<size_t>(y) * 1024 + x) * 4];
            for (int ch = 0; ch < 3; ++ch) o[ch] = (top[ch] + (bot[ch] - top[ch]) * t) * wob;
            o[3] = 1.0f;
        }
    s.camera = default_camera();
    s.camera.position = v3p(8.0f, 1.5f, 0.0f);
    s.camera.look_at = v3p(0.0f, 0.0f, 0.0f);
    s.camera.up = v3p(0.0f, 1.0f, 0.0f);
    s.camera.fov = 35.0f;
    s.camera.width = 512;
    s.camera.height = 512;
    s.settings.sample_per_frame = 32;
    s.settings.target_samples = 32;
    s.settings.max_depth = 10;
    s.settings.sampler = 0;
    s.settings.seed = 0;
    s.output_file = "diff_views.hdr";
    s.bvh_opts = {32, 5, 4, 0};
    s.has_bvh_opts = true;
    return true;
}

}  // namespace ptamd
