// glTF 2.0 / GLB scene loading (io/gltf_loader.h).
#pragma once

#include <string>
#include <vector>

#include "host_scene.h"

namespace ptamd {

// GltfLoadOptions (io/gltf_loader.h:12-19).
struct GltfOptions {
    float transform[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    bool load_camera = true;         // use the first perspective camera of the file
    bool auto_fit_camera = true;     // otherwise fit the camera to the geometry bounds
    bool add_default_lights = true;  // two directional lights when the scene has no light at all
    bool load_environment = true;    // look for the default sky .hdr in a "models" directory above the file
    std::vector<std::string>* material_names = nullptr;  // receives the file's material names
    std::vector<std::string>* warnings = nullptr;        // images that could not be decoded
};

// GltfSceneLoader::load (io/gltf_loader.cpp:583-732): appends the file's
// textures, materials, triangles and lights to s.
bool load_gltf(const std::string& path, HostScene& s, const GltfOptions& opts, std::string* err);

}  // namespace ptamd
