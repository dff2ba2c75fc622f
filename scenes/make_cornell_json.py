"""Writes scenes/cornell_box.json: the 36-triangle Cornell workload expressed in
the reference's scene schema (quads + named materials + camera + render).

    python scenes/make_cornell_json.py
"""
import ctypes as C
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import pathtracingtutorial_amd as pt  # noqa: E402
from pathtracingtutorial_amd import _capi  # noqa: E402

NAMES = ["m0_white", "m1_red", "m2_green", "m3_light", "m4_mirror"]  # sorted order == index order


def f(x):
    return float(str(np.float32(x)))


def main():
    sm = pt.SceneManager()
    sm.generate(0)
    v = _capi.HostSceneView()
    _capi.check(pt.lib().pt_scene_host_view(sm.handle, C.byref(v)))
    va = np.frombuffer(C.string_at(v.vertices, 64 * v.num_vertices), np.float32).reshape(-1, 16)
    fa = np.frombuffer(C.string_at(v.faces, 16 * v.num_faces), np.uint32).reshape(-1, 4)
    mats = {}
    for k in range(v.num_materials):
        m = _capi.Material.from_buffer_copy(C.string_at(C.addressof(v.materials.contents) + 256 * k, 256))
        e = {"baseColor": [f(m.base_color.x), f(m.base_color.y), f(m.base_color.z)], "roughness": f(m.roughness),
             "metallic": f(m.metallic)}
        if m.emissive.x or m.emissive.y or m.emissive.z:
            e["emissive"] = [f(m.emissive.x), f(m.emissive.y), f(m.emissive.z)]
            e["emissiveStrength"] = f(m.emissive_strength)
        mats[NAMES[k]] = e
    objs = []
    for q in range(0, v.num_vertices, 4):
        tri = fa[q // 2]
        objs.append({"type": "quad", "vertices": [[f(c) for c in va[q + i, 0:3]] for i in range(4)],
                     "material": NAMES[int(tri[3])]})
    c = v.camera
    scene = {
        "camera": {"position": c.position.tolist(), "lookAt": c.look_at.tolist(), "up": c.up.tolist(),
                   "fov": c.fov, "resolution": [c.width, c.height]},
        "render": {"spp": 16, "samplePerFrame": 16, "maxDepth": 8, "sampler": "pcg", "output": "cornell_box.hdr"},
        "materials": mats,
        "objects": objs,
    }
    with open(os.path.join(HERE, "cornell_box.json"), "w") as fh:
        json.dump(scene, fh, indent=1)


if __name__ == "__main__":
    main()
