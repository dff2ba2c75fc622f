"""A glTF export of the Cornell box JSON scene renders the same image.

The test writes the JSON scene's triangles (one primitive per run of faces
with one material, so the face order and therefore the BVH are kept),
vertex attributes and materials (base colour, metallic, roughness,
emissive + KHR_materials_emissive_strength) as a GLB, loads it with
pt_scene_load, gives it the JSON scene's camera and render settings and
renders one batch with both integrators.  The loader passes normals and
tangents through the normal matrix and renormalizes them (as the
reference's glTF loader does), which may move a unit vector by one ulp, so
the images are the same paths up to float rounding: nearly all pixels
bit-identical, every channel within 1e-4 relative but a few where a
rounding difference changed a path's branch."""
import ctypes as C
import json
import os
import struct

import numpy as np
import pytest

import pathtracingtutorial_amd as pt
from pathtracingtutorial_amd import _capi

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
CORNELL = os.path.join(HERE, "..", "scenes", "cornell_box.json")


def _view(sm):
    v = _capi.HostSceneView()
    _capi.check(pt.lib().pt_scene_host_view(sm.handle, C.byref(v)))
    return v


def _export_glb(sm, path):
    v = _view(sm)
    nv, nf = v.num_vertices, v.num_faces
    verts = np.ctypeslib.as_array(C.cast(v.vertices, C.POINTER(C.c_float)), shape=(nv * 16,)).reshape(nv, 16)
    faces = np.ctypeslib.as_array(C.cast(v.faces, C.POINTER(C.c_uint32)), shape=(nf * 4,)).reshape(nf, 4)
    pos, nor, uv, tan = verts[:, 0:3].copy(), verts[:, 4:7].copy(), verts[:, 8:10].copy(), verts[:, 12:16].copy()
    bin_ = bytearray()
    views, accs = [], []

    def add(arr, ctype, typ):
        while len(bin_) % 4:
            bin_.append(0)
        views.append({"buffer": 0, "byteOffset": len(bin_), "byteLength": arr.nbytes})
        bin_.extend(arr.tobytes())
        accs.append({"bufferView": len(views) - 1, "componentType": ctype, "count": int(arr.shape[0]), "type": typ})
        return len(accs) - 1

    a = {"POSITION": add(pos.astype(np.float32), 5126, "VEC3"), "NORMAL": add(nor.astype(np.float32), 5126, "VEC3"),
         "TEXCOORD_0": add(uv.astype(np.float32), 5126, "VEC2"), "TANGENT": add(tan.astype(np.float32), 5126, "VEC4")}
    prims = []
    start = 0
    for i in range(1, nf + 1):
        if i == nf or faces[i, 3] != faces[start, 3]:
            idx = faces[start:i, :3].reshape(-1).astype(np.uint32)
            prims.append({"attributes": a, "indices": add(idx, 5125, "SCALAR"), "material": int(faces[start, 3])})
            start = i
    mats = []
    for i in range(v.num_materials):
        m = v.materials[i]
        g = {"pbrMetallicRoughness": {"baseColorFactor": [m.base_color.x, m.base_color.y, m.base_color.z, m.base_alpha],
                                      "metallicFactor": m.metallic, "roughnessFactor": m.roughness}}
        if m.emissive.x > 0 or m.emissive.y > 0 or m.emissive.z > 0:
            g["emissiveFactor"] = [m.emissive.x, m.emissive.y, m.emissive.z]
            g["extensions"] = {"KHR_materials_emissive_strength": {"emissiveStrength": m.emissive_strength}}
        mats.append(g)
    doc = {"asset": {"version": "2.0"}, "buffers": [{"byteLength": len(bin_)}], "bufferViews": views,
           "accessors": accs, "materials": mats, "meshes": [{"primitives": prims}], "nodes": [{"mesh": 0}],
           "scenes": [{"nodes": [0]}], "scene": 0}
    js = json.dumps(doc).encode()
    js += b" " * ((4 - len(js) % 4) % 4)
    while len(bin_) % 4:
        bin_.append(0)
    total = 12 + 8 + len(js) + 8 + len(bin_)
    with open(path, "wb") as f:
        f.write(struct.pack("<III", 0x46546C67, 2, total) + struct.pack("<II", len(js), 0x4E4F534A) + js +
                struct.pack("<II", len(bin_), 0x004E4942) + bytes(bin_))
    return len(prims)


def _render(sm, integ):
    r = pt.Renderer(sm, integ)
    r.renderBatch()
    return r.downloadCurrentImage(1).copy()


def test_gltf_export_of_cornell_renders_identically(tmp_path):
    js = pt.SceneManager()
    js.load(CORNELL)
    cam = js.camera()
    cam.width, cam.height = 96, 96
    js.updateCamera(cam)
    js.setSamplePerFrame(4)
    nprims = _export_glb(js, str(tmp_path / "cornell.glb"))
    assert nprims >= 5
    gl = pt.SceneManager()
    gl.load(str(tmp_path / "cornell.glb"))
    info_j = js.info()
    assert gl.info().num_faces == info_j.num_faces and gl.info().num_lights == 0  # the light quad emits
    gl.updateCamera(cam)
    gl.setSamplePerFrame(4)
    gl.setMaxDepth(info_j.max_depth)
    gl.setSampler(info_j.sampler)
    for sm in (js, gl):
        sm.buildAccel()
        sm.uploadToDevice()
    assert js.numBvhNodes() == gl.numBvhNodes()
    for integ in (pt.PathIntegrator, pt.WavefrontIntegrator):
        a = _render(js, integ())
        js.uploadToDevice()
        b = _render(gl, integ())
        gl.uploadToDevice()
        assert np.isfinite(a).all() and a[..., :3].mean() > 0.01
        same = np.all(a == b, axis=-1).mean()
        close = np.abs(a - b) <= 1e-4 * (1.0 + np.abs(a))
        assert same > 0.5 and close.mean() >= 0.99, (integ.__name__, same, close.mean())


def _textured_scene(path):
    """A box lit by an emissive-textured ceiling quad and a point light, with
    quads using every texture slot: base colour (uv transform, clamp/mirror
    wrap), metallic-roughness, normal maps with and without mesh tangents,
    alpha mask, clearcoat, sheen and transmission textures."""
    import io

    from PIL import Image
    from test_gltf import Builder

    rng = np.random.default_rng(21)
    b = Builder()
    d = b.doc
    y, x = np.mgrid[0:16, 0:16]

    def png(rgba):
        bio = io.BytesIO()
        Image.fromarray(rgba.astype(np.uint8), "RGBA").save(bio, "PNG")
        return {"bufferView": b.view(bio.getvalue()), "mimeType": "image/png"}

    checker = ((x // 4 + y // 4) % 2).astype(np.float64)
    base = np.stack([80 + 150 * checker, 60 + 100 * (1 - checker), 40 + 30 * rng.random((16, 16)),
                     np.full((16, 16), 255)], -1)
    mr = np.stack([np.zeros((16, 16)), 40 + 200 * (x / 15), 255 * (y > 7), np.full((16, 16), 255)], -1)
    bump = np.stack([np.sin(x * 0.8), np.cos(y * 0.6), np.full((16, 16), 2.0)], -1)
    bump /= np.linalg.norm(bump, axis=-1, keepdims=True)
    normal = np.concatenate([(bump * 0.5 + 0.5) * 255, np.full((16, 16, 1), 255)], -1)
    emis = np.stack([255 * (x / 15), 200 * np.ones((16, 16)), 255 * (y / 15), np.full((16, 16), 255)], -1)
    alpha = np.stack([np.full((16, 16), 200), np.full((16, 16), 120), np.full((16, 16), 90), 255 * checker], -1)
    misc = rng.integers(30, 256, (16, 16, 4))
    d["images"] = [png(a) for a in (base, mr, normal, emis, alpha, misc)]
    d["samplers"] = [{"wrapS": 33071, "wrapT": 33648}, {}]
    d["textures"] = [{"source": 0, "sampler": 0}, {"source": 1, "sampler": 1}, {"source": 2}, {"source": 3},
                     {"source": 4}, {"source": 5}]
    d["materials"] = [
        {"name": "wall", "pbrMetallicRoughness": {"baseColorFactor": [0.75, 0.75, 0.75, 1], "metallicFactor": 0.0}},
        {"name": "textured", "pbrMetallicRoughness": {
            "baseColorTexture": {"index": 0, "extensions": {"KHR_texture_transform": {
                "offset": [0.2, -0.1], "scale": [1.7, 0.8], "rotation": 0.4}}},
            "metallicRoughnessTexture": {"index": 1}, "metallicFactor": 1.0, "roughnessFactor": 1.0}},
        {"name": "bumpy", "pbrMetallicRoughness": {"baseColorFactor": [0.6, 0.7, 0.8, 1], "metallicFactor": 0.0,
                                                   "roughnessFactor": 0.4},
         "normalTexture": {"index": 2, "scale": 0.9}},
        {"name": "light", "emissiveFactor": [1, 1, 1], "emissiveTexture": {"index": 3},
         "extensions": {"KHR_materials_emissive_strength": {"emissiveStrength": 6.0}},
         "pbrMetallicRoughness": {"metallicFactor": 0.0}},
        {"name": "leaf", "alphaMode": "MASK", "alphaCutoff": 0.5, "doubleSided": True,
         "pbrMetallicRoughness": {"baseColorTexture": {"index": 4}, "metallicFactor": 0.0, "roughnessFactor": 0.8}},
        {"name": "coated", "pbrMetallicRoughness": {"baseColorFactor": [0.3, 0.5, 0.2, 1], "metallicFactor": 0.2,
                                                    "roughnessFactor": 0.5},
         "extensions": {"KHR_materials_clearcoat": {"clearcoatFactor": 1.0, "clearcoatRoughnessFactor": 0.6,
                                                    "clearcoatTexture": {"index": 5},
                                                    "clearcoatRoughnessTexture": {"index": 5}},
                        "KHR_materials_sheen": {"sheenRoughnessFactor": 0.7, "sheenColorFactor": [1, 0.8, 0.6],
                                                "sheenColorTexture": {"index": 5}, "sheenRoughnessTexture": {"index": 5}},
                        "KHR_materials_transmission": {"transmissionFactor": 0.5,
                                                       "transmissionTexture": {"index": 5}}}},
    ]
    prims = []

    def quad(a, bb, c, dd, mat, tangent=True):
        P = np.array([a, bb, c, dd], np.float32)
        n = np.cross(P[1] - P[0], P[2] - P[0])
        n /= np.linalg.norm(n)
        N = np.repeat(n[None].astype(np.float32), 4, 0)
        UV = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], np.float32)
        attrs = {"POSITION": b.accessor(P, 5126, "VEC3"), "NORMAL": b.accessor(N, 5126, "VEC3"),
                 "TEXCOORD_0": b.accessor(UV, 5126, "VEC2")}
        if tangent:
            t = (P[1] - P[0]) / np.linalg.norm(P[1] - P[0])
            attrs["TANGENT"] = b.accessor(np.repeat(np.r_[t, 1.0][None].astype(np.float32), 4, 0), 5126, "VEC4")
        prims.append({"attributes": attrs, "indices": b.accessor(np.array([0, 1, 2, 0, 2, 3], np.uint16), 5123, "SCALAR"),
                      "material": mat})

    s = 1.0
    quad([-s, -s, -s], [s, -s, -s], [s, -s, s], [-s, -s, s], 0)      # floor (normal up after winding)
    quad([-s, s, s], [s, s, s], [s, s, -s], [-s, s, -s], 0)          # ceiling
    quad([-s, -s, -s], [-s, s, -s], [s, s, -s], [s, -s, -s], 1)      # back wall: textured
    quad([-s, -s, s], [-s, s, s], [-s, s, -s], [-s, -s, -s], 2)      # left: normal map + tangents
    quad([s, -s, -s], [s, s, -s], [s, s, s], [s, -s, s], 2, False)   # right: normal map, no tangents
    quad([-0.3, 0.99, 0.3], [0.3, 0.99, 0.3], [0.3, 0.99, -0.3], [-0.3, 0.99, -0.3], 3)  # light
    quad([-0.7, -0.9, 0.2], [-0.1, -0.9, 0.2], [-0.1, -0.1, -0.2], [-0.7, -0.1, -0.2], 4)  # leaf
    quad([0.1, -0.9, -0.4], [0.7, -0.9, -0.4], [0.7, 0.0, -0.1], [0.1, 0.0, -0.1], 5)    # coated
    d["meshes"] = [{"primitives": prims}]
    d["nodes"] = [{"mesh": 0}, {"translation": [0, 0, 3.2], "camera": 0},
                  {"translation": [0.4, 0.5, 0.5], "extensions": {"KHR_lights_punctual": {"light": 0}}}]
    d["cameras"] = [{"type": "perspective", "perspective": {"yfov": 0.75, "znear": 0.1}}]
    d["extensions"] = {"KHR_lights_punctual": {"lights": [{"type": "point", "color": [1, 0.9, 0.8],
                                                           "intensity": 1.5}]}}
    d["scenes"] = [{"nodes": [0, 1, 2]}]
    open(path, "wb").write(b.glb())


def test_textured_gltf_matches_oracle(tmp_path):
    import oracle
    from test_gpu_parity import _paired_batches, _zscore_check
    path = str(tmp_path / "textured.glb")
    _textured_scene(path)
    sm = pt.SceneManager()
    sm.load(path)
    cam = sm.camera()
    cam.width, cam.height = 64, 64
    sm.updateCamera(cam)
    sm.setSamplePerFrame(4)
    sm.setMaxDepth(6)
    sm.buildAccel()
    sm.uploadToDevice()
    info = sm.info()
    assert info.num_textures == 6 and info.num_lights == 1 and info.num_materials == 6
    ora = oracle.OracleScene.from_view(_view(sm))
    for i in range(info.num_textures):
        d = _capi.TextureDesc()
        _capi.check(pt.lib().pt_scene_texture(sm.handle, i, C.byref(d)))
        rgba = np.ctypeslib.as_array(d.rgba, shape=(d.height * d.width * 4,)).reshape(d.height, d.width, 4).copy()
        ora.add_texture(rgba, d.wrap_s, d.wrap_t, d.linear)
    gpu, cpu, _ = _paired_batches(sm, ora, 4, 16)
    assert np.isfinite(gpu).all() and gpu.mean() > 1e-3
    _zscore_check(gpu, cpu)
