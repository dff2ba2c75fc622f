"""glTF 2.0 scenes (GltfSceneLoader, Part_2/src/io/gltf_loader.cpp) through
pt_scene_load, checked on the host scene against values computed here from
the glTF spec and the reference loader's conventions: node transforms
(T*R*S, column-major matrices, parent*child), normals by the inverse
transpose, generated area-weighted normals, stride-aware accessors, index
types, materials with their KHR extensions and glTF defaults, textures with
sampler wrap modes (PNG inside a GLB buffer view, JPEG as a data URI, a
missing image -> one texel of zeros), the first perspective camera,
KHR_lights_punctual, the fitted camera and default lights, and "gltf"
objects inside JSON scenes (transform, materialOverrides)."""
import base64
import ctypes as C
import io
import json
import struct

import numpy as np
import pytest

import pathtracingtutorial_amd as pt
from pathtracingtutorial_amd import _capi

PIL = pytest.importorskip("PIL.Image")


def _view(sm):
    v = _capi.HostSceneView()
    _capi.check(pt.lib().pt_scene_host_view(sm.handle, C.byref(v)))
    return v


def _arrays(sm):
    v = _view(sm)
    P = np.array([[v.vertices[i].position.x, v.vertices[i].position.y, v.vertices[i].position.z]
                  for i in range(v.num_vertices)], np.float64)
    N = np.array([[v.vertices[i].normal.x, v.vertices[i].normal.y, v.vertices[i].normal.z]
                  for i in range(v.num_vertices)], np.float64)
    UV = np.array([[v.vertices[i].u, v.vertices[i].v] for i in range(v.num_vertices)], np.float64)
    T = np.array([list(v.vertices[i].tangent) for i in range(v.num_vertices)], np.float64)
    F = np.array([[v.faces[i].v0, v.faces[i].v1, v.faces[i].v2, v.faces[i].material_id] for i in range(v.num_faces)])
    return v, P, N, UV, T, F


def _quat_mat(q):
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y), 0],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x), 0],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y), 0], [0, 0, 0, 1]])


def _trs(t, q, s):
    T = np.eye(4)
    T[:3, 3] = t
    S = np.diag([s[0], s[1], s[2], 1.0])
    return T @ _quat_mat(q) @ S


class Builder:
    """A glTF document and its binary buffer."""

    def __init__(self):
        self.bin = bytearray()
        self.doc = {"asset": {"version": "2.0"}, "buffers": [], "bufferViews": [], "accessors": []}

    def view(self, data: bytes, stride=None):
        while len(self.bin) % 4:
            self.bin.append(0)
        bv = {"buffer": 0, "byteOffset": len(self.bin), "byteLength": len(data)}
        if stride:
            bv["byteStride"] = stride
        self.bin += data
        self.doc["bufferViews"].append(bv)
        return len(self.doc["bufferViews"]) - 1

    def accessor(self, arr: np.ndarray, ctype: int, typ: str, stride=None, offset=0, view=None):
        if view is None:
            view = self.view(arr.tobytes(), stride)
        acc = {"bufferView": view, "componentType": ctype, "count": int(arr.shape[0]), "type": typ}
        if offset:
            acc["byteOffset"] = offset
        self.doc["accessors"].append(acc)
        return len(self.doc["accessors"]) - 1

    def glb(self) -> bytes:
        doc = dict(self.doc)
        doc["buffers"] = [{"byteLength": len(self.bin)}]
        js = json.dumps(doc).encode()
        js += b" " * ((4 - len(js) % 4) % 4)
        b = bytes(self.bin) + b"\0" * ((4 - len(self.bin) % 4) % 4)
        total = 12 + 8 + len(js) + 8 + len(b)
        return (struct.pack("<III", 0x46546C67, 2, total) + struct.pack("<II", len(js), 0x4E4F534A) + js +
                struct.pack("<II", len(b), 0x004E4942) + b)

    def gltf(self) -> str:
        doc = dict(self.doc)
        doc["buffers"] = [{"byteLength": len(self.bin),
                           "uri": "data:application/octet-stream;base64," + base64.b64encode(bytes(self.bin)).decode()}]
        return json.dumps(doc)


def _png_bytes(rgba):
    b = io.BytesIO()
    PIL.fromarray(rgba, "RGBA").save(b, "PNG")
    return b.getvalue()


def _scene_doc():
    b = Builder()
    rng = np.random.default_rng(0)
    # quad: positions + normals interleaved in one strided view, uvs separate, u16 indices
    pos = np.array([[-1, 0, -1], [1, 0, -1], [1, 0, 1], [-1, 0, 1]], np.float32)
    nor = np.array([[0, 1, 0]] * 4, np.float32)
    inter = np.concatenate([pos, nor], axis=1).astype(np.float32)
    iv = b.view(inter.tobytes(), stride=24)
    a_pos = b.accessor(pos, 5126, "VEC3", view=iv)
    a_nor = b.accessor(nor, 5126, "VEC3", view=iv, offset=12)
    uv = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], np.float32)
    a_uv = b.accessor(uv, 5126, "VEC2")
    tan = np.array([[1, 0, 0, -1]] * 4, np.float32)
    a_tan = b.accessor(tan, 5126, "VEC4")
    a_idx = b.accessor(np.array([0, 2, 1, 0, 3, 2], np.uint16), 5123, "SCALAR")
    # a triangle fan without normals, u8 indices (area-weighted normals generated)
    p2 = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], np.float32)
    a_p2 = b.accessor(p2, 5126, "VEC3")
    a_i2 = b.accessor(np.array([0, 1, 2, 0, 2, 3, 0, 3, 1], np.uint8), 5121, "SCALAR")
    # non-indexed triangle
    p3 = rng.uniform(-1, 1, (3, 3)).astype(np.float32)
    a_p3 = b.accessor(p3, 5126, "VEC3")
    # textures: PNG in a buffer view, JPEG as a data URI, a missing image
    tex_rgba = rng.integers(0, 256, (4, 5, 4), dtype=np.uint8)
    png_view = b.view(_png_bytes(tex_rgba))
    jb = io.BytesIO()
    PIL.fromarray(rng.integers(0, 256, (8, 8, 3), dtype=np.uint8), "RGB").save(jb, "JPEG", quality=95)
    d = b.doc
    d["images"] = [{"bufferView": png_view, "mimeType": "image/png"},
                   {"uri": "data:image/jpeg;base64," + base64.b64encode(jb.getvalue()).decode()},
                   {"uri": "missing_file.png"}]
    d["samplers"] = [{"wrapS": 33071, "wrapT": 33648}]
    d["textures"] = [{"source": 0, "sampler": 0}, {"source": 1}, {"source": 2}]
    d["materials"] = [
        {"name": "full", "pbrMetallicRoughness": {
            "baseColorFactor": [0.2, 0.4, 0.6, 0.5], "metallicFactor": 0.3, "roughnessFactor": 0.7,
            "baseColorTexture": {"index": 0, "extensions": {"KHR_texture_transform": {
                "offset": [0.1, 0.2], "scale": [2.0, 3.0], "rotation": 0.5}}},
            "metallicRoughnessTexture": {"index": 1}},
         "normalTexture": {"index": 1, "scale": 0.8}, "emissiveTexture": {"index": 2},
         "occlusionTexture": {"index": 1, "strength": 0.4}, "emissiveFactor": [1.0, 0.5, 0.0],
         "doubleSided": True, "alphaMode": "MASK", "alphaCutoff": 0.3,
         "extensions": {"KHR_materials_emissive_strength": {"emissiveStrength": 4.0},
                        "KHR_materials_ior": {"ior": 1.33},
                        "KHR_materials_transmission": {"transmissionFactor": 0.6, "transmissionTexture": {"index": 2}},
                        "KHR_materials_clearcoat": {"clearcoatFactor": 0.9, "clearcoatRoughnessFactor": 0.1,
                                                    "clearcoatTexture": {"index": 0},
                                                    "clearcoatRoughnessTexture": {"index": 1},
                                                    "clearcoatNormalTexture": {"index": 2}},
                        "KHR_materials_sheen": {"sheenRoughnessFactor": 0.25, "sheenColorFactor": [0.1, 0.2, 0.3],
                                                "sheenColorTexture": {"index": 0},
                                                "sheenRoughnessTexture": {"index": 1}},
                        "KHR_materials_specular": {"specularFactor": 0.45, "specularColorFactor": [0.9, 0.8, 0.7]},
                        "KHR_materials_volume": {"thicknessFactor": 0.05, "attenuationDistance": 2.5,
                                                 "attenuationColor": [0.5, 0.6, 0.7]},
                        "KHR_materials_anisotropy": {"anisotropyStrength": 0.35, "anisotropyRotation": 1.2}}},
        {"name": "plain"},
        {"name": "blend", "alphaMode": "BLEND"},
    ]
    d["meshes"] = [
        {"primitives": [{"attributes": {"POSITION": a_pos, "NORMAL": a_nor, "TEXCOORD_0": a_uv, "TANGENT": a_tan},
                         "indices": a_idx, "material": 0}]},
        {"primitives": [{"attributes": {"POSITION": a_p2}, "indices": a_i2, "material": 1},
                        {"attributes": {"POSITION": a_p3}},
                        {"attributes": {"POSITION": a_p3}, "mode": 1}]},  # lines: skipped
    ]
    q = np.array([0.1, 0.3, -0.2, 0.9])
    q = (q / np.linalg.norm(q)).tolist()
    child_m = _trs([0.5, -1.0, 2.0], [0, 0, 0, 1], [1.0, 2.0, 1.0])
    cam_m = _trs([0, 1, 5], [0, 0, 0, 1], [1, 1, 1])
    d["nodes"] = [
        {"translation": [1, 2, 3], "rotation": q, "scale": [2, 2, 2], "mesh": 0, "children": [1, 3]},
        {"matrix": child_m.T.reshape(-1).tolist(), "mesh": 1},
        {"mesh": 0},  # not in the scene
        {"translation": [0, 3, 0], "extensions": {"KHR_lights_punctual": {"light": 0}}},
        {"matrix": cam_m.T.reshape(-1).tolist(), "camera": 0},
        {"rotation": [-0.70710678, 0, 0, 0.70710678], "extensions": {"KHR_lights_punctual": {"light": 1}}},
    ]
    d["cameras"] = [{"type": "perspective", "perspective": {"yfov": 0.6, "aspectRatio": 1.5, "znear": 0.1}}]
    d["extensions"] = {"KHR_lights_punctual": {"lights": [
        {"type": "spot", "color": [1, 0.5, 0.25], "intensity": 7.0, "range": 9.0,
         "spot": {"innerConeAngle": 0.2, "outerConeAngle": 0.5}},
        {"type": "directional", "intensity": 2.0}]}}
    d["scenes"] = [{"nodes": [0, 4, 5]}]
    d["scene"] = 0
    root = _trs([1, 2, 3], q, [2, 2, 2])
    return b, dict(root=root, child=root @ child_m, pos=pos, nor=nor, uv=uv, tan=tan, p2=p2, p3=p3,
                   tex_rgba=tex_rgba, q=q)


@pytest.fixture(params=["glb", "gltf"])
def loaded(request, tmp_path):
    b, exp = _scene_doc()
    path = tmp_path / ("scene." + request.param)
    if request.param == "glb":
        path.write_bytes(b.glb())
    else:
        path.write_text(b.gltf())
    sm = pt.SceneManager()
    sm.load(str(path))
    return sm, exp


def test_gltf_geometry_and_transforms(loaded):
    sm, e = loaded
    v, P, N, UV, T, F = _arrays(sm)
    # node 0: quad under T*R*S; node 1 (child, matrix): fan + triangle; line primitive skipped
    assert v.num_vertices == 4 + 4 + 3 and v.num_faces == 2 + 3 + 1
    R, C_ = e["root"], e["child"]
    exp_p = (R @ np.c_[e["pos"], np.ones(4)].T).T[:, :3]
    assert np.allclose(P[:4], exp_p, atol=1e-5)
    nm = np.linalg.inv(R).T[:3, :3]
    en = (nm @ e["nor"].T).T
    en /= np.linalg.norm(en, axis=1, keepdims=True)
    assert np.allclose(N[:4], en, atol=1e-5)
    assert np.allclose(UV[:4], e["uv"])
    et = (R[:3, :3] @ e["tan"][:, :3].T).T
    et /= np.linalg.norm(et, axis=1, keepdims=True)
    assert np.allclose(T[:4, :3], et, atol=1e-5) and np.all(T[:4, 3] == -1)
    assert F[:2].tolist() == [[0, 2, 1, 0], [0, 3, 2, 0]]
    # child: fan with generated area-weighted normals, u8 indices, material 1
    exp_p2 = (C_ @ np.c_[e["p2"], np.ones(4)].T).T[:, :3]
    assert np.allclose(P[4:8], exp_p2, atol=1e-5)
    assert F[2:5, :3].tolist() == [[4, 5, 6], [4, 6, 7], [4, 7, 5]] and np.all(F[2:5, 3] == 1)
    acc = np.zeros((4, 3))
    for f in F[2:5, :3] - 4:
        n = np.cross(exp_p2[f[1]] - exp_p2[f[0]], exp_p2[f[2]] - exp_p2[f[0]])
        acc[f] += n
    acc /= np.linalg.norm(acc, axis=1, keepdims=True)
    assert np.allclose(N[4:8], acc, atol=1e-5)
    # non-indexed triangle without a material: material 0 of the scene
    assert F[5].tolist() == [8, 9, 10, 0]
    assert np.allclose(P[8:11], (C_ @ np.c_[e["p3"], np.ones(3)].T).T[:, :3], atol=1e-5)


def test_gltf_materials(loaded):
    sm, _ = loaded
    v = _view(sm)
    assert v.num_materials == 3
    m = v.materials[0]
    f = lambda x: pytest.approx(x, rel=1e-6, abs=1e-7)
    assert (m.base_color.x, m.base_color.y, m.base_color.z) == (f(0.2), f(0.4), f(0.6))
    assert m.base_alpha == f(0.5) and m.metallic == f(0.3) and m.roughness == f(0.7)
    assert (m.base_color_tex_id, m.metallic_roughness_tex_id, m.normal_tex_id, m.emissive_tex_id,
            m.occlusion_tex_id) == (0, 1, 1, 2, 1)
    assert (m.uv_offset_x, m.uv_offset_y, m.uv_scale_x, m.uv_scale_y, m.uv_rotation) == (
        f(0.1), f(0.2), f(2.0), f(3.0), f(0.5))
    assert m.normal_scale == f(0.8) and m.occlusion_strength == f(0.4)
    assert (m.emissive.x, m.emissive.y, m.emissive.z, m.emissive_strength) == (f(1.0), f(0.5), 0.0, f(4.0))
    assert m.double_sided == 1 and m.alpha_mode == 1 and m.alpha_cutoff == f(0.3)
    assert m.ior == f(1.33) and m.spec_trans == f(0.6) and m.transmission_tex_id == 2
    assert (m.clearcoat, m.clearcoat_roughness) == (f(0.9), f(0.1))
    assert (m.clearcoat_tex_id, m.clearcoat_roughness_tex_id, m.clearcoat_normal_tex_id) == (0, 1, 2)
    assert m.sheen == f(0.25) and (m.sheen_color_r, m.sheen_color_g, m.sheen_color_b) == (f(0.1), f(0.2), f(0.3))
    assert (m.sheen_color_tex_id, m.sheen_roughness_tex_id) == (0, 1)
    assert m.specular_factor == f(0.45)
    assert (m.specular_color_r, m.specular_color_g, m.specular_color_b) == (f(0.9), f(0.8), f(0.7))
    assert (m.thickness_factor, m.attenuation_distance) == (f(0.05), f(2.5))
    assert (m.att_color_r, m.att_color_g, m.att_color_b) == (f(0.5), f(0.6), f(0.7))
    assert (m.anisotropic, m.anisotropy_rotation) == (f(0.35), f(1.2))
    p = v.materials[1]  # glTF defaults: metallic 1, roughness 1, base colour 1, opaque
    assert (p.metallic, p.roughness, p.base_color.x, p.alpha_mode, p.base_color_tex_id) == (1.0, 1.0, 1.0, 0, -1)
    # emissive [0,0,0] by default (strength keeps the material default): not an emitter
    assert (p.emissive.x, p.emissive.y, p.emissive.z) == (0.0, 0.0, 0.0) and p.double_sided == 0
    assert v.materials[2].alpha_mode == 2


def test_gltf_textures(loaded):
    sm, e = loaded
    assert sm.info().num_textures == 3
    d = _capi.TextureDesc()
    _capi.check(pt.lib().pt_scene_texture(sm.handle, 0, C.byref(d)))
    t0 = np.ctypeslib.as_array(d.rgba, shape=(d.height * d.width * 4,)).reshape(d.height, d.width, 4)
    assert (d.width, d.height, d.wrap_s, d.wrap_t, d.linear) == (5, 4, 33071, 33648, 0)
    assert np.array_equal(t0, e["tex_rgba"].astype(np.float32) / np.float32(255.0))  # row 0 = image top
    _capi.check(pt.lib().pt_scene_texture(sm.handle, 1, C.byref(d)))
    assert (d.width, d.height, d.wrap_s, d.wrap_t) == (8, 8, 10497, 10497)
    _capi.check(pt.lib().pt_scene_texture(sm.handle, 2, C.byref(d)))  # missing image: one zero texel
    t2 = np.ctypeslib.as_array(d.rgba, shape=(4,))
    assert (d.width, d.height) == (1, 1) and np.all(t2 == 0)


def test_gltf_camera_and_lights(loaded):
    sm, e = loaded
    cam = sm.camera()
    assert (cam.position.x, cam.position.y, cam.position.z) == (0.0, 1.0, 5.0)
    assert (cam.look_at.x, cam.look_at.y, cam.look_at.z) == (0.0, 1.0, 4.0)
    assert cam.fov == pytest.approx(0.6 * 180 / np.pi, rel=1e-6)
    assert cam.width == int(1.5 * cam.height)
    v = _view(sm)
    assert v.num_lights == 2
    s = v.lights[0]
    R = e["root"]
    exp_pos = R @ np.array([0, 3, 0, 1.0])  # light node is a child of node 0
    assert s.type == 2 and (s.position.x, s.position.y, s.position.z) == tuple(pytest.approx(x, abs=1e-5) for x in exp_pos[:3])
    d = R[:3, :3] @ np.array([0, 0, -1.0])
    d /= np.linalg.norm(d)
    assert np.allclose([s.direction.x, s.direction.y, s.direction.z], d, atol=1e-5)
    assert (s.intensity, s.range) == (7.0, 9.0)
    assert (s.color.x, s.color.y, s.color.z) == (1.0, 0.5, 0.25)
    assert s.inner_cone == pytest.approx(np.cos(0.2)) and s.outer_cone == pytest.approx(np.cos(0.5))
    dl = v.lights[1]  # rotated -90 deg about x: -z turns into -y
    assert dl.type == 1 and np.allclose([dl.direction.x, dl.direction.y, dl.direction.z], [0, -1, 0], atol=1e-6)
    assert dl.range == pytest.approx(1e30) and dl.outer_cone == pytest.approx(np.cos(np.pi / 4))


def test_gltf_fallbacks(tmp_path):
    """No camera: fitted to the bounds (gltf_loader.cpp:650-679); nothing emits: two default lights."""
    b = Builder()
    pos = np.array([[0, 0, 0], [2, 0, 0], [0, 4, 0]], np.float32)
    a = b.accessor(pos, 5126, "VEC3")
    b.doc.update(meshes=[{"primitives": [{"attributes": {"POSITION": a}}]}], nodes=[{"mesh": 0}],
                 scenes=[{"nodes": [0]}])
    p = tmp_path / "t.glb"
    p.write_bytes(b.glb())
    sm = pt.SceneManager()
    sm.load(str(p))
    v = _view(sm)
    assert v.num_materials == 1 and v.num_lights == 2
    assert all(v.lights[i].type == 1 for i in range(2))
    assert (v.lights[0].intensity, v.lights[1].intensity) == (3.0, 1.0)
    cam = sm.camera()
    center = np.array([1.0, 2.0, 0.0])
    radius = np.linalg.norm([2, 4, 0]) * 0.5
    dist = radius / np.tan(np.radians(cam.fov) * 0.5) * 1.2
    assert np.allclose([cam.position.x, cam.position.y, cam.position.z],
                       center + [0, radius * 0.3, dist], rtol=1e-5)
    assert np.allclose([cam.look_at.x, cam.look_at.y, cam.look_at.z], center)
    # an emissive material: no default lights
    b.doc["materials"] = [{"emissiveFactor": [1, 1, 1]}]
    b.doc["meshes"][0]["primitives"][0]["material"] = 0
    p.write_bytes(b.glb())
    sm2 = pt.SceneManager()
    sm2.load(str(p))
    assert _view(sm2).num_lights == 0


def test_json_scene_with_gltf_object(tmp_path):
    b, e = _scene_doc()
    (tmp_path / "obj.glb").write_bytes(b.glb())
    scene = {"camera": {"position": [0, 0, 5], "lookAt": [0, 0, 0], "up": [0, 1, 0], "fov": 40,
                        "resolution": [64, 48]},
             "materials": {"white": {"baseColor": [0.8, 0.8, 0.8]}},
             "objects": [{"type": "quad", "vertices": [[-1, -1, 0], [1, -1, 0], [1, 1, 0], [-1, 1, 0]],
                          "material": "white"},
                         {"type": "gltf", "file": "obj.glb",
                          "transform": {"translate": [0, 0, -3], "rotate": {"axis": [0, 1, 0], "angle": 90},
                                        "scale": 0.5},
                          "materialOverrides": {"plain": {"baseColor": [0.1, 0.9, 0.1], "roughness": 0.2}}}]}
    (tmp_path / "s.json").write_text(json.dumps(scene))
    sm = pt.SceneManager()
    sm.load(str(tmp_path / "s.json"))
    v, P, N, UV, T, F = _arrays(sm)
    assert v.num_materials == 1 + 3 and v.num_lights == 2  # JSON material first, then the file's
    assert F[2:4, 3].tolist() == [1, 1]  # the file's material 0 -> scene material 1
    ov = v.materials[2]
    assert (ov.base_color.x, ov.base_color.y, ov.base_color.z) == (pytest.approx(0.1), pytest.approx(0.9), pytest.approx(0.1))
    assert ov.roughness == pytest.approx(0.2) and ov.metallic == 1.0
    # object transform T * R(y, 90 deg) * S(0.5) applied before the file's node transforms
    X = np.eye(4)
    X[:3, 3] = [0, 0, -3]
    Ry = np.array([[0, 0, 1, 0], [0, 1, 0, 0], [-1, 0, 0, 0], [0, 0, 0, 1.0]])
    X = X @ Ry @ np.diag([0.5, 0.5, 0.5, 1.0])
    exp = (X @ e["root"] @ np.c_[e["pos"], np.ones(4)].T).T[:, :3]
    assert np.allclose(P[4:8], exp, atol=1e-5)
    cam = sm.camera()
    assert (cam.width, cam.height) == (64, 48) and cam.position.z == 5.0  # the file's camera is not used


def test_gltf_errors(tmp_path):
    p = tmp_path / "bad.gltf"
    p.write_text("{ not json")
    with pytest.raises(pt.PtError):
        pt.SceneManager().load(str(p))
    b = Builder()
    a = b.accessor(np.zeros((3, 3), np.float32), 5126, "VEC3")
    b.doc["accessors"][a]["count"] = 100  # reaches past the buffer
    b.doc.update(meshes=[{"primitives": [{"attributes": {"POSITION": a}}]}], nodes=[{"mesh": 0}],
                 scenes=[{"nodes": [0]}])
    p2 = tmp_path / "short.glb"
    p2.write_bytes(b.glb())
    with pytest.raises(pt.PtError):
        pt.SceneManager().load(str(p2))
