"""Host side of the boundary: scene description loading (JSON schema, OBJ,
.hdr), procedural workloads and the C-ABI symbols.  CPU only; nothing here
launches a kernel."""
import ctypes as C
import hashlib
import json
import os
import re

import numpy as np
import pytest

import pathtracingtutorial_amd as pt
from conftest import ROOT, golden_npz
from pathtracingtutorial_amd import _capi

REF_TREE = "/root/reference/Part_2"
needs_ref = pytest.mark.skipif(not os.path.isdir(REF_TREE), reason="reference assets not present on this machine")


def sha(b):
    return hashlib.sha256(b).hexdigest()


def load_obj_arrays(name, g):
    """Vertex/face bytes of a reference mesh as the reference loader produced them (positions only)."""
    if name != "sphere2.obj":
        raise KeyError(name)
    m = golden_npz("trace_sphere2.npz")
    pos, faces = m["positions"], m["faces"]
    verts = np.zeros((pos.shape[0], 16), np.float32)
    verts[:, 0:3] = pos
    f = np.zeros((faces.shape[0], 4), np.uint32)
    f[:, :3] = faces
    return verts.tobytes(), f.tobytes(), pos.shape[0], faces.shape[0]


def host_view(sm):
    v = _capi.HostSceneView()
    _capi.check(pt.lib().pt_scene_host_view(sm.handle, C.byref(v)))
    return v


def view_arrays(v):
    va = np.frombuffer(C.string_at(v.vertices, 64 * v.num_vertices), np.float32).reshape(-1, 16)
    fa = np.frombuffer(C.string_at(v.faces, 16 * v.num_faces), np.uint32).reshape(-1, 4)
    return va, fa


# ------------------------------------------------------------------ C ABI
def test_library_exports_every_declared_symbol():
    header = open(os.path.join(ROOT, "include", "pt_amd.h")).read()
    declared = set(re.findall(r"\b(pt_[a-z0-9_]+)\s*\(", header))
    declared -= {"pt_status"}
    lib = pt.lib()
    missing = [s for s in sorted(declared) if not hasattr(lib, s)]
    assert not missing, missing
    assert declared <= set(_capi.SIGNATURES), sorted(declared - set(_capi.SIGNATURES))


def test_struct_sizes_agree_with_bindings():
    sizes = (C.c_uint32 * 8)()
    _capi.check(pt.lib().pt_struct_sizes(sizes, 8))
    assert list(sizes) == [C.sizeof(_capi.Vertex), C.sizeof(_capi.Face), C.sizeof(_capi.Material),
                           C.sizeof(_capi.PunctualLight), C.sizeof(_capi.Camera), C.sizeof(_capi.CameraFrame),
                           C.sizeof(_capi.HostSceneDesc), C.sizeof(_capi.Counters)]
    assert list(sizes)[:5] == [64, 16, 256, 112, 80]


def test_default_material_matches_reference(golden, tmp_path):
    # materials/material.h defaults, byte for byte (the reference's Material{})
    sm = pt.SceneManager()
    sm.generate(0)
    ref = bytes.fromhex(golden["default_material_hex"])
    d = _capi.Material.from_buffer_copy(ref)
    assert abs(d.base_color.x - 0.8) < 1e-7 and d.roughness == np.float32(0.5) and d.ior == np.float32(1.5)
    # a JSON material with no overrides is the default material
    path = str(tmp_path / "default_mat_scene.json")
    with open(path, "w") as f:
        json.dump({"materials": {"m": {}}, "objects": [
            {"type": "quad", "vertices": [[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], "material": "m"}]}, f)
    sm.load(path)
    v = host_view(sm)
    got = C.string_at(v.materials, 256)
    # compare every field except padding
    a, b = _capi.Material.from_buffer_copy(got), d
    for name, _t in _capi.Material._fields_:
        if name.startswith("_"):
            continue
        va, vb = getattr(a, name), getattr(b, name)
        if isinstance(va, _capi.Vec3p):
            assert va.tolist() == vb.tolist(), name
        else:
            assert va == vb, name


def test_errors_are_status_codes_not_exits():
    sm = pt.SceneManager()
    with pytest.raises(pt.PtError) as e:
        sm.load("/nonexistent/scene.json")
    assert e.value.status == 2  # PT_ERR_IO
    with pytest.raises(pt.PtError) as e:
        sm.load(os.path.join(ROOT, "pytest.ini"))
    assert e.value.status == 6  # unknown format
    bad = os.path.join("/tmp", "pt_bad_scene.json")
    open(bad, "w").write('{"camera": {"fov": }')
    with pytest.raises(pt.PtError) as e:
        sm.load(bad)
    assert e.value.status == 3  # PT_ERR_PARSE
    open(bad, "w").write(json.dumps({"objects": [
        {"type": "quad", "vertices": [[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], "material": 3}]}))
    with pytest.raises(pt.PtError) as e:  # material index out of range is rejected, not read out of bounds
        sm.load(bad)
    assert e.value.status == 1


def test_scene_create_validates_indices():
    verts = (_capi.Vertex * 3)()
    faces = (_capi.Face * 1)(_capi.Face(0, 1, 5, 0))
    d = _capi.HostSceneDesc()
    d.vertices, d.num_vertices = verts, 3
    d.faces, d.num_faces = faces, 1
    sm = pt.SceneManager()
    with pytest.raises(pt.PtError):
        sm.create(d)
    faces[0] = _capi.Face(0, 1, 2, 0)
    verts[1].position = _capi.Vec3p.of((float("nan"), 0, 0))
    with pytest.raises(pt.PtError):
        sm.create(d)


# --------------------------------------------------------- JSON scene schema
@needs_ref
@pytest.mark.parametrize("scene", ["nee_area_light.json", "env_diffuse.json"])
def test_reference_scene_files_load(scene):
    sm = pt.SceneManager()
    sm.load(os.path.join(REF_TREE, "scenes", scene))
    i = sm.info()
    assert i.num_faces == 12 if scene == "nee_area_light.json" else 10
    assert (i.width, i.height) == (800, 800)
    assert i.target_samples == 512 and i.sample_per_frame == 4 and i.max_depth == 8
    v = host_view(sm)
    mats = [_capi.Material.from_buffer_copy(C.string_at(C.addressof(v.materials.contents) + 256 * k, 256))
            for k in range(v.num_materials)]
    # materials are indexed in sorted-name order (nlohmann::json object iteration)
    if scene == "nee_area_light.json":
        assert i.num_materials == 4
        assert [round(m.base_color.z, 3) for m in mats] == [0.9, 1.0, 0.25, 0.9]  # blue, light, red, white
        assert mats[1].emissive_strength == 40.0
    else:
        assert (i.env_width, i.env_height) == (800, 800)


def test_json_scene_schema_and_defaults(tmp_path):
    scene = {
        "camera": {"position": [1, 2, 3], "lookAt": [0, 0, -1], "up": [0, 1, 0], "fov": 60, "aperture": 0.1,
                   "focalDistance": 3.5, "resolution": [320, 200]},
        "render": {"spp": 64, "samplePerFrame": 8, "maxDepth": 5, "sampler": "pcg", "output": "x.hdr", "seed": 9},
        "materials": {"zeta": {"baseColor": [0.1, 0.2, 0.3], "roughness": 0.25, "metallic": 0.5, "doubleSided": True,
                               "alphaMode": 2},
                      "alpha": {"emissive": [1, 2, 3], "emissiveStrength": 4}},
        "lights": [{"type": "spot", "position": [0, 5, 0], "direction": [0, -2, 0], "innerCone": 0.2,
                    "outerCone": 0.5, "intensity": 3, "color": [1, 0.5, 0.25], "range": 10}],
        "objects": [{"type": "quad", "vertices": [[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], "material": "zeta"},
                    {"type": "quad", "vertices": [[0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], "material": 0}],
    }
    p = tmp_path / "s.json"
    p.write_text(json.dumps(scene))
    sm = pt.SceneManager()
    sm.load(str(p))
    i = sm.info()
    assert (i.width, i.height, i.target_samples, i.sample_per_frame, i.max_depth, i.sampler, i.seed) == \
        (320, 200, 64, 8, 5, 0, 9)
    v = host_view(sm)
    cam = v.camera
    assert cam.fov == 60 and abs(cam.aperture - 0.1) < 1e-7 and abs(cam.focal_distance - 3.5) < 1e-7
    va, fa = view_arrays(v)
    assert fa[:, 3].tolist() == [1, 1, 0, 0]  # "zeta" sorts after "alpha"
    m1 = _capi.Material.from_buffer_copy(C.string_at(C.addressof(v.materials.contents) + 256, 256))
    assert m1.double_sided == 1 and m1.alpha_mode == 2 and m1.roughness == np.float32(0.25)
    light = v.lights[0]
    assert light.type == 2
    assert abs(light.inner_cone - np.cos(0.2)) < 1e-6 and abs(light.outer_cone - np.cos(0.5)) < 1e-6
    assert light.direction.tolist() == [0.0, -1.0, 0.0]
    # quad geometry: two triangles (a,b,c), (a,c,d), flat normal, uv square (geometry/shapes.cpp:5-22)
    assert fa[0, :3].tolist() == [0, 1, 2] and fa[1, :3].tolist() == [0, 2, 3]
    assert va[:4, 4:7].tolist() == [[0, 0, 1]] * 4
    assert va[:4, 8:10].tolist() == [[0, 0], [1, 0], [1, 1], [0, 1]]


def test_quad_matches_reference_builder(golden, tmp_path):
    q = golden["quad"]
    scene = {"materials": {}, "objects": [{"type": "quad", "vertices": q["in"], "material": 7}]}
    scene["materials"] = {f"m{k}": {} for k in range(8)}
    p = tmp_path / "q.json"
    p.write_text(json.dumps(scene))
    sm = pt.SceneManager()
    sm.load(str(p))
    v = host_view(sm)
    got = np.frombuffer(C.string_at(v.vertices, 4 * 64), np.float32).reshape(4, 16)
    ref = np.frombuffer(bytes.fromhex(q["vertices_hex"]), np.float32).reshape(4, 16)
    cols = [0, 1, 2, 4, 5, 6, 8, 9, 12, 13, 14, 15]  # position, normal, uv, tangent (no padding)
    assert np.array_equal(got[:, cols], ref[:, cols])
    assert C.string_at(v.faces, 32) == bytes.fromhex(q["faces_hex"])


def test_auto_fit_camera(tmp_path):
    # scene_loader.cpp:212-238: no camera block -> look at the bounds centre from +z
    p = tmp_path / "a.json"
    p.write_text(json.dumps({"objects": [{"type": "quad", "vertices": [[-1, -1, 0], [1, -1, 0], [1, 1, 0],
                                                                       [-1, 1, 0]], "material": 0}]}))
    sm = pt.SceneManager()
    sm.load(str(p))
    c = sm.camera()
    r = np.sqrt(8) / 2
    dist = r / np.tan(np.radians(90) / 2) * 1.2
    assert np.allclose(c.position.tolist(), [0, r * 0.3, dist], rtol=1e-6)
    assert c.look_at.tolist() == [0, 0, 0]


def test_camera_frame_reference_known_answers():
    # tests/test_math.cpp:51-85
    import oracle
    cam = _capi.Camera()
    cam.position, cam.look_at, cam.up = _capi.Vec3p.of((0, 0, 4)), _capi.Vec3p.of((0, 0, 0)), _capi.Vec3p.of((0, 1, 0))
    cam.fov, cam.width, cam.height = 90.0, 800, 800
    for f in (oracle.camera_frame(cam),):
        assert abs(f.forward[2] + 1) < 1e-5 and abs(f.right[0] - 1) < 1e-5 and abs(f.up[1] - 1) < 1e-5
        assert abs(f.tan_half_fov - 1) < 1e-4 and abs(f.aspect - 1) < 1e-5
    sm = pt.SceneManager()
    sm.generate(0)
    sm.updateCamera(cam)
    f = sm.cameraFrame()
    assert abs(f.forward[2] + 1) < 1e-5 and abs(f.right[0] - 1) < 1e-5 and abs(f.up[1] - 1) < 1e-5
    assert abs(f.tan_half_fov - 1) < 1e-4 and abs(f.aspect - 1) < 1e-5
    cam.position, cam.look_at = _capi.Vec3p.of((0, 0, 0)), _capi.Vec3p.of((1, 0, 0))
    cam.fov, cam.width, cam.height = 60.0, 1920, 1080
    sm.updateCamera(cam)
    f = sm.cameraFrame()
    assert abs(f.forward[0] - 1) < 1e-5 and abs(f.forward[1]) < 1e-5 and abs(f.forward[2]) < 1e-5
    assert abs(f.tan_half_fov - np.tan(np.radians(30))) < 1e-4 and abs(f.aspect - 1920 / 1080) < 1e-4
    o = oracle.camera_frame(cam)
    assert list(o.forward) == list(f.forward) and list(o.right) == list(f.right) and list(o.up) == list(f.up)
    assert o.tan_half_fov == f.tan_half_fov and o.aspect == f.aspect


# ---------------------------------------------------------------- OBJ + HDR
@needs_ref
@pytest.mark.parametrize("name", ["sphere2.obj", "Stanford Bunny.obj", "teapot.obj", "sphere.obj"])
def test_obj_loader_matches_reference(golden, tmp_path, name):
    g = golden["obj"][name]
    scene = {"materials": {f"m{k}": {} for k in range(3)},
             "objects": [{"type": "mesh", "file": os.path.join(REF_TREE, "models", name), "material": 2,
                          "transform": {"translate": g["translate"], "rotate": {"axis": g["axis"], "angle": g["angle"]},
                                        "scale": g["scale"]}}]}
    p = tmp_path / "m.json"
    p.write_text(json.dumps(scene))
    sm = pt.SceneManager()
    sm.load(str(p))
    v = host_view(sm)
    va, fa = view_arrays(v)
    assert (v.num_vertices, v.num_faces) == (g["nv"], g["nf"])
    assert sha(np.ascontiguousarray(va[:, 0:3]).tobytes()) == g["pos_sha"]
    assert sha(np.ascontiguousarray(va[:, 8:10]).tobytes()) == g["uv_sha"]
    assert sha(fa.tobytes()) == g["faces_sha"]
    # normals go through the inverse-transpose; compare to a tolerance
    nref = np.array([r[3:6] for r in g["first_vertices"]], np.float32)
    assert np.allclose(va[:4, 4:7], nref, atol=2e-6)


@needs_ref
def test_hdr_decode_matches_stb(golden, tmp_path):
    g = golden["hdr_default"]
    p = tmp_path / "e.json"
    p.write_text(json.dumps({"environment": {"file": os.path.join(REF_TREE, "default.hdr")}}))
    sm = pt.SceneManager()
    sm.load(str(p))
    v = host_view(sm)
    assert (v.env_width, v.env_height) == (g["w"], g["h"])
    img = np.ctypeslib.as_array(v.env_rgba, shape=(g["w"] * g["h"] * 4,)).reshape(g["h"], g["w"], 4)
    assert sha(img.tobytes()) == g["sha"]


def test_hdr_roundtrip_python_writer(tmp_path):
    from pathtracingtutorial_amd.imageio import read_hdr, write_hdr
    rng = np.random.default_rng(0)
    img = np.concatenate([rng.exponential(2.0, (9, 13, 3)), np.ones((9, 13, 1))], -1).astype(np.float32)
    write_hdr(str(tmp_path / "x.hdr"), img, flip_y=False)
    back = read_hdr(str(tmp_path / "x.hdr"))
    # RGBE keeps 8 mantissa bits relative to the largest channel of each pixel
    err = np.abs(back[..., :3] - img[..., :3])
    assert np.all(err <= img[..., :3].max(-1, keepdims=True) / 128 + 1e-30)
    sm = pt.SceneManager()
    (tmp_path / "s.json").write_text(json.dumps({"environment": {"file": "x.hdr"}}))
    sm.load(str(tmp_path / "s.json"))
    v = host_view(sm)
    c_img = np.ctypeslib.as_array(v.env_rgba, shape=(9 * 13 * 4,)).reshape(9, 13, 4)
    assert np.array_equal(c_img, back)


# ---------------------------------------------------------------- workloads
def test_cornell_workload():
    sm = pt.SceneManager()
    sm.generate(0)
    i = sm.info()
    assert (i.num_faces, i.width, i.height, i.target_samples, i.sample_per_frame, i.max_depth, i.sampler) == \
        (36, 512, 512, 16, 16, 8, 0)
    va, fa = view_arrays(host_view(sm))
    assert va[:, 0:3].min() >= -0.5 - 1e-6 and va[:, 0:3].max() <= 0.5 + 1e-6


def test_cornell_json_equals_generator():
    sm = pt.SceneManager()
    sm.load(os.path.join(ROOT, "scenes", "cornell_box.json"))
    a = host_view(sm)
    sg = pt.SceneManager()
    sg.generate(0)
    b = host_view(sg)
    va, fa = view_arrays(a)
    vb, fb = view_arrays(b)
    assert np.array_equal(fa, fb)
    assert np.allclose(va[:, 0:3], vb[:, 0:3], atol=1e-6)
    assert a.camera.fov == b.camera.fov and (a.camera.width, a.camera.height) == (b.camera.width, b.camera.height)


def test_spheres_workload_is_seeded():
    a, b, c = pt.SceneManager(), pt.SceneManager(), pt.SceneManager()
    a.generate(2, 1)
    b.generate(2, 1)
    c.generate(2, 2)
    va, _ = view_arrays(host_view(a))
    vb, _ = view_arrays(host_view(b))
    vc, _ = view_arrays(host_view(c))
    assert np.array_equal(va, vb) and not np.array_equal(va, vc)
