"""The GPU integrators against the mathematics of a plane under a constant
environment (see test_oracle_furnace.py): with every escaping path the
environment and maxDepth >= 2, a pixel's expectation is 2 pi L * the
hemisphere integral of evalDisneyBsdf's f (cosine included), reached only
through the NEE + BSDF-sample MIS combination.  Checked for the megakernel and
the wavefront integrator (binary and 4-wide traversal)."""
import ctypes as C

import numpy as np
import pytest

import pathtracingtutorial_amd as pt
from pathtracingtutorial_amd import _capi
from test_oracle_furnace import LTEX, _hemisphere_integral, _material

pytestmark = pytest.mark.gpu


def _plane(mat):
    verts = (_capi.Vertex * 4)()
    for i, p in enumerate([(-1000, 0, -1000), (1000, 0, -1000), (1000, 0, 1000), (-1000, 0, 1000)]):
        verts[i].position = _capi.Vec3p.of(p)
        verts[i].normal = _capi.Vec3p.of((0, 1, 0))
    faces = (_capi.Face * 2)()
    faces[0] = _capi.Face(0, 2, 1, 0)
    faces[1] = _capi.Face(0, 3, 2, 0)
    mats = (_capi.Material * 1)(mat)
    env = np.zeros((32, 64, 4), np.float32)
    env[..., :3] = LTEX
    env[..., 3] = 1.0
    cam = _capi.Camera()
    cam.position = _capi.Vec3p.of((0.0, 1.0, 1.2))
    cam.look_at = _capi.Vec3p.of((0.0, 0.0, 0.0))
    cam.up = _capi.Vec3p.of((0.0, 1.0, 0.0))
    cam.fov, cam.focal_distance = 0.5, 1.0
    cam.width = cam.height = 16
    st = _capi.RenderSettings()
    st.sample_per_frame, st.target_samples, st.max_depth, st.sampler = 64, 64, 4, 0
    d = _capi.HostSceneDesc()
    d.vertices, d.num_vertices = verts, 4
    d.faces, d.num_faces = faces, 2
    d.materials, d.num_materials = mats, 1
    d.camera = C.pointer(cam)
    d.env_rgba = env.ctypes.data_as(C.POINTER(C.c_float))
    d.env_width, d.env_height = 64, 32
    d.settings = C.pointer(st)
    sm = pt.SceneManager()
    sm.create(d)
    sm.buildAccel()
    sm.uploadToDevice()
    return sm


@pytest.mark.parametrize("case", ["diffuse", "glossy-metal"])
def test_gpu_plane_under_constant_environment(case):
    kw = dict(metallic=0.0, roughness=0.8) if case == "diffuse" else dict(metallic=1.0, roughness=0.4)
    mat = _material(**kw)
    wo = np.array([0.0, 1.0, 1.2]) / np.linalg.norm([0.0, 1.0, 1.2])
    exact = 2 * np.pi * LTEX * _hemisphere_integral(mat, wo, np.array([0.0, 1.0, 0.0]))
    sm = _plane(mat)
    for integ, width in ((pt.PathIntegrator, 2), (pt.WavefrontIntegrator, 2), (pt.WavefrontIntegrator, 4)):
        sm.setWavefrontConfig(bvh_width=width)
        sm.uploadToDevice()
        r = pt.Renderer(sm, integ())
        means = []
        for _ in range(16):
            r.renderBatch()
            means.append(r.downloadCurrentImage(1)[..., :3].reshape(-1, 3).mean(0))
        means = np.array(means, np.float64)
        est, se = means.mean(0), means.std(0, ddof=1) / np.sqrt(len(means))
        assert np.all(np.abs(est - exact) <= 4 * se + 0.01 * exact), (integ.__name__, width, est, exact, se)


@pytest.mark.parametrize("case", ["diffuse", "glossy-metal"])
def test_gpu_plane_under_punctual_lights(case):
    from test_oracle_furnace import _lights, punctual_closed_form
    kw = dict(metallic=0.0, roughness=0.8) if case == "diffuse" else dict(metallic=1.0, roughness=0.4)
    mat = _material(**kw)
    wo = np.array([0.0, 1.0, 1.2]) / np.linalg.norm([0.0, 1.0, 1.2])
    exact = punctual_closed_form(mat, wo)
    verts = (_capi.Vertex * 4)()
    for i, p in enumerate([(-1000, 0, -1000), (1000, 0, -1000), (1000, 0, 1000), (-1000, 0, 1000)]):
        verts[i].position = _capi.Vec3p.of(p)
        verts[i].normal = _capi.Vec3p.of((0, 1, 0))
    faces = (_capi.Face * 2)()
    faces[0] = _capi.Face(0, 2, 1, 0)
    faces[1] = _capi.Face(0, 3, 2, 0)
    mats = (_capi.Material * 1)(mat)
    lights = _lights()
    cam = _capi.Camera()
    cam.position = _capi.Vec3p.of((0.0, 1.0, 1.2))
    cam.look_at = _capi.Vec3p.of((0.0, 0.0, 0.0))
    cam.up = _capi.Vec3p.of((0.0, 1.0, 0.0))
    cam.fov, cam.focal_distance = 0.2, 1.0
    cam.width = cam.height = 4
    st = _capi.RenderSettings()
    st.sample_per_frame, st.target_samples, st.max_depth, st.sampler = 4, 4, 3, 0
    d = _capi.HostSceneDesc()
    d.vertices, d.num_vertices = verts, 4
    d.faces, d.num_faces = faces, 2
    d.materials, d.num_materials = mats, 1
    d.lights, d.num_lights = lights, 2
    d.camera = C.pointer(cam)
    d.settings = C.pointer(st)
    sm = pt.SceneManager()
    sm.create(d)
    sm.buildAccel()
    sm.uploadToDevice()
    for integ in (pt.PathIntegrator, pt.WavefrontIntegrator):
        sm.uploadToDevice()
        r = pt.Renderer(sm, integ())
        r.renderBatch()
        img = r.downloadCurrentImage(1)[..., :3].reshape(-1, 3).astype(np.float64)
        assert np.allclose(img.mean(0), exact, rtol=1e-3), (integ.__name__, img.mean(0), exact)
