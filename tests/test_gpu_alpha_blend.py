"""A blended (alpha mode 2) surface under an area light and an environment,
through both GPU integrators and the oracle, with either sampler.

A blended surface draws one random number for its coverage test before the
light samples (path.cu:147-158), so the light samples the shading bounce
draws ahead of the surface gathers start one position too early there and
the bounce draws them again in place; elsewhere the drawn-ahead samples are
used.  Both integrators and the oracle use the same random numbers per path,
so their images agree pixel by pixel up to float rounding, and the
integrators' image means equal the oracle's."""
import ctypes as C

import numpy as np
import pytest

import oracle
import pathtracingtutorial_amd as pt
from pathtracingtutorial_amd import _capi
from test_gpu_wavefront import _same_paths
from test_oracle_furnace import _material

pytestmark = pytest.mark.gpu


def _scene():
    quads = [  # (y, half extent, material, normal y)
        (0.0, 2.0, 0, 1.0),
        (0.5, 1.0, 1, 1.0),
    ]
    verts, faces = [], []
    for y, e, m, ny in quads:
        b = len(verts)
        verts += [((-e, y, -e), ny), ((e, y, -e), ny), ((e, y, e), ny), ((-e, y, e), ny)]
        faces += [(b, b + 2, b + 1, m), (b, b + 3, b + 2, m)]
    b = len(verts)
    verts += [((-0.5, 2.0, -0.5), -1.0), ((0.5, 2.0, -0.5), -1.0), ((0.0, 2.0, 0.5), -1.0)]
    faces += [(b, b + 1, b + 2, 2)]
    V = (_capi.Vertex * len(verts))()
    for i, (p, ny) in enumerate(verts):
        V[i].position = _capi.Vec3p.of(p)
        V[i].normal = _capi.Vec3p.of((0.0, ny, 0.0))
    F = (_capi.Face * len(faces))()
    for i, f in enumerate(faces):
        F[i] = _capi.Face(*f)
    floor = _material(metallic=0.0, roughness=0.7)
    veil = _material(metallic=0.0, roughness=0.4, alpha_mode=2, base_alpha=0.4, double_sided=1)
    veil.base_color = _capi.Vec3p.of((0.2, 0.5, 0.9))
    lamp = _material(metallic=0.0, roughness=0.9, emissive_strength=1.0)
    lamp.emissive = _capi.Vec3p.of((6.0, 5.0, 4.0))
    M = (_capi.Material * 3)(floor, veil, lamp)
    env = np.zeros((16, 32, 4), np.float32)
    env[..., :3] = (0.25, 0.3, 0.4)
    env[..., 3] = 1.0
    cam = _capi.Camera()
    cam.position = _capi.Vec3p.of((0.0, 3.0, 2.5))
    cam.look_at = _capi.Vec3p.of((0.0, 0.0, 0.0))
    cam.up = _capi.Vec3p.of((0.0, 1.0, 0.0))
    cam.fov, cam.focal_distance = 0.7, 1.0
    cam.width, cam.height = 48, 32
    return V, F, M, env, cam


@pytest.mark.parametrize("sampler", [0, 1], ids=["pcg", "sobol"])
def test_blended_surface_under_area_and_environment_light(sampler):
    V, F, M, env, cam = _scene()
    st = _capi.RenderSettings()
    st.sample_per_frame, st.target_samples, st.max_depth, st.sampler = 16, 16, 6, sampler
    d = _capi.HostSceneDesc()
    d.vertices, d.num_vertices = V, len(V)
    d.faces, d.num_faces = F, len(F)
    d.materials, d.num_materials = M, len(M)
    d.camera = C.pointer(cam)
    d.env_rgba = env.ctypes.data_as(C.POINTER(C.c_float))
    d.env_width, d.env_height = 32, 16
    d.settings = C.pointer(st)
    sm = pt.SceneManager()
    sm.create(d)
    sm.buildAccel()
    imgs = {}
    for integ in (pt.PathIntegrator, pt.WavefrontIntegrator):
        sm.uploadToDevice()  # frame 0: the same sample indices
        r = pt.Renderer(sm, integ())
        r.renderBatch()
        imgs[integ.__name__] = r.downloadCurrentImage(1)[..., :3].astype(np.float64)
    mk, wf = imgs["PathIntegrator"], imgs["WavefrontIntegrator"]
    assert np.isfinite(wf).all() and wf.mean() > 0.05
    _same_paths(wf, mk, tol_frac=0.01)
    sc = oracle.OracleScene(bytes(V), len(V), bytes(F), len(F), bytes(M), len(M), env=env)
    ref, _ = sc.render(cam, 16, 6, sampler, 0, frame=0)
    ref = ref.astype(np.float64)
    _same_paths(wf, ref, tol_frac=0.05)
    assert np.allclose(wf.mean(axis=(0, 1)), ref.mean(axis=(0, 1)), rtol=5e-3)
    assert np.allclose(mk.mean(axis=(0, 1)), ref.mean(axis=(0, 1)), rtol=5e-3)
