"""Edge cases of the render path through the C ABI: a scene without
geometry (every camera ray escapes), a 1x1 film, films whose sides are not
multiples of the 16x16 tile, one sample per frame.  The megakernel's and the
wavefront's images are bit-identical; the scene without geometry also
matches the oracle per pixel (both sides draw the same random numbers)."""
import ctypes as C

import numpy as np
import pytest

import oracle
import pathtracingtutorial_amd as pt
from pathtracingtutorial_amd import _capi

pytestmark = pytest.mark.gpu


def _host_view(sm):
    v = _capi.HostSceneView()
    _capi.check(pt.lib().pt_scene_host_view(sm.handle, C.byref(v)))
    return v


def _sky_only(w, h, spf, sampler):
    """No triangles; a 32x16 environment with a gradient in u and in v."""
    mats = (_capi.Material * 1)()
    verts = (_capi.Vertex * 1)()
    faces = (_capi.Face * 1)()
    env = np.zeros((16, 32, 4), np.float32)
    env[..., 0] = np.linspace(0.2, 2.0, 32, dtype=np.float32)[None, :]
    env[..., 1] = np.linspace(0.5, 1.5, 16, dtype=np.float32)[:, None]
    env[..., 2] = 0.7
    env[..., 3] = 1.0
    cam = _capi.Camera()
    cam.position = _capi.Vec3p.of((0.0, 0.0, 0.0))
    cam.look_at = _capi.Vec3p.of((0.3, 0.2, -1.0))
    cam.up = _capi.Vec3p.of((0.0, 1.0, 0.0))
    cam.fov, cam.focal_distance = 70.0, 1.0
    cam.width, cam.height = w, h
    st = _capi.RenderSettings()
    st.sample_per_frame, st.target_samples, st.max_depth, st.sampler = spf, 4 * spf, 4, sampler
    d = _capi.HostSceneDesc()
    d.vertices, d.num_vertices = verts, 0
    d.faces, d.num_faces = faces, 0
    d.materials, d.num_materials = mats, 1
    d.camera = C.pointer(cam)
    d.env_rgba = env.ctypes.data_as(C.POINTER(C.c_float))
    d.env_width, d.env_height = 32, 16
    d.settings = C.pointer(st)
    sm = pt.SceneManager()
    sm.create(d)
    sm._env_keepalive = env
    sm.buildAccel()
    sm.uploadToDevice()
    return sm


def _frames(sm, integ, batches):
    r = pt.Renderer(sm, integ)
    out = []
    for _ in range(batches):
        r.renderBatch()
        out.append(r.downloadCurrentImage(1).copy())
    return np.array(out), r.downloadCurrentImage(0).copy()


def _cornell(w, h, spf):
    sm = pt.SceneManager()
    sm.generate(0)
    cam = sm.camera()
    cam.width, cam.height = w, h
    sm.updateCamera(cam)
    sm.setSamplePerFrame(spf)
    sm.setMaxDepth(8)
    sm.buildAccel()
    sm.uploadToDevice()
    return sm


def _both_integrators_equal(sm, batches=2):
    mk, mk_acc = _frames(sm, pt.PathIntegrator(), batches)
    assert np.isfinite(mk).all() and np.isfinite(mk_acc).all()
    for width in (2, 4):
        sm.setWavefrontConfig(bvh_width=width)
        sm.uploadToDevice()
        wf, wf_acc = _frames(sm, pt.WavefrontIntegrator(), batches)
        assert np.array_equal(wf, mk), f"{width}-wide: {np.count_nonzero(wf != mk)} channels differ"
        assert np.array_equal(wf_acc, mk_acc)
    return mk, mk_acc


@pytest.mark.parametrize("sampler", [0, 1], ids=["pcg", "sobol"])
def test_scene_without_geometry_sees_the_environment(sampler):
    sm = _sky_only(24, 10, 4, sampler)
    info = sm.info()
    assert info.num_faces == 0
    mk, _ = _both_integrators_equal(sm)
    # the oracle on the same random numbers: each sample's radiance is the
    # environment along its camera ray (x 2 pi, envmap.h:43-45); the two
    # sides differ only where libm's atan2/acos round differently
    ora = oracle.OracleScene.from_view(_host_view(sm))
    for b in range(mk.shape[0]):
        img, _ = ora.render(sm.camera(), 4, info.max_depth, info.sampler, info.seed, frame=b)
        gpu = mk[b][..., :3].astype(np.float64)
        cpu = img[..., :3].astype(np.float64)
        rel = np.abs(gpu - cpu) / np.maximum(np.abs(cpu), 1e-6)
        assert (rel <= 1e-4).mean() >= 0.98, (b, (rel <= 1e-4).mean())
        assert abs(gpu.mean() - cpu.mean()) <= 1e-3 * cpu.mean()
    # the green channel's texels average 1 over v: the image is near 2 pi
    assert 0.5 * 2 * np.pi < mk[..., 1].mean() < 1.5 * 2 * np.pi


@pytest.mark.parametrize("size", [(1, 1), (17, 9), (33, 7), (5, 40)], ids=lambda s: f"{s[0]}x{s[1]}")
def test_films_not_a_multiple_of_the_tile(size):
    sm = _cornell(size[0], size[1], 4)
    mk, acc = _both_integrators_equal(sm)
    assert mk.shape[1:3] == (size[1], size[0])
    # the running mean of two batches is their mean (renderer.cu accumulation)
    assert np.allclose(acc[..., :3], 0.5 * (mk[0][..., :3] + mk[1][..., :3]), rtol=1e-6, atol=1e-7)


def test_one_sample_per_frame():
    sm = _cornell(32, 24, 1)
    mk, acc = _both_integrators_equal(sm, batches=3)
    assert mk[..., :3].mean() > 0
    assert np.allclose(acc[..., :3], mk[..., :3].mean(0), rtol=1e-5, atol=1e-7)
