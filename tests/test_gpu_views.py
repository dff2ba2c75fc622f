"""Many cameras of one scene: the forward in one wavefront schedule and the
adjoint in one launch equal the per-camera renders and adjoints."""
import ctypes as C

import numpy as np
import pytest

import pathtracingtutorial_amd as pt
from pathtracingtutorial_amd import _capi

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _scene():
    sm = pt.SceneManager()
    sm.generate(3, 3)
    cam = sm.camera()
    cam.width, cam.height = 32, 32
    sm.updateCamera(cam)
    sm.setSamplePerFrame(4)
    sm.setMaxDepth(5)
    sm.buildAccel()
    sm.uploadToDevice()
    return sm


def _cams(sm, n=3):
    base = sm.camera()
    cams = []
    for i in range(n):
        c = _capi.Camera.from_buffer_copy(bytes(base))
        c.position.x = base.position.x + 0.15 * (i - 1)
        c.position.y = base.position.y + 0.05 * i
        cams.append(c)
    return cams


def test_views_forward_equals_single_renders():
    from pathtracingtutorial_amd.diff import MultiViewRenderer
    sm = _scene()
    cams = _cams(sm)
    mv = MultiViewRenderer(sm, cams)
    imgs = mv.forward_images(2).cpu().numpy()
    for i, c in enumerate(cams):
        sm.updateCamera(c)
        sm.uploadToDevice()
        _capi.check(pt.lib().pt_scene_set_frame(sm.handle, 2 * len(cams) + i))
        r = pt.Renderer(sm, pt.WavefrontIntegrator())
        r.renderBatch()
        assert np.array_equal(r.downloadCurrentImage(1), imgs[i])


def test_views_adjoint_equals_sum_of_single_adjoints():
    from pathtracingtutorial_amd.diff import DiffRenderer, MultiViewRenderer
    sm = _scene()
    cams = _cams(sm)
    mv = MultiViewRenderer(sm, cams)
    rng = np.random.default_rng(4)
    g = torch.from_numpy(rng.uniform(-1, 1, (len(cams), 32, 32, 4)).astype(np.float32)).cuda()
    multi = mv.adjoint(1, g)
    total = None
    for i, c in enumerate(cams):
        sm.updateCamera(c)
        dr = DiffRenderer(sm)
        out = dr.adjoint(1 * len(cams) + i, g[i].contiguous())
        total = out if total is None else {k: (v + out[k] if v is not None else None) for k, v in total.items()}
    for k in ("base_color", "roughness", "metallic", "texture", "env"):
        a, b = multi[k], total[k]
        scale = float(b.abs().max()) + 1e-12
        assert torch.allclose(a, b, rtol=1e-4, atol=1e-5 * scale), k


def test_views_autograd_shapes():
    from pathtracingtutorial_amd.diff import MultiViewRenderer
    sm = _scene()
    cams = _cams(sm, 2)
    mv = MultiViewRenderer(sm, cams)
    rough = torch.full((sm.info().num_materials,), 0.5, device="cuda", requires_grad=True)
    env_shape = (sm.info().env_height, sm.info().env_width, 4)
    env = torch.ones(env_shape, device="cuda", requires_grad=True)
    img = mv.render(0, roughness=rough, env=env)
    assert tuple(img.shape) == (2, 32, 32, 4)
    (img[..., :3] ** 2).mean().backward()
    assert rough.grad is not None and torch.isfinite(rough.grad).all()
    assert env.grad is not None and env.grad.shape == env.shape and env.grad.abs().sum() > 0


def test_ring_views_adjoint_matches_host_reference_adjoint():
    # The 16-view workload (generate(5): 2,048,002 triangles, 16 ring cameras at
    # 512^2, depth 10, gradients of the 1024^2 ground texture and the
    # 1024x512 environment) at 2 spp, all 16 views in one adjoint schedule.
    # The weights sit on three 24x24 tiles (ground, spheres, sky) of four
    # views, on pixels whose forward image equals the oracle's; the whole
    # texture and environment gradient arrays are compared with the sum of
    # the oracle's FP64 reverse-mode adjoints of those views, and 32 texels
    # and environment texels with the oracle's FP64 central differences.
    import oracle
    from pathtracingtutorial_amd.diff import MultiViewRenderer, ring_cameras
    from test_gpu_adjoint import _oracle_with_textures, _same_path_pixels, compare_arrays
    sm = pt.SceneManager()
    sm.generate(5, 1)
    sm.setSamplePerFrame(2)
    sm.buildAccel()
    sm.uploadToDevice()
    info = sm.info()
    assert info.num_faces == 2048002 and (info.env_width, info.env_height) == (1024, 512)
    cams = ring_cameras(16)
    mv = MultiViewRenderer(sm, cams)
    frame = 0
    imgs = mv.forward_images(frame).cpu().numpy()
    ora = _oracle_with_textures(sm)
    rng = np.random.default_rng(17)
    W = H = 512
    weights = np.zeros((16, H, W, 3), np.float32)
    tiles = [(244, 268, 60, 84), (244, 268, 244, 268), (244, 268, 440, 464)]  # ground, spheres, sky rows
    views = [0, 5, 10, 13]
    for v in views:
        for (x0, x1, y0, y1) in tiles:
            o, _ = ora.render(cams[v], 2, info.max_depth, info.sampler, info.seed, frame=frame * 16 + v,
                              rect=(x0, x1, y0, y1))
            same = _same_path_pixels(imgs[v, y0:y1, x0:x1], o)
            assert same.mean() >= 0.85, (v, same.mean())
            weights[v, y0:y1, x0:x1] = rng.uniform(-1, 1, (y1 - y0, x1 - x0, 3)) * same[..., None]
    g = torch.zeros((16, H, W, 4), dtype=torch.float32, device="cuda")
    g[..., :3] = torch.from_numpy(weights).cuda()
    out = mv.adjoint(frame, g, (True, True, True, True, True))
    tex = np.zeros((1024 * 1024, 4))
    env = np.zeros((512, 1024, 4))
    for v in views:
        h = ora.adjoint(cams[v], 2, info.max_depth, info.sampler, info.seed, frame * 16 + v, weights[v])
        tex += h["textures"][0].reshape(-1, 4)
        env += h["env"]
    # the 1024^2 ground texture is seen from afar: most touched texels get one
    # or two lookups, so a 1/256 filter-weight step (see compare_arrays) shows
    # on a larger share of them than in the close-up Cornell texture
    rt = compare_arrays("ring views texture", out["texture"].cpu().numpy()[:, :3], tex[:, :3], frac=0.95)
    re = compare_arrays("ring views env", out["env"].cpu().numpy()[..., :3], env[..., :3], frac=0.99)
    print(f"ring views FP32 (GPU atomics) vs FP64 (host): max |d| / max |g|: texture {rt:.2e}, env {re:.2e}")
    assert (np.abs(tex[:, :3]) > 0).sum() > 1000 and (np.abs(env[..., :3]) > 0).sum() > 1000
    # central differences of the oracle's detached estimator on 32 scalars
    flat_t = np.argsort(-np.abs(tex[:, :3]).reshape(-1))[:16]
    flat_e = np.argsort(-np.abs(env[..., :3]).reshape(-1))[:16]
    params = [(3, 0, int(k // 3), int(k % 3)) for k in flat_t] + [(4, 0, int(k // 3), int(k % 3)) for k in flat_e]
    fd = np.zeros(len(params))
    for v in views:
        d, _ = ora.adjoint_fd(cams[v], 2, info.max_depth, info.sampler, info.seed, frame * 16 + v, weights[v],
                              params, h=1e-4)
        fd += d
    gt, ge = out["texture"].cpu().numpy(), out["env"].cpu().numpy().reshape(-1, 4)
    got = np.array([gt[b, c] if k == 3 else ge[b, c] for (k, a, b, c) in params])
    scale = np.abs(fd).max()
    assert np.all(np.abs(got - fd) <= 1e-3 * np.abs(fd) + 1e-5 * scale), list(zip(params, got, fd))


def _reused(sm):
    r = C.c_int(-1)
    _capi.check(pt.lib().pt_scene_get_keep_radiance(sm.handle, None, C.byref(r)))
    return bool(r.value)


def test_adjoint_seeded_from_kept_forward_radiance_equals_recomputed():
    # The forward keeps each path's radiance; the adjoint of the same batch
    # seeds its replay from it (no primal pass) and gives the gradients of the
    # adjoint that recomputes the primal pass.  A parameter update, another
    # frame or keep turned off each make the adjoint recompute.
    from pathtracingtutorial_amd.diff import MultiViewRenderer
    sm = _scene()
    cams = _cams(sm)
    mv = MultiViewRenderer(sm, cams)
    rng = np.random.default_rng(9)
    g = torch.from_numpy(rng.uniform(-1, 1, (len(cams), 32, 32, 4)).astype(np.float32)).cuda()
    mv.forward_images(1)
    kept = mv.adjoint(1, g)
    assert mv.last_backward_reused and _reused(sm)
    _capi.check(pt.lib().pt_scene_set_keep_radiance(sm.handle, 0))
    fresh = mv.adjoint(1, g)
    assert not mv.last_backward_reused
    for k in ("base_color", "roughness", "metallic", "texture", "env"):
        a, b = kept[k], fresh[k]
        scale = float(b.abs().max()) + 1e-12
        assert float(b.abs().sum()) > 0, k
        # the same terms summed by FP32 atomics in another order (the bar of
        # the views-vs-single-views comparison above)
        assert torch.allclose(a, b, rtol=1e-4, atol=1e-5 * scale), (k, float((a - b).abs().max()), scale)
    _capi.check(pt.lib().pt_scene_set_keep_radiance(sm.handle, 1))
    mv.forward_images(1)
    mv.adjoint(2, g)  # another batch
    assert not mv.last_backward_reused
    mv.forward_images(1)
    mv.push(roughness=torch.full((sm.info().num_materials,), 0.4, device="cuda"))
    mv.adjoint(1, g)  # parameters changed since the forward
    assert not mv.last_backward_reused
    mv.forward_images(1)
    mv.adjoint(1, g)
    assert mv.last_backward_reused


def test_single_view_wavefront_forward_keeps_radiance_for_adjoint():
    from pathtracingtutorial_amd.diff import DiffRenderer
    sm = _scene()
    dr = DiffRenderer(sm)
    _capi.check(pt.lib().pt_scene_set_keep_radiance(sm.handle, 1))
    rng = np.random.default_rng(11)
    g = torch.from_numpy(rng.uniform(-1, 1, (32, 32, 4)).astype(np.float32)).cuda()
    _capi.check(pt.lib().pt_scene_set_frame(sm.handle, 3))
    f = C.c_void_p()
    _capi.check(pt.lib().pt_film_create(32, 32, C.byref(f)))
    ms = C.c_float(0.0)
    _capi.check(pt.lib().pt_render_batch(sm.handle, f, _capi.PT_INTEGRATOR_WAVEFRONT, C.byref(ms)))
    img = np.zeros((32, 32, 4), np.float32)
    _capi.check(pt.lib().pt_film_download(f, 1, img.ctypes.data_as(C.POINTER(C.c_float))))
    kept = dr.adjoint(3, g, replay=True)
    assert _reused(sm)
    _capi.check(pt.lib().pt_scene_set_keep_radiance(sm.handle, 0))
    fresh = dr.adjoint(3, g, replay=True)
    assert not _reused(sm)
    # the replay's image: the forward's batch image, both ways
    assert np.array_equal(kept["replay"].cpu().numpy()[..., :3], img[..., :3])
    assert torch.equal(kept["replay"], fresh["replay"])
    for k in ("base_color", "roughness", "metallic", "texture", "env"):
        a, b = kept[k], fresh[k]
        scale = float(b.abs().max()) + 1e-12
        assert torch.allclose(a, b, rtol=1e-4, atol=1e-5 * scale), (k, float((a - b).abs().max()), scale)
    pt.lib().pt_film_destroy(f)


def test_kept_radiance_across_chunks():
    # chunk_spp 1 with 4 samples per pixel: the batch runs as four chunks.  Each
    # chunk's forward keeps its paths' radiance at their global index and the
    # adjoint's chunks read it back: the replay's image is the forward's and the
    # gradients are those of the adjoint that recomputes the primal pass.
    from pathtracingtutorial_amd.diff import DiffRenderer
    sm = _scene()
    sm.setWavefrontConfig(chunk_spp=1)
    dr = DiffRenderer(sm)
    rng = np.random.default_rng(12)
    g = torch.from_numpy(rng.uniform(-1, 1, (32, 32, 4)).astype(np.float32)).cuda()
    img = dr.forward_image(2)
    kept = dr.adjoint(2, g, replay=True)
    assert dr.last_backward_reused
    assert torch.equal(kept["replay"][..., :3], img[..., :3])
    _capi.check(pt.lib().pt_scene_set_keep_radiance(sm.handle, 0))
    fresh = dr.adjoint(2, g, replay=True)
    assert not dr.last_backward_reused
    assert torch.equal(kept["replay"], fresh["replay"])
    for k in ("base_color", "roughness", "metallic", "texture", "env"):
        a, b = kept[k], fresh[k]
        scale = float(b.abs().max()) + 1e-12
        assert float(b.abs().sum()) > 0, k
        assert torch.allclose(a, b, rtol=1e-4, atol=1e-5 * scale), (k, float((a - b).abs().max()), scale)
    # one chunk (the default) gives the same image and, to atomic order, gradients
    sm.setWavefrontConfig(chunk_spp=64)
    _capi.check(pt.lib().pt_scene_set_keep_radiance(sm.handle, 1))
    img1 = dr.forward_image(2)
    one = dr.adjoint(2, g, replay=True)
    assert dr.last_backward_reused
    assert torch.equal(img1, img)
    for k in ("base_color", "roughness", "metallic", "texture", "env"):
        a, b = one[k], fresh[k]
        scale = float(b.abs().max()) + 1e-12
        assert torch.allclose(a, b, rtol=1e-4, atol=1e-5 * scale), (k, float((a - b).abs().max()), scale)
