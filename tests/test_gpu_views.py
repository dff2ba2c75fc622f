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
