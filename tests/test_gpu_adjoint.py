"""Path-replay adjoint on the GPU against the oracle's FP64 detached estimator.

For a weight image w, the adjoint returns d/dtheta sum(w * I) for every
material factor, texel and environment texel.  The oracle records each path
of the same batch (same draws, same decisions) and takes central differences
of the detached estimator (directions, pdfs, MIS weights and Russian
roulette held) in double precision: the two agree path by path, so the
tolerance is the 1e-3 relative bar, plus a small absolute floor for
parameters whose gradient is (nearly) zero.  Both forms of the adjoint (one
thread per pixel, and the wavefront schedule's primal + replay passes) are
checked; they compute the same estimator and agree to float summation order."""
import ctypes as C
import json
import os

import numpy as np
import pytest

import oracle
import pathtracingtutorial_amd as pt
from pathtracingtutorial_amd import _capi
from test_gpu_parity import host_view

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

FORMS = {"path": pt.PathIntegrator, "wavefront": pt.WavefrontIntegrator}
CORNELL = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "scenes", "cornell_box.json")


def _oracle_with_textures(sm):
    ora = oracle.OracleScene.from_view(host_view(sm))
    for i in range(sm.info().num_textures):
        d = _capi.TextureDesc()
        _capi.check(pt.lib().pt_scene_texture(sm.handle, i, C.byref(d)))
        rgba = np.ctypeslib.as_array(d.rgba, shape=(d.height * d.width * 4,)).reshape(d.height, d.width, 4).copy()
        ora.add_texture(rgba, d.wrap_s, d.wrap_t, d.linear)
    return ora


def _diff_cornell(w=48, h=48, spf=4, depth=6, seed=3):
    sm = pt.SceneManager()
    sm.generate(3, seed)
    cam = sm.camera()
    cam.width, cam.height = w, h
    sm.updateCamera(cam)
    sm.setSamplePerFrame(spf)
    sm.setMaxDepth(depth)
    sm.buildAccel()
    sm.uploadToDevice()
    # the binary BVH, which the per-pixel form also walks: both forms then
    # resolve an exact-distance tie the same way, so they trace the same paths
    # (the 4-wide tree may resolve such a tie the other way; checked on its
    # own below)
    sm.setWavefrontConfig(bvh_width=2)
    return sm


def test_diff_cornell_workload():
    sm = _diff_cornell()
    info = sm.info()
    assert info.num_materials == 8 and info.num_textures == 1
    assert (info.env_width, info.env_height) == (512, 256)
    assert info.num_faces == 36


@pytest.mark.parametrize("form", sorted(FORMS))
def test_adjoint_replay_reproduces_forward(form):
    from pathtracingtutorial_amd.diff import DiffRenderer
    sm = _diff_cornell()
    sm.setAdjointIntegrator(FORMS[form]())
    dr = DiffRenderer(sm)
    img = dr.forward_image(5)
    g = torch.ones_like(img)
    out = dr.adjoint(5, g, replay=True)
    rep = out["replay"]
    assert torch.isfinite(rep).all()
    # the replay pass walks the same paths with the same arithmetic
    assert torch.allclose(rep[..., :3], img[..., :3], rtol=1e-5, atol=1e-6)


def test_wavefront_replay_sums_radiance_as_the_primal_pass():
    # the replay starts from the stored sums and adds each path's emitter,
    # escape and light terms with the same roundings and in the same order as
    # the primal pass, so the two images agree to the bit (the pixel adjoint
    # the replay applies is L - lo, exact when they do)
    from pathtracingtutorial_amd.diff import DiffRenderer
    sm = _diff_cornell()
    sm.setAdjointIntegrator(pt.WavefrontIntegrator())
    dr = DiffRenderer(sm)
    frame = 5
    _capi.check(pt.lib().pt_scene_set_frame(sm.handle, frame))
    ms = C.c_float(0.0)
    _capi.check(pt.lib().pt_render_batch(sm.handle, dr._film, _capi.PT_INTEGRATOR_WAVEFRONT, C.byref(ms)))
    img = torch.empty((dr.height, dr.width, 4), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    _capi.check(pt.lib().pt_film_copy_device(dr._film, 1, C.c_void_p(img.data_ptr())))
    out = dr.adjoint(frame, torch.ones_like(img), replay=True)
    rep = out["replay"][..., :3]
    diff = (rep - img[..., :3]).abs()
    assert torch.equal(rep, img[..., :3]), f"{int((diff > 0).sum())} channels differ, max {float(diff.max()):.3g}"


def _params(sm, rng, n):
    info = sm.info()
    mats = [0, 1, 2, 3, 4, 6, 7]
    ps = []
    for _ in range(n):
        k = int(rng.integers(0, 5))
        if k == 0:
            ps.append((0, int(rng.choice([0, 1, 3, 4, 6, 7])), 0, int(rng.integers(0, 3))))
        elif k in (1, 2):
            ps.append((k, int(rng.choice(mats)), 0, 0))
        elif k == 3:  # a texel in the visible middle of the back-wall texture
            x, y = int(rng.integers(64, 192)), int(rng.integers(64, 192))
            ps.append((3, 0, y * 256 + x, int(rng.integers(0, 3))))
        else:  # an environment texel near the horizon band seen through the opening
            x, y = int(rng.integers(0, info.env_width)), int(rng.integers(96, 160))
            ps.append((4, 0, y * info.env_width + x, int(rng.integers(0, 3))))
    return ps


def _gpu_grad(out, p, tex_w=256):
    kind, a, b, c = p
    if kind == 0:
        return float(out["base_color"][a, c])
    if kind == 1:
        return float(out["roughness"][a])
    if kind == 2:
        return float(out["metallic"][a])
    if kind == 3:
        return float(out["texture"][b, c])
    return float(out["env"].reshape(-1, 4)[b, c])


@pytest.mark.parametrize("form", sorted(FORMS))
def test_adjoint_matches_fp64_detached_fd(form):
    from pathtracingtutorial_amd.diff import DiffRenderer
    sm = _diff_cornell(w=48, h=48, spf=4)
    sm.setAdjointIntegrator(FORMS[form]())
    cam, info = sm.camera(), sm.info()
    rng = np.random.default_rng(11)
    weights = rng.uniform(-1.0, 1.0, size=(info.height, info.width, 3)).astype(np.float32)
    dr = DiffRenderer(sm)
    frame = 2
    img = dr.forward_image(frame)
    g = torch.zeros((info.height, info.width, 4), dtype=torch.float32, device="cuda")
    g[..., :3] = torch.from_numpy(weights).cuda()
    out = dr.adjoint(frame, g)
    # plus gradients that exercise every kind: the most-hit texel and env texel
    tex_g = out["texture"].abs().sum(1).cpu().numpy()
    env_g = out["env"].reshape(-1, 4).abs().sum(1).cpu().numpy()
    params = _params(sm, rng, 32) + [(3, 0, int(tex_g.argmax()), 1), (4, 0, int(env_g.argmax()), 2)]
    ora = _oracle_with_textures(sm)
    fd, value = ora.adjoint_fd(cam, info.sample_per_frame, info.max_depth, info.sampler, info.seed, frame, weights,
                               params, h=1e-4)
    # the recorded paths evaluate to the forward image
    gpu_value = float((img[..., :3].cpu().numpy().astype(np.float64) * weights).sum())
    assert abs(gpu_value - value) <= 1e-4 * abs(value) + 1e-3
    scale = np.abs(fd).max()
    assert scale > 0
    bad = []
    for p, f in zip(params, fd):
        gg = _gpu_grad(out, p)
        if abs(gg - f) > 1e-3 * abs(f) + 1e-5 * scale:
            bad.append((p, gg, f))
    assert not bad, f"{len(bad)} of {len(params)} gradients off: {bad[:6]}"
    assert sum(abs(f) > 1e-3 * scale for f in fd) >= 16  # most checks carry a real gradient


def test_autograd_function_and_descent():
    # recover the base colours of the diffuse-looking materials from a
    # 64-spp target: the parameter error must shrink under Adam
    from pathtracingtutorial_amd.diff import DiffRenderer
    sm = _diff_cornell(w=32, h=32, spf=16)
    dr = DiffRenderer(sm)
    v = host_view(sm)
    mats = (_capi.Material * v.num_materials).from_buffer_copy(C.string_at(v.materials, 256 * v.num_materials))
    gt = torch.tensor([[m.base_color.x, m.base_color.y, m.base_color.z] for m in mats], device="cuda")
    target = torch.stack([dr.forward_image(f) for f in range(4)]).mean(0)
    bc = gt.clone()
    learn = [0, 1, 3, 4, 7]
    bc[learn] = 0.5
    bc.requires_grad_(True)
    opt = torch.optim.Adam([bc], lr=2e-2)
    err0 = float((bc.detach()[learn] - gt[learn]).abs().mean())
    for it in range(40):
        opt.zero_grad()
        img = dr.render(100 + it, base_color=bc)
        loss = ((img[..., :3] - target[..., :3]) ** 2).mean()
        loss.backward()
        assert bc.grad is not None and torch.isfinite(bc.grad).all()
        with torch.no_grad():
            bc.grad[[2, 5, 6]] = 0.0  # textured wall, light, tall box stay
        opt.step()
        with torch.no_grad():
            bc.clamp_(0.0, 1.0)
    err1 = float((bc.detach()[learn] - gt[learn]).abs().mean())
    assert err1 < 0.6 * err0, (err0, err1)


def _both_forms(sm, frame, g, replay=False):
    from pathtracingtutorial_amd.diff import DiffRenderer
    dr = DiffRenderer(sm)
    out = {}
    for form in ("path", "wavefront"):
        sm.setAdjointIntegrator(FORMS[form]())
        out[form] = dr.adjoint(frame, g, replay=replay)
    return out


def _assert_same_grads(a, b, keys=("base_color", "roughness", "metallic", "texture", "env")):
    for k in keys:
        if a[k] is None:
            assert b[k] is None
            continue
        scale = float(a[k].abs().max()) + 1e-12
        diff = float((a[k] - b[k]).abs().max())
        assert diff <= 2e-4 * scale, (k, diff, scale)


def test_wavefront_adjoint_equals_megakernel_adjoint():
    sm = _diff_cornell(w=64, h=48, spf=8)
    rng = np.random.default_rng(5)
    g = torch.from_numpy(rng.uniform(-1, 1, (48, 64, 4)).astype(np.float32)).cuda()
    out = _both_forms(sm, 3, g, replay=True)
    assert sm.adjointIntegrator() == "Wavefront"
    _assert_same_grads(out["path"], out["wavefront"])
    # both replays are the forward batch
    assert torch.allclose(out["path"]["replay"], out["wavefront"]["replay"], rtol=1e-5, atol=1e-6)


def test_wavefront_adjoint_wide_bvh_matches_binary():
    # The 4-wide BVH traces the same paths (exact-distance ties aside): the
    # gradients and the replayed image agree with the binary tree's.
    sm = _diff_cornell(w=64, h=48, spf=8)
    rng = np.random.default_rng(5)
    g = torch.from_numpy(rng.uniform(-1, 1, (48, 64, 4)).astype(np.float32)).cuda()
    from pathtracingtutorial_amd.diff import DiffRenderer
    dr = DiffRenderer(sm)
    sm.setAdjointIntegrator(pt.WavefrontIntegrator())
    out = {}
    for width in (2, 4):
        sm.setWavefrontConfig(bvh_width=width)
        out[width] = dr.adjoint(3, g, replay=True)
    _assert_same_grads(out[2], out[4])
    a, b = out[2]["replay"][..., :3], out[4]["replay"][..., :3]
    close = (a - b).abs() <= 1e-5 * (1 + a.abs())
    assert float(close.float().mean()) >= 0.999


def test_wavefront_adjoint_chunked_equals_single_chunk():
    # one chunk per sample (chunk_spp=1) against all samples in flight
    sm = _diff_cornell(w=40, h=40, spf=8)
    g = torch.ones((40, 40, 4), dtype=torch.float32, device="cuda")
    from pathtracingtutorial_amd.diff import DiffRenderer
    dr = DiffRenderer(sm)
    sm.setAdjointIntegrator(pt.WavefrontIntegrator())
    outs = []
    for cfg in (dict(chunk_spp=8), dict(chunk_spp=1, use_graph=0), dict(chunk_spp=2, sort_by_material=0)):
        sm.setWavefrontConfig(**cfg)
        outs.append(dr.adjoint(1, g))
    for o in outs[1:]:
        _assert_same_grads(outs[0], o)


def test_wavefront_adjoint_punctual_lights(tmp_path):
    scene = json.load(open(CORNELL))
    scene["camera"]["resolution"] = [40, 40]
    scene["render"]["samplePerFrame"] = 4
    scene["render"]["spp"] = 4
    scene["lights"] = [
        {"type": "point", "position": [0.2, 0.3, 0.1], "color": [1, 0.9, 0.8], "intensity": 0.5},
        {"type": "spot", "position": [-0.2, 0.4, 0.2], "direction": [0.1, -1, -0.1], "color": [0.3, 0.6, 1.0],
         "intensity": 2.0, "innerCone": 0.3, "outerCone": 0.6},
    ]
    path = tmp_path / "cornell_lights.json"
    path.write_text(json.dumps(scene))
    sm = pt.SceneManager()
    sm.load(str(path))
    sm.buildAccel()
    sm.uploadToDevice()
    assert sm.info().num_lights == 2
    rng = np.random.default_rng(9)
    g = torch.from_numpy(rng.uniform(0, 1, (40, 40, 4)).astype(np.float32)).cuda()
    out = _both_forms(sm, 0, g, replay=True)
    _assert_same_grads(out["path"], out["wavefront"], keys=("base_color", "roughness", "metallic"))
    assert float(out["wavefront"]["base_color"].abs().sum()) > 0
    assert torch.allclose(out["path"]["replay"], out["wavefront"]["replay"], rtol=1e-4, atol=1e-5)


def _same_path_pixels(img, ref, tol=1e-5):
    """Pixels whose batch mean equals the oracle's to float rounding in every
    channel: all their paths took the oracle's branches."""
    img = np.asarray(img, np.float64)[..., :3]
    ref = np.asarray(ref, np.float64)
    return np.all(np.abs(img - ref) <= tol * (1.0 + np.abs(ref)), axis=-1)


def compare_arrays(name, gpu, host, rel=1e-3, floor=1e-5, frac=0.995):
    """Element by element, |gpu - host| <= rel |host| + floor max|host| for at
    least `frac` of the elements either side gives a gradient (all of them
    when frac = 1), and the whole array within rel in the L1 norm.  Texel gradients pass through bilinear filter weights
    quantized to 1/256 (the texture unit's, cuda_texture.cu:56-60), taken at
    a lookup coordinate that the GPU's and the host's libm (atan2, asin of
    the environment mapping) may round one ulp apart: now and then a weight
    step moves between neighbouring texels; the L1 norm does not see it."""
    gpu = np.asarray(gpu, np.float64)
    host = np.asarray(host, np.float64)
    scale = float(np.abs(host).max())
    err = np.abs(gpu - host)
    ok = err <= rel * np.abs(host) + floor * scale
    nz = host != 0
    big = np.abs(host) > 1e-3 * scale
    worst = float((err[big] / np.abs(host[big])).max()) if big.any() else 0.0
    l1 = float(err.sum() / max(np.abs(host).sum(), 1e-300))
    print(f"{name}: {host.size} elements, {int(nz.sum())} nonzero, {int((~ok).sum())} outside tolerance, "
          f"max |d| / max|host| {float(err.max()) / max(scale, 1e-300):.2e}, worst relative (|host| > 1e-3 max) "
          f"{worst:.2e}, L1 {l1:.2e}")
    touched = nz | (gpu != 0)  # a weight step can also reach a texel the other side gives weight 0
    assert ok[touched].mean() >= frac, (name, np.argwhere(~ok)[:5], gpu[~ok][:5], host[~ok][:5])
    assert l1 <= rel, (name, l1)
    return float(err.max()) / max(scale, 1e-300)


@pytest.mark.parametrize("form", sorted(FORMS))
def test_adjoint_full_arrays_match_host_reference_adjoint(form):
    # Every element of every gradient array (8 x 3 base colours, 8 roughness,
    # 8 metallic, the 256^2 albedo texture, the 512x256 environment) against
    # the oracle's FP64 reverse-mode adjoint of the same batch.  The weights
    # are restricted to pixels whose forward image equals the oracle's
    # (nearly all): there the GPU traced the oracle's paths.
    from pathtracingtutorial_amd.diff import DiffRenderer
    sm = _diff_cornell(w=32, h=32, spf=4)
    sm.setAdjointIntegrator(FORMS[form]())
    cam, info = sm.camera(), sm.info()
    dr = DiffRenderer(sm)
    frame = 1
    img = dr.forward_image(frame).cpu().numpy()
    ora = _oracle_with_textures(sm)
    ref_img, _ = ora.render(cam, info.sample_per_frame, info.max_depth, info.sampler, info.seed, frame=frame)
    same = _same_path_pixels(img, ref_img)
    assert same.mean() >= 0.98, same.mean()
    rng = np.random.default_rng(21)
    weights = (rng.uniform(-1.0, 1.0, (info.height, info.width, 3)) * same[..., None]).astype(np.float32)
    g = torch.zeros((info.height, info.width, 4), dtype=torch.float32, device="cuda")
    g[..., :3] = torch.from_numpy(weights).cuda()
    out = dr.adjoint(frame, g)
    host = ora.adjoint(cam, info.sample_per_frame, info.max_depth, info.sampler, info.seed, frame, weights)
    compare_arrays("base_color", out["base_color"].cpu().numpy(), host["base_color"], frac=1.0)
    compare_arrays("roughness", out["roughness"].cpu().numpy(), host["roughness"], frac=1.0)
    compare_arrays("metallic", out["metallic"].cpu().numpy(), host["metallic"], frac=1.0)
    compare_arrays("texture", out["texture"].cpu().numpy()[:, :3], host["textures"][0].reshape(-1, 4)[:, :3])
    compare_arrays("env", out["env"].cpu().numpy()[..., :3], host["env"][..., :3])


def _quads_under_sky(ew, eh, seed=4):
    """A diffuse ground quad and a tilted glossy quad under a random ew x eh
    environment, seen from above with the sky behind them."""
    from test_oracle_furnace import _material
    rng = np.random.default_rng(seed)
    pts = [(-4, 0, -4), (4, 0, -4), (4, 0, 4), (-4, 0, 4), (-1, 0.3, -1), (1, 1.3, -1), (1, 1.3, 1), (-1, 0.3, 1)]
    verts = (_capi.Vertex * 8)()
    for i, p in enumerate(pts):
        verts[i].position = _capi.Vec3p.of(p)
        n = np.array([0.0, 1.0, 0.0]) if i < 4 else np.array([-1.0, 2.0, 0.0]) / np.sqrt(5.0)
        verts[i].normal = _capi.Vec3p.of(tuple(n))
    faces = (_capi.Face * 4)()
    faces[0], faces[1] = _capi.Face(0, 2, 1, 0), _capi.Face(0, 3, 2, 0)
    faces[2], faces[3] = _capi.Face(4, 6, 5, 1), _capi.Face(4, 7, 6, 1)
    mats = (_capi.Material * 2)(_material(roughness=0.7), _material(metallic=0.8, roughness=0.3))
    env = np.ones((eh, ew, 4), np.float32)
    env[..., :3] = rng.uniform(0.2, 3.0, (eh, ew, 3)).astype(np.float32)
    cam = _capi.Camera()
    cam.position = _capi.Vec3p.of((0.0, 3.0, 6.0))
    cam.look_at = _capi.Vec3p.of((0.0, 0.5, 0.0))
    cam.up = _capi.Vec3p.of((0.0, 1.0, 0.0))
    cam.fov, cam.focal_distance = 60.0, 1.0
    cam.width, cam.height = 48, 32
    st = _capi.RenderSettings()
    st.sample_per_frame, st.target_samples, st.max_depth, st.sampler = 8, 8, 5, 0
    d = _capi.HostSceneDesc()
    d.vertices, d.num_vertices = verts, 8
    d.faces, d.num_faces = faces, 4
    d.materials, d.num_materials = mats, 2
    d.camera = C.pointer(cam)
    d.env_rgba = env.ctypes.data_as(C.POINTER(C.c_float))
    d.env_width, d.env_height = ew, eh
    d.settings = C.pointer(st)
    sm = pt.SceneManager()
    sm.create(d)
    sm.buildAccel()
    sm.uploadToDevice()
    sm.setWavefrontConfig(bvh_width=2)
    return sm


@pytest.mark.parametrize("size", [(90, 45), (33, 17), (2080, 1040)], ids=lambda s: f"{s[0]}x{s[1]}")
def test_environment_gradient_splat_on_odd_sizes(size):
    # The wavefront adjoint adds the environment gradients of later escapes and
    # of environment shadow rays through per-tile LDS sums; an environment
    # whose sides are not multiples of the tile has partial tiles, and
    # bilinear taps that wrap round the seam fall outside their tile; one of
    # more than 4096 tiles (2080 x 1040) adds its taps directly.  The
    # gradients equal those of the per-pixel adjoint, which adds every tap
    # directly, and of the wavefront without the material sort (no splat).
    sm = _quads_under_sky(*size)
    rng = np.random.default_rng(8)
    g = torch.from_numpy(rng.uniform(-1, 1, (32, 48, 4)).astype(np.float32)).cuda()
    out = _both_forms(sm, 2, g)
    env = out["wavefront"]["env"]
    assert env is not None and float(env.abs().sum()) > 0
    touched = int((env.reshape(-1, 4)[:, :3] != 0).any(1).sum())
    assert touched > min(size[0] * size[1] // 4, 2000)  # spread over the map
    _assert_same_grads(out["path"], out["wavefront"], keys=("base_color", "roughness", "metallic", "env"))
    from pathtracingtutorial_amd.diff import DiffRenderer
    sm.setWavefrontConfig(sort_by_material=0)
    sm.setAdjointIntegrator(pt.WavefrontIntegrator())
    unsorted = DiffRenderer(sm).adjoint(2, g)
    _assert_same_grads(out["wavefront"], unsorted, keys=("base_color", "roughness", "metallic", "env"))
