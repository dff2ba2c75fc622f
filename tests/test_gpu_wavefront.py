"""Wavefront integrator vs the megakernel and the oracle, through the C ABI.

The wavefront schedule runs the megakernel's estimator with the same random
numbers per path, so without punctual lights (whose shadow contributions are
added with float atomics) the images agree pixel by pixel up to floating
point contraction differences between the two compiled kernels; every
wavefront launch configuration gives bit-identical images (per-path work and
per-pixel summation order do not depend on the schedule)."""
import json
import os

import numpy as np
import pytest

import oracle
import pathtracingtutorial_amd as pt
from test_gpu_parity import _zscore_check, host_view

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
CORNELL = os.path.join(HERE, "..", "scenes", "cornell_box.json")


def _cornell(w=96, h=96, spf=8, depth=8, sampler=None):
    sm = pt.SceneManager()
    sm.generate(0)
    cam = sm.camera()
    cam.width, cam.height = w, h
    sm.updateCamera(cam)
    sm.setSamplePerFrame(spf)
    sm.setMaxDepth(depth)
    if sampler is not None:
        sm.setSampler(sampler)
    sm.buildAccel()
    sm.uploadToDevice()
    return sm


def _env_scene(w=80, h=48, spf=8, depth=6):
    sm = pt.SceneManager()
    sm.generate(2, 5)
    cam = sm.camera()
    cam.width, cam.height = w, h
    sm.updateCamera(cam)
    sm.setSamplePerFrame(spf)
    sm.setMaxDepth(depth)
    sm.buildAccel(24, 4, 0)
    sm.uploadToDevice()
    return sm


def _render(sm, integ, batches=1):
    r = pt.Renderer(sm, integ)
    news = []
    for _ in range(batches):
        r.renderBatch()
        news.append(r.downloadCurrentImage(1).copy())
    return np.array(news), r.downloadCurrentImage(0).copy()


def _same_paths(a, b, tol_frac=0.01):
    """Images made from the same random numbers: nearly all pixels equal to
    float rounding; a few may differ where a rounding difference changed a
    path's branch."""
    a = a[..., :3].astype(np.float64)
    b = b[..., :3].astype(np.float64)
    close = np.abs(a - b) <= 1e-4 * (1.0 + np.abs(b))
    frac_bad = 1.0 - close.mean()
    assert frac_bad <= tol_frac, f"{frac_bad:.3%} of pixel channels differ"
    return frac_bad


def _bit_identical(a, b, label):
    """The two integrators trace the same paths (same tree, same random
    numbers), round every radiance term alike (no multiply-add contraction
    in the shading code) and sum a path's terms and a pixel's samples in the
    same order: the images agree to the bit."""
    d = np.abs(a[..., :3].astype(np.float64) - b[..., :3].astype(np.float64))
    assert np.array_equal(a[..., :3], b[..., :3]), f"{label}: {int((d > 0).sum())} channels differ, max {d.max():.3g}"


def test_wavefront_matches_megakernel_cornell():
    sm = _cornell()
    mk, mk_acc = _render(sm, pt.PathIntegrator(), 2)
    sm.uploadToDevice()  # frame_count back to 0: the same sample indices
    wf, wf_acc = _render(sm, pt.WavefrontIntegrator(), 2)
    assert np.isfinite(wf).all() and wf[..., :3].max() <= 100.0
    assert np.all(wf[..., 3] == 1.0)
    _bit_identical(wf, mk, "cornell batches")
    _bit_identical(wf_acc, mk_acc, "cornell running mean")


@pytest.mark.parametrize("width", [2, 4])
def test_wavefront_matches_megakernel_env_sobol(width):
    sm = _env_scene()
    sm.setWavefrontConfig(bvh_width=width)
    sm.uploadToDevice()
    mk, _ = _render(sm, pt.PathIntegrator(), 2)
    sm.uploadToDevice()
    wf, _ = _render(sm, pt.WavefrontIntegrator(), 2)
    assert np.isfinite(wf).all()
    _bit_identical(wf, mk, f"env scene, {width}-wide tree")


def test_wavefront_matches_oracle():
    sm = _cornell(w=64, h=64, spf=4)
    ora = oracle.OracleScene.from_view(host_view(sm))
    cam = sm.camera()
    info = sm.info()
    r = pt.Renderer(sm, pt.WavefrontIntegrator())
    gpu, cpu = [], []
    for b in range(10):
        r.renderBatch()
        gpu.append(r.downloadCurrentImage(1)[..., :3].astype(np.float64))
        img, _ = ora.render(cam, 4, info.max_depth, info.sampler, info.seed, frame=b)
        cpu.append(img.astype(np.float64))
    _zscore_check(np.array(gpu), np.array(cpu))


CONFIGS = [
    dict(chunk_spp=1),
    dict(chunk_spp=3),  # not a divisor of 8: the largest divisor below (2) is used
    dict(chunk_spp=64),  # capped at sample_per_frame
    dict(sort_by_material=0),
    dict(use_graph=0),
    dict(extend_block=64, shade_block=256),
    dict(extend_block=256, shade_block=64),
    dict(waves_per_simd=1),
    dict(stack_lds=16),
    dict(stack_lds=16, extend_block=256, waves_per_simd=2),
    dict(grab_extend=64, grab_shade=64, grab_shadow=64),
    dict(grab_extend=1024, grab_shade=1024, grab_shadow=1024, chunk_spp=2),
]


def test_wavefront_configs_are_bit_identical():
    sm = _env_scene(w=72, h=40, spf=8)
    base, base_acc = _render(sm, pt.WavefrontIntegrator(), 2)
    default = sm.wavefrontConfig()
    for cfg in CONFIGS:
        sm.setWavefrontConfig(**default)
        sm.setWavefrontConfig(**cfg)
        sm.uploadToDevice()
        img, acc = _render(sm, pt.WavefrontIntegrator(), 2)
        assert np.array_equal(img, base), f"config {cfg} changed the image"
        assert np.array_equal(acc, base_acc), f"config {cfg} changed the accumulation"


def test_wavefront_pcg_cornell_configs_bit_identical():
    sm = _cornell(w=48, h=40, spf=6)
    base, _ = _render(sm, pt.WavefrontIntegrator(), 1)
    for cfg in (dict(chunk_spp=2), dict(sort_by_material=0, use_graph=0)):
        sm.setWavefrontConfig(**cfg)
        sm.uploadToDevice()
        img, _ = _render(sm, pt.WavefrontIntegrator(), 1)
        assert np.array_equal(img, base), f"config {cfg} changed the image"


def test_wavefront_counters_match_megakernel():
    sm = _cornell(w=64, h=64, spf=2)
    sm.setWavefrontConfig(bvh_width=2)  # the megakernel's tree
    sm.enableCounters(True)
    sm.resetCounters()
    _render(sm, pt.PathIntegrator(), 1)
    a = sm.counters()
    sm.resetCounters()
    sm.uploadToDevice()
    _render(sm, pt.WavefrontIntegrator(), 1)
    b = sm.counters()
    assert b["paths"] == a["paths"] == 64 * 64 * 2
    for k in ("closest_rays", "shadow_rays"):
        assert abs(a[k] - b[k]) <= 0.01 * a[k] + 10, (k, a[k], b[k])
    # the megakernel's closest-hit stack keeps entry distances and drops a popped
    # node beyond the hit found since (as the reference does at its pop); the
    # wavefront's does not: the same rays, fewer node and triangle fetches
    for k in ("node_visits", "tri_tests"):
        assert 0.8 * b[k] <= a[k] <= b[k] + 10, (k, a[k], b[k])
    assert b["stack_overflows"] == 0


@pytest.mark.parametrize("scene", ["cornell", "env"])
def test_wavefront_wide_bvh_same_paths_fewer_node_bytes(scene):
    # The 4-wide BVH gives the binary tree's closest hits (up to exact-distance
    # ties), so the images agree pixel by pixel with the same random numbers,
    # every ray is traced alike, and the node data fetched per ray drops.
    sm = _cornell(w=64, h=64, spf=4) if scene == "cornell" else _env_scene(w=72, h=40, spf=8)
    sm.enableCounters(True)
    out = {}
    for width in (2, 4):
        sm.setWavefrontConfig(bvh_width=width)
        sm.uploadToDevice()
        sm.resetCounters()
        out[width] = _render(sm, pt.WavefrontIntegrator(), 2)[0], sm.counters()
    (a, ca), (b, cb) = out[2], out[4]
    bad = _same_paths(b, a, tol_frac=0.002)
    print(f"{scene}: {bad:.4%} of channels differ between the binary and the 4-wide BVH")
    for k in ("closest_rays", "shadow_rays", "paths"):
        assert abs(ca[k] - cb[k]) <= 0.002 * ca[k] + 2, (k, ca[k], cb[k])
    assert cb["node_visits"] < 0.9 * ca["node_visits"], (ca["node_visits"], cb["node_visits"])
    assert cb["stack_overflows"] == 0


def test_wavefront_depth_limits():
    # maxDepth 0 renders black; maxDepth 1 sees only emitters and direct light
    for depth in (0, 1):
        sm = _cornell(w=32, h=32, spf=2, depth=depth)
        mk, _ = _render(sm, pt.PathIntegrator(), 1)
        sm.uploadToDevice()
        wf, _ = _render(sm, pt.WavefrontIntegrator(), 1)
        if depth == 0:
            assert np.all(wf[..., :3] == 0.0)
        _bit_identical(wf, mk, f"depth {depth}")


def test_wavefront_punctual_lights(tmp_path):
    scene = json.load(open(CORNELL))
    scene["camera"]["resolution"] = [48, 48]
    scene["render"]["samplePerFrame"] = 4
    scene["render"]["spp"] = 4
    scene["lights"] = [
        {"type": "point", "position": [0.2, 0.3, 0.1], "color": [1, 0.9, 0.8], "intensity": 0.5},
        {"type": "spot", "position": [-0.2, 0.4, 0.2], "direction": [0.1, -1, -0.1], "color": [0.3, 0.6, 1.0],
         "intensity": 2.0, "innerCone": 0.3, "outerCone": 0.6},
        {"type": "directional", "direction": [0.3, -1.0, -0.4], "color": [1, 1, 1], "intensity": 0.2},
    ]
    path = tmp_path / "cornell_lights.json"
    path.write_text(json.dumps(scene))
    sm = pt.SceneManager()
    sm.load(str(path))
    sm.buildAccel()
    sm.uploadToDevice()
    assert sm.info().num_lights == 3
    mk, _ = _render(sm, pt.PathIntegrator(), 2)
    sm.uploadToDevice()
    wf, _ = _render(sm, pt.WavefrontIntegrator(), 2)
    # per-light contributions are summed in another order: equal to 1e-4
    _same_paths(wf, mk, tol_frac=0.02)
    ora = oracle.OracleScene.from_view(host_view(sm))
    cam, info = sm.camera(), sm.info()
    sm.uploadToDevice()
    r = pt.Renderer(sm, pt.WavefrontIntegrator())
    gpu, cpu = [], []
    for b in range(8):
        r.renderBatch()
        gpu.append(r.downloadCurrentImage(1)[..., :3].astype(np.float64))
        img, _ = ora.render(cam, 4, info.max_depth, info.sampler, info.seed, frame=b)
        cpu.append(img.astype(np.float64))
    _zscore_check(np.array(gpu), np.array(cpu), frac=0.01)


def test_wavefront_autotune_keeps_image():
    sm = _env_scene(w=64, h=32, spf=8)
    base, _ = _render(sm, pt.WavefrontIntegrator(), 1)
    best, ms = sm.tuneWavefront(max_trials=6)
    assert ms > 0.0
    assert best == sm.wavefrontConfig()
    assert best["extend_block"] in (64, 128, 256) and best["shade_block"] in (64, 128, 256)
    sm.uploadToDevice()
    img, _ = _render(sm, pt.WavefrontIntegrator(), 1)
    assert np.array_equal(img, base)


def test_wavefront_bad_config_rejected():
    sm = _cornell(w=16, h=16, spf=1)
    with pytest.raises(pt.PtError):
        sm.setWavefrontConfig(extend_block=100)
    with pytest.raises(pt.PtError):
        sm.setWavefrontConfig(chunk_spp=0)
    with pytest.raises(pt.PtError):
        sm.setWavefrontConfig(stack_lds=8)
    with pytest.raises(pt.PtError):
        sm.setWavefrontConfig(grab_shade=96)
    with pytest.raises(pt.PtError):
        sm.setWavefrontConfig(bvh_width=8)


@pytest.mark.parametrize("width", [2, 4])
def test_wavefront_split_stack_deep_bvh_bit_identical(width):
    # One triangle per leaf over 2M triangles: a BVH deeper than the 16 LDS
    # stack entries, so traversal stacks spill into device memory.
    sm = pt.SceneManager()
    sm.generate(5)
    cam = sm.camera()
    cam.width, cam.height = 96, 64
    sm.updateCamera(cam)
    sm.setSamplePerFrame(2)
    sm.buildAccel(24, 1, 1)
    assert sm.info().bvh_depth > 20
    sm.uploadToDevice()
    sm.enableCounters(True)
    out = {}
    for lds in (32, 16):
        sm.setWavefrontConfig(stack_lds=lds, bvh_width=width)
        sm.uploadToDevice()
        sm.resetCounters()
        out[lds] = _render(sm, pt.WavefrontIntegrator(), 1)[0], sm.counters()
    assert np.array_equal(out[16][0], out[32][0])
    for k in ("node_visits", "tri_tests", "shadow_node_visits", "shadow_tri_tests", "stack_overflows"):
        assert out[16][1][k] == out[32][1][k], k


@pytest.mark.parametrize("which", ["env", "cornell"])
def test_tree_rebuilt_after_upload_renders_like_a_fresh_scene(which):
    # a rebuilt tree reorders the triangles: the scene must be uploaded again
    # (rendering before that is refused), and the upload rebuilds everything
    # indexed in tree order (shading records, emissive-triangle tables)
    make = _env_scene if which == "env" else _cornell
    a = make()
    a.buildAccel(32, 4, 4)
    with pytest.raises(pt.PtError):
        _render(a, pt.WavefrontIntegrator())
    a.uploadToDevice()
    b = pt.SceneManager()
    b.generate(2, 5) if which == "env" else b.generate(0)
    b.updateCamera(a.camera())
    b.setSamplePerFrame(8)
    b.setMaxDepth(6 if which == "env" else 8)
    b.buildAccel(32, 4, 4)
    b.uploadToDevice()
    for integ in (pt.WavefrontIntegrator, pt.PathIntegrator):
        ia, _ = _render(a, integ())
        ib, _ = _render(b, integ())
        assert np.isfinite(ia).all() and ia.mean() > 0
        assert np.array_equal(ia, ib), integ.__name__
