"""Light-selection tables built on the GPU (pt_alias_build_device,
pt_scene_build_env_tables_device).

An alias table samples bucket i with probability 1/n and keeps it when u2 <
prob[i], else takes alias[i] (direct_lighting.cuh:55-59), so item j is drawn
with probability (prob[j] + sum over buckets i aliased to j of (1 - prob[i])) / n.
That implied distribution is computed here from the table and must equal the
weights' pmf for every weight set, and the environment tables built from the
device texels must match the host's pmf; an environment scene renders the
same image (within 3 sigma, independent samples) with either table."""
import ctypes as C

import numpy as np
import pytest

import pathtracingtutorial_amd as pt
from pathtracingtutorial_amd import _capi

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _implied(table):
    n = table.shape[0]
    prob = table[:, 0].astype(np.float64)
    alias = table[:, 1].copy().view(np.uint32)
    assert np.all((prob >= 0) & (prob <= 1)) and np.all(alias < n)
    mass = prob.copy()
    np.add.at(mass, alias, 1.0 - prob)
    return mass / n


def _weights(case, rng):
    if case == "uniform":
        return np.ones(1000, np.float32)
    if case == "random":
        return rng.random(100003).astype(np.float32)
    if case == "spiky":  # a few huge weights among many tiny ones
        w = rng.random(50000).astype(np.float32) * 1e-3
        w[rng.integers(0, 50000, 5)] = 1e4
        return w
    if case == "zeros":
        w = rng.random(4097).astype(np.float32)
        w[rng.random(4097) < 0.6] = 0.0
        return w
    if case == "heavy-tail":
        return (rng.pareto(1.2, 300000) + 1e-6).astype(np.float32)
    if case == "single":
        return np.array([3.0], np.float32)
    return np.array([1.0, 3.0], np.float32)


@pytest.mark.parametrize("case", ["uniform", "random", "spiky", "zeros", "heavy-tail", "single", "two"])
def test_alias_table_samples_the_pmf(case):
    w = _weights(case, np.random.default_rng(7))
    n = w.size
    wd = torch.from_numpy(w).cuda()
    pmf = torch.empty(n, dtype=torch.float32, device="cuda")
    alias = torch.empty((n, 2), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    _capi.check(pt.lib().pt_alias_build_device(C.c_void_p(wd.data_ptr()), n, C.c_void_p(pmf.data_ptr()),
                                               C.c_void_p(alias.data_ptr())))
    exact = w.astype(np.float64) / w.astype(np.float64).sum()
    assert np.allclose(pmf.cpu().numpy(), exact, rtol=1e-6, atol=1e-12)
    implied = _implied(alias.cpu().numpy())
    assert np.allclose(implied, exact, rtol=1e-5, atol=1e-9)
    assert np.all(implied[exact == 0] == 0)  # zero-weight items are never drawn


def test_alias_build_rejects_all_zero():
    wd = torch.zeros(16, dtype=torch.float32, device="cuda")
    pmf = torch.empty(16, device="cuda")
    alias = torch.empty((16, 2), device="cuda")
    with pytest.raises(pt.PtError):
        _capi.check(pt.lib().pt_alias_build_device(C.c_void_p(wd.data_ptr()), 16, C.c_void_p(pmf.data_ptr()),
                                                   C.c_void_p(alias.data_ptr())))


def _env_scene():
    sm = pt.SceneManager()
    sm.generate(2, 5)
    cam = sm.camera()
    cam.width, cam.height = 64, 40
    sm.updateCamera(cam)
    sm.setSamplePerFrame(8)
    sm.setMaxDepth(5)
    sm.buildAccel(24, 4, 0)
    sm.uploadToDevice()
    return sm


def _mean_image(sm, batches):
    r = pt.Renderer(sm, pt.WavefrontIntegrator())
    imgs = []
    for _ in range(batches):
        r.renderBatch()
        imgs.append(r.downloadCurrentImage(1)[..., :3].astype(np.float64))
    return np.array(imgs)


def _exact_env_pmf(sm):
    """lum * sin(theta) weights (light_sampler.cpp:112-149) normalized in double precision."""
    v = _capi.HostSceneView()
    _capi.check(pt.lib().pt_scene_host_view(sm.handle, C.byref(v)))
    h, w = v.env_height, v.env_width
    env = np.ctypeslib.as_array(v.env_rgba, shape=(h * w * 4,)).reshape(h, w, 4).astype(np.float64)
    st = np.maximum(np.sin(np.pi * (np.arange(h) + 0.5) / h), 0.0)[:, None]
    wts = np.maximum(0.2126 * env[..., 0] + 0.7152 * env[..., 1] + 0.0722 * env[..., 2], 0.0) * st
    return (wts / wts.sum()).reshape(-1)


def test_env_tables_on_device_match_host_and_render_alike():
    sm = _env_scene()
    host = sm.lightTables()
    assert host["env_pmf"].size > 0
    exact = _exact_env_pmf(sm)
    # the host table (like the reference) sums the weights in single precision
    assert np.allclose(host["env_pmf"], exact, rtol=1e-3, atol=1e-12)
    _capi.check(pt.lib().pt_scene_build_env_tables_device(sm.handle))
    dev = sm.lightTables()
    assert np.allclose(dev["env_pmf"], exact, rtol=1e-5, atol=1e-12)
    assert np.allclose(_implied(dev["env_alias"]), exact, rtol=1e-4, atol=1e-10)
    assert not np.array_equal(dev["env_alias"], host["env_alias"])  # another layout
    # images: device tables, then the host tables again (independent estimates of the same image)
    a = _mean_image(sm, 12)
    _capi.check(pt.lib().pt_scene_rebuild_env_tables(sm.handle))
    assert np.array_equal(sm.lightTables()["env_alias"], host["env_alias"])
    b = _mean_image(sm, 12)
    var = (a.var(0, ddof=1) + b.var(0, ddof=1)) / a.shape[0]
    d = np.abs(a.mean(0) - b.mean(0))
    z = np.where(var > 0, d / np.sqrt(np.maximum(var, 1e-30)), np.where(d > 1e-5, np.inf, 0.0))
    assert (z > 3.5).mean() <= 0.01
    assert abs(a.mean() - b.mean()) <= 3 * np.sqrt(var.mean() / var.size) + 1e-3 * b.mean()


def _emissive_exact(sm, emissive, strength):
    """area * luminance(emissive * strength) weights (light_sampler.cpp:67-81) of the
    BVH-ordered faces, normalized in double precision."""
    v = _capi.HostSceneView()
    _capi.check(pt.lib().pt_scene_host_view(sm.handle, C.byref(v)))
    verts = np.ctypeslib.as_array(C.cast(v.vertices, C.POINTER(C.c_float)),
                                  shape=(v.num_vertices * C.sizeof(_capi.Vertex) // 4,))
    verts = verts.reshape(v.num_vertices, -1)[:, :3].astype(np.float64)
    faces = np.ctypeslib.as_array(C.cast(v.faces, C.POINTER(C.c_int32)), shape=(v.num_faces * 4,)).reshape(-1, 4)
    _, order = sm.bvh(0)
    f = faces[order]
    a, b, c = verts[f[:, 0]], verts[f[:, 1]], verts[f[:, 2]]
    area = 0.5 * np.linalg.norm(np.cross(b - a, c - a), axis=1)
    e = emissive.astype(np.float64) * strength.astype(np.float64)[:, None]
    lum = 0.2126 * e[:, 0] + 0.7152 * e[:, 1] + 0.0722 * e[:, 2]
    emits = (strength > 0) & np.any(emissive > 0, axis=1)
    w = np.where(emits[f[:, 3]], area * np.maximum(lum[f[:, 3]], 0.0), 0.0)
    return w / w.sum()


def test_emissive_tables_on_device_match_host_and_render_alike():
    sm = _env_scene()
    nm = sm.info().num_materials
    assert nm >= 3
    rng = np.random.default_rng(11)
    em = (rng.random((nm, 3)) * 4).astype(np.float32)
    em[::2] = 0.0  # every other material stays dark
    st = (rng.random(nm) + 0.5).astype(np.float32)
    sm.updateEmission(em, st)
    sm.rebuildEmissiveTables(on_device=False)
    host = sm.lightTables()
    n_host = sm.info().num_emissive_triangles
    assert n_host > 100
    exact = _emissive_exact(sm, em, st)
    assert np.allclose(host["tri_pmf"], exact, rtol=1e-3, atol=1e-12)
    sm.rebuildEmissiveTables(on_device=True)
    dev = sm.lightTables()
    assert sm.info().num_emissive_triangles == n_host
    # the same emissive faces, in face order; the same pmf up to rounding
    assert np.array_equal(dev["tri_pmf"] > 0, exact > 0)
    assert np.allclose(dev["tri_pmf"], exact, rtol=1e-4, atol=1e-12)
    sel = dev["tri_pmf"][dev["tri_pmf"] > 0].astype(np.float64)
    assert np.allclose(_implied(dev["tri_alias"]), sel, rtol=1e-4, atol=1e-10)
    a = _mean_image(sm, 12)
    sm.rebuildEmissiveTables(on_device=False)
    assert np.array_equal(sm.lightTables()["tri_alias"], host["tri_alias"])
    b = _mean_image(sm, 12)
    var = (a.var(0, ddof=1) + b.var(0, ddof=1)) / a.shape[0]
    d = np.abs(a.mean(0) - b.mean(0))
    z = np.where(var > 0, d / np.sqrt(np.maximum(var, 1e-30)), np.where(d > 1e-5, np.inf, 0.0))
    assert (z > 3.5).mean() <= 0.01
    assert abs(a.mean() - b.mean()) <= 3 * np.sqrt(var.mean() / var.size) + 1e-3 * b.mean()


def test_emissive_tables_on_device_follow_emission_changes():
    sm = pt.SceneManager()
    sm.generate(0)
    sm.setSamplePerFrame(4)
    sm.buildAccel()
    sm.uploadToDevice()
    host = sm.lightTables()
    n0 = sm.info().num_emissive_triangles
    assert n0 == 2
    sm.rebuildEmissiveTables(on_device=True)
    dev = sm.lightTables()
    assert np.array_equal(dev["tri_pmf"] > 0, host["tri_pmf"] > 0)
    assert np.allclose(dev["tri_pmf"], host["tri_pmf"], rtol=1e-6)
    assert np.allclose(_implied(dev["tri_alias"]), [0.5, 0.5], rtol=1e-6)
    # every emission off: no emissive faces, and the render still runs (no area lights)
    nm = sm.info().num_materials
    sm.updateEmission(strength=np.zeros(nm, np.float32))
    sm.rebuildEmissiveTables(on_device=True)
    assert sm.info().num_emissive_triangles == 0
    assert not sm.lightTables()["tri_pmf"].any()
    r = pt.Renderer(sm, pt.WavefrontIntegrator())
    r.renderBatch()
    img = r.downloadCurrentImage(1)
    assert np.isfinite(img).all()
    # back on: the two faces return
    sm.updateEmission(strength=np.ones(nm, np.float32))
    sm.rebuildEmissiveTables(on_device=True)
    assert sm.info().num_emissive_triangles == n0
    assert np.allclose(sm.lightTables()["tri_pmf"], host["tri_pmf"], rtol=1e-6)
