"""Many scenes in one wavefront schedule (pt_render_jobs).

Each job's film must be exactly the film its scene gets from a wavefront
batch of its own (same kernels, same random numbers, samples summed in the
same order), and agree with the megakernel to float rounding."""
import numpy as np
import pytest

import pathtracingtutorial_amd as pt
from test_gpu_wavefront import _same_paths

pytestmark = pytest.mark.gpu


def _scene(kind, w, h, spf, depth, seed=0, bvh=None):
    sm = pt.SceneManager()
    sm.generate(kind, seed)
    cam = sm.camera()
    cam.width, cam.height = w, h
    sm.updateCamera(cam)
    sm.setSamplePerFrame(spf)
    sm.setMaxDepth(depth)
    if bvh:
        sm.buildAccel(*bvh)
    else:
        sm.buildAccel()
    sm.uploadToDevice()
    return sm


def _jobs():
    return [
        _scene(0, 64, 48, 4, 8),                       # Cornell, PCG, area light
        _scene(2, 40, 32, 8, 6, seed=5, bvh=(24, 4, 0)),  # icospheres + sky, Sobol
        _scene(0, 33, 17, 3, 2),                       # odd size, shallow
        _scene(3, 24, 24, 2, 6, seed=3),               # textured diff Cornell + env
    ]


def test_jobs_equal_single_scene_batches():
    scenes = _jobs()
    singles = []
    for sm in scenes:
        r = pt.Renderer(sm, pt.WavefrontIntegrator())
        r.renderBatch()
        r.renderBatch()
        singles.append((r.downloadCurrentImage(1).copy(), r.downloadCurrentImage(0).copy()))
        sm.uploadToDevice()  # frame_count back to 0
    rs = [pt.Renderer(sm, pt.WavefrontIntegrator()) for sm in scenes]
    for _ in range(2):
        ms = pt.render_jobs(rs)
        assert ms > 0
    for sm, r, (new, acc) in zip(scenes, rs, singles):
        assert sm.info().frame_count == 2
        assert np.array_equal(r.downloadCurrentImage(1), new)
        assert np.array_equal(r.downloadCurrentImage(0), acc)


def test_jobs_match_megakernel():
    scenes = _jobs()
    mk = []
    for sm in scenes:
        r = pt.Renderer(sm, pt.PathIntegrator())
        r.renderBatch()
        mk.append(r.downloadCurrentImage(1).copy())
        sm.uploadToDevice()
    rs = [pt.Renderer(sm, pt.WavefrontIntegrator()) for sm in scenes]
    pt.render_jobs(rs)
    for r, ref in zip(rs, mk):
        _same_paths(r.downloadCurrentImage(1), ref, tol_frac=0.02)


def test_jobs_chunking_and_counters():
    scenes = _jobs()
    base = None
    for chunk in (1, 3, 64):
        for sm in scenes:
            sm.uploadToDevice()
        scenes[0].setWavefrontConfig(chunk_spp=chunk)
        scenes[0].enableCounters(True)
        scenes[0].resetCounters()
        rs = [pt.Renderer(sm, pt.WavefrontIntegrator()) for sm in scenes]
        pt.render_jobs(rs)
        c = scenes[0].counters()
        imgs = [r.downloadCurrentImage(1).copy() for r in rs]
        expect = sum(i.width * i.height * i.sample_per_frame for i in (sm.info() for sm in scenes))
        assert c["paths"] == expect
        assert c["closest_rays"] >= expect - sum(0 for _ in scenes)
        if base is None:
            base = imgs
        else:
            for a, b in zip(imgs, base):
                assert np.array_equal(a, b)


def test_jobs_reject_shared_film_and_bad_input():
    sm = _scene(0, 16, 16, 1, 2)
    r = pt.Renderer(sm, pt.WavefrontIntegrator())
    with pytest.raises(pt.PtError):
        pt.render_jobs([r, r])
