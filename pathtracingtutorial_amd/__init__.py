"""MI355X path tracer: Python mirror of the reference renderer's entry points.

Classes keep the reference's names and methods (Part_2/src/scene/
scene_manager.h:11-45, renderer.h:17-55, integrators/integrator.h:10-20),
over the C ABI of libptamd.so:

    sm = SceneManager(); sm.load("scene.json"); sm.buildAccel(); sm.uploadToDevice()
    r = Renderer(sm, PathIntegrator()); r.renderAll(); r.saveImage("out.hdr")

Errors raise PtError (status code + message) instead of exiting the process.
"""
from __future__ import annotations

import ctypes as C
import math
import os
from dataclasses import dataclass

import numpy as np

from . import _capi
from ._capi import (BvhOptions, Camera, CameraFrame, Counters, PtError, SceneInfo, Vec3p, check, fptr, i32ptr,
                    lib, u32ptr)

__all__ = [
    "SceneManager", "Renderer", "Integrator", "PathIntegrator", "WavefrontIntegrator", "PtError", "QualityMetrics",
    "render_jobs",
    "compute_rmse", "compute_relmse", "probe_sampler", "probe_reciprocal", "trace_rays", "bvh_options", "REFERENCE_BVH", "lib",
]

# Reference builder settings (accel/bvh_builder.cpp:10-13).
REFERENCE_BVH = (24, 4, 0)


def bvh_options(bins: int = 24, leaf_threshold: int = 4, max_leaf: int = 0) -> BvhOptions:
    return BvhOptions(bins, leaf_threshold, max_leaf, 0)


class Integrator:
    """Integrator plugin (integrators/integrator.h:10-20): selects the kernel."""

    kind = _capi.PT_INTEGRATOR_PATH

    def name(self) -> str:
        raise NotImplementedError


class PathIntegrator(Integrator):
    """The plain megakernel (integrators/path_integrator.h:7-15)."""

    kind = _capi.PT_INTEGRATOR_PATH

    def name(self) -> str:
        return "Path"


class WavefrontIntegrator(Integrator):
    """Queue-based wavefront form of the same estimator."""

    kind = _capi.PT_INTEGRATOR_WAVEFRONT

    def name(self) -> str:
        return "Wavefront"


class SceneManager:
    """SceneManager (scene/scene_manager.h:11-45) over a pt_scene handle."""

    def __init__(self):
        self._h = C.c_void_p()
        self._output = "output.hdr"
        self._path = ""

    # -- construction -------------------------------------------------------
    def load(self, path: str) -> None:
        self._free()
        check(lib().pt_scene_load(os.fsencode(path), C.byref(self._h)))
        self._path = path

    def generate(self, which: int, seed: int = 0) -> None:
        """Procedural workloads: 0 Cornell box, 1 64 displaced icospheres, 2 small icospheres."""
        self._free()
        check(lib().pt_scene_generate(int(which), int(seed), C.byref(self._h)))

    def create(self, desc) -> None:
        self._free()
        check(lib().pt_scene_create(C.byref(desc), C.byref(self._h)))

    def buildAccel(self, bins: int | None = None, leaf_threshold: int | None = None, max_leaf: int | None = None):
        opts = None
        if bins is not None or leaf_threshold is not None or max_leaf is not None:
            opts = C.byref(bvh_options(bins or 24, leaf_threshold or 4, max_leaf or 0))
        check(lib().pt_scene_build_accel(self.handle, opts))

    def uploadToDevice(self) -> None:
        check(lib().pt_scene_upload(self.handle))

    # -- settings (scene_manager.h:32-36) ------------------------------------
    def setSpp(self, spp: int) -> None:
        check(lib().pt_scene_set_spp(self.handle, int(spp)))

    def setSamplePerFrame(self, spf: int) -> None:
        check(lib().pt_scene_set_sample_per_frame(self.handle, int(spf)))

    def setMaxDepth(self, depth: int) -> None:
        check(lib().pt_scene_set_max_depth(self.handle, int(depth)))

    def setSampler(self, sampler) -> None:
        if isinstance(sampler, str):
            sampler = {"pcg": 0, "sobol": 1}[sampler]
        check(lib().pt_scene_set_sampler(self.handle, int(sampler)))

    def setSeed(self, seed: int) -> None:
        check(lib().pt_scene_set_seed(self.handle, int(seed) & 0xFFFFFFFF))

    def setOutput(self, path: str) -> None:
        self._output = path

    def updateCamera(self, cam: Camera) -> None:
        check(lib().pt_scene_set_camera(self.handle, C.byref(cam)))

    # -- queries ----------------------------------------------------------------
    @property
    def handle(self) -> C.c_void_p:
        if not self._h:
            raise PtError(5, "no scene loaded")
        return self._h

    def info(self) -> SceneInfo:
        i = SceneInfo()
        check(lib().pt_scene_get_info(self.handle, C.byref(i)))
        return i

    def camera(self) -> Camera:
        c = Camera()
        check(lib().pt_scene_get_camera(self.handle, C.byref(c)))
        return c

    def cameraFrame(self) -> CameraFrame:
        f = CameraFrame()
        check(lib().pt_scene_camera_frame(self.handle, C.byref(f)))
        return f

    def outputPath(self) -> str:
        return self._output

    def scenePath(self) -> str:
        return self._path

    def numFaces(self) -> int:
        return self.info().num_faces

    def numBvhNodes(self) -> int:
        return self.info().num_bvh_nodes

    def numEmissiveTriangles(self) -> int:
        return self.info().num_emissive_triangles

    def bvh(self, layout: int = 0):
        """(nodes[N,8] float32 view with data words as uint32, face_order[F])."""
        nn, nf = C.c_uint32(0), C.c_uint32(0)
        check(lib().pt_scene_get_bvh(self.handle, layout, None, C.byref(nn), None, C.byref(nf)))
        nodes = np.zeros((nn.value, 8), np.float32)
        order = np.zeros(nf.value, np.uint32)
        check(lib().pt_scene_get_bvh(self.handle, layout, fptr(nodes), C.byref(nn), u32ptr(order), C.byref(nf)))
        return nodes, order

    def wideBvh(self):
        """(the scene's 4-wide BVH as built on the GPU, (n, 16) float32 with
        child words and quantized planes as uint32 bits; its stack bound)."""
        nw, ms = C.c_uint32(0), C.c_uint32(0)
        check(lib().pt_scene_get_wide_bvh(self.handle, None, C.byref(nw), C.byref(ms)))
        out = np.zeros((nw.value, 16), np.float32)
        check(lib().pt_scene_get_wide_bvh(self.handle, fptr(out), C.byref(nw), C.byref(ms)))
        return out, int(ms.value)

    def lightTables(self):
        a, b, c = C.c_uint32(0), C.c_uint32(0), C.c_uint32(0)
        check(lib().pt_scene_get_light_tables(self.handle, None, C.byref(a), None, C.byref(b), None, None, C.byref(c)))
        tp = np.zeros(a.value, np.float32)
        ta = np.zeros((b.value, 2), np.float32)
        ep = np.zeros(c.value, np.float32)
        ea = np.zeros((c.value, 2), np.float32)
        check(lib().pt_scene_get_light_tables(self.handle, fptr(tp), C.byref(a), fptr(ta), C.byref(b), fptr(ep),
                                              fptr(ea), C.byref(c)))
        return {"tri_pmf": tp, "tri_alias": ta, "env_pmf": ep, "env_alias": ea}

    def updateEmission(self, emissive=None, strength=None) -> None:
        """Sets the materials' emissive color (num_materials x 3) and/or strength
        (num_materials) of a built scene; the light tables keep their old
        distribution until rebuildEmissiveTables."""
        def arr(a, n):
            if a is None:
                return None
            a = np.ascontiguousarray(np.asarray(a, np.float32).reshape(-1))
            if a.size != n:
                raise ValueError(f"expected {n} values, got {a.size}")
            return a
        nm = self.info().num_materials
        e, s = arr(emissive, 3 * nm), arr(strength, nm)
        check(lib().pt_scene_update_emission(self.handle, fptr(e) if e is not None else None,
                                             fptr(s) if s is not None else None))

    def rebuildEmissiveTables(self, on_device: bool = True) -> None:
        """Emissive-triangle sampling tables for the current emission: built on
        the GPU (default; same distribution, another alias layout) or on the
        host (the reference's Vose table)."""
        if on_device:
            check(lib().pt_scene_build_emissive_tables_device(self.handle))
        else:
            check(lib().pt_scene_rebuild_emissive_tables(self.handle))

    def wavefrontConfig(self) -> dict:
        c = _capi.WavefrontConfig()
        check(lib().pt_scene_get_wavefront_config(self.handle, C.byref(c)))
        return {k: int(getattr(c, k)) for k, _ in _capi.WavefrontConfig._fields_}

    def setWavefrontConfig(self, **fields) -> None:
        """Changes fields of the wavefront launch configuration (speed only)."""
        cur = self.wavefrontConfig()
        for k, v in fields.items():
            if k not in cur:
                raise KeyError(f"unknown wavefront config field {k!r}")
            cur[k] = int(v)
        check(lib().pt_scene_set_wavefront_config(self.handle, C.byref(_capi.WavefrontConfig(**cur))))

    def tuneWavefront(self, max_trials: int = 16) -> tuple[dict, float]:
        """Autotunes the wavefront configuration; returns it and its ms per sample per pixel."""
        c = _capi.WavefrontConfig()
        ms = C.c_float(0.0)
        check(lib().pt_scene_tune_wavefront(self.handle, int(max_trials), C.byref(c), C.byref(ms)))
        return {k: int(getattr(c, k)) for k, _ in _capi.WavefrontConfig._fields_}, float(ms.value)

    def wavefrontStats(self) -> dict:
        """Per-stage GPU ms and launch counts of the last batch rendered with time_stages=1."""
        st = _capi.WavefrontStats()
        check(lib().pt_scene_wavefront_stats(self.handle, C.byref(st)))
        return {name: {"ms": float(st.stage_ms[i]), "launches": int(st.launches[i])}
                for i, name in enumerate(_capi.WAVEFRONT_STAGES)}

    def setAdjointIntegrator(self, integrator: Integrator) -> None:
        """Form of the path-replay adjoint: WavefrontIntegrator() (default) or PathIntegrator()."""
        check(lib().pt_scene_set_adjoint_integrator(self.handle, integrator.kind))

    def adjointIntegrator(self) -> str:
        k = C.c_int(0)
        check(lib().pt_scene_get_adjoint_integrator(self.handle, C.byref(k)))
        return "Wavefront" if k.value == _capi.PT_INTEGRATOR_WAVEFRONT else "Path"

    def enableCounters(self, on: bool = True) -> None:
        check(lib().pt_counters_enable(self.handle, 1 if on else 0))

    def resetCounters(self) -> None:
        check(lib().pt_counters_reset(self.handle))

    def counters(self) -> dict:
        c = Counters()
        check(lib().pt_counters_get(self.handle, C.byref(c)))
        return {k: int(getattr(c, k)) for k, _ in Counters._fields_}

    def _free(self):
        if self._h:
            lib().pt_scene_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self._free()
        except Exception:
            pass


def render_jobs(renderers) -> float:
    """One batch of every renderer's scene in one persistent wavefront schedule
    (pt_render_jobs); returns the GPU ms.  Each renderer's film and scene
    frame count advance as with renderBatch()."""
    arr = (_capi.Job * len(renderers))()
    for i, r in enumerate(renderers):
        arr[i].scene = r._scene.handle.value
        arr[i].film = r._film.value
    ms = C.c_float(0.0)
    check(lib().pt_render_jobs(arr, len(renderers), C.byref(ms)))
    for r in renderers:
        r._last_ms = ms.value / len(renderers)
        r._batches += 1
    return float(ms.value)


@dataclass
class QualityMetrics:
    """film/quality_metrics.h:13-18, plus relMSE."""
    spp: int = 0
    rmse: float = 0.0
    psnr: float = 0.0
    time_ms: float = 0.0
    relmse: float = 0.0


def compute_rmse(img: np.ndarray, ref: np.ndarray, channels: int = 3) -> float:
    """computeRMSE (film/quality_metrics.h:20-34), accumulated in float64."""
    d = img[..., :channels].astype(np.float64) - ref[..., :channels].astype(np.float64)
    return float(np.sqrt(np.mean(d * d))) if d.size else 0.0


def compute_relmse(img: np.ndarray, ref: np.ndarray, channels: int = 3, eps: float = 1e-2) -> float:
    """mean over pixels and channels of (I-R)^2 / (R^2 + eps)."""
    a = img[..., :channels].astype(np.float64)
    r = ref[..., :channels].astype(np.float64)
    return float(np.mean((a - r) ** 2 / (r * r + eps)))


def rmse_to_db(rmse: float) -> float:
    return 100.0 if rmse <= 0 else -20.0 * math.log10(rmse)


class Renderer:
    """Renderer (renderer.h:17-55): owns the integrator and the film."""

    def __init__(self, scene: SceneManager, integrator: Integrator | None = None):
        self._scene = scene
        self._integrator = integrator or PathIntegrator()
        self._film = C.c_void_p()
        self._total_ms = 0.0
        self._last_ms = 0.0
        self._batches = 0
        cam = scene.camera()
        check(lib().pt_film_create(cam.width, cam.height, C.byref(self._film)))
        self._w, self._h = cam.width, cam.height

    def renderBatch(self) -> int:
        ms = C.c_float(0.0)
        check(lib().pt_render_batch(self._scene.handle, self._film, self._integrator.kind, C.byref(ms)))
        self._last_ms = ms.value
        self._total_ms += ms.value
        self._batches += 1
        return self.currentSpp()

    def renderAll(self) -> float:
        while not self.isComplete():
            self.renderBatch()
        return self._total_ms

    def currentSpp(self) -> int:
        i = self._scene.info()
        return i.frame_count * i.sample_per_frame

    def targetSpp(self) -> int:
        return self._scene.info().target_samples

    def isComplete(self) -> bool:
        return self.currentSpp() >= self.targetSpp()

    def samplePerFrame(self) -> int:
        return self._scene.info().sample_per_frame

    def lastBatchMs(self) -> float:
        return self._last_ms

    def totalTraceMs(self) -> float:
        return self._total_ms

    def resetAccumulation(self) -> None:
        check(lib().pt_film_reset(self._film, self._scene.handle))
        self._total_ms = self._last_ms = 0.0
        self._batches = 0

    def resize(self, w: int, h: int) -> None:
        if (w, h) == (self._w, self._h):
            return
        cam = self._scene.camera()
        cam.width, cam.height = int(w), int(h)
        self._scene.updateCamera(cam)
        lib().pt_film_destroy(self._film)
        self._film = C.c_void_p()
        check(lib().pt_film_create(int(w), int(h), C.byref(self._film)))
        self._w, self._h = int(w), int(h)
        self.resetAccumulation()

    def filmWidth(self) -> int:
        return self._w

    def filmHeight(self) -> int:
        return self._h

    def downloadCurrentImage(self, which: int = 0) -> np.ndarray:
        """Running-mean RGBA float image, shape (H, W, 4), row 0 = bottom."""
        out = np.zeros((self._h, self._w, 4), np.float32)
        check(lib().pt_film_download(self._film, int(which), fptr(out)))
        return out

    def filmDevicePtr(self, which: int = 0) -> int:
        p = C.c_void_p()
        check(lib().pt_film_device_ptr(self._film, int(which), C.byref(p)))
        return int(p.value or 0)

    def saveImage(self, path: str) -> None:
        """ImageIO::write of the accumulated film (.hdr .pfm .png .jpg .jpeg .bmp .tga)."""
        from .imageio import write_image
        write_image(path, self.downloadCurrentImage())

    def runConvergenceBenchmark(self, ref: np.ndarray, checkpoints) -> list[QualityMetrics]:
        """Renderer::runConvergenceBenchmark (renderer.cu:138-173) with relMSE."""
        self.resetAccumulation()
        out = []
        k = 0
        elapsed = 0.0
        while k < len(checkpoints):
            spp = self.renderBatch()
            elapsed += self._last_ms
            if spp >= checkpoints[k]:
                img = self.downloadCurrentImage()
                r = compute_rmse(img, ref)
                out.append(QualityMetrics(spp, r, rmse_to_db(r), elapsed, compute_relmse(img, ref)))
                k += 1
        return out

    def __del__(self):
        try:
            if self._film:
                lib().pt_film_destroy(self._film)
        except Exception:
            pass


def probe_sampler(sampler: int, xyi: np.ndarray, dims: int, seed: int = 0):
    """Device sampler values: xyi is (n, 3) int32 of (x, y, index)."""
    xyi = np.ascontiguousarray(xyi, np.int32)
    n = xyi.shape[0]
    u = np.zeros((n, dims), np.uint32)
    f = np.zeros((n, dims), np.float32)
    check(lib().pt_probe_sampler(int(sampler), int(seed), i32ptr(xyi), n, int(dims), u32ptr(u), fptr(f)))
    return u, f


def trace_rays(scene: SceneManager, rays: np.ndarray, any_hit: bool = False, wide: bool = False):
    """Device closest/any hit for (n, 8) float32 rays {o, d, tmax, _}; wide
    traverses the 4-wide BVH instead of the binary one."""
    rays = np.ascontiguousarray(rays, np.float32)
    n = rays.shape[0]
    t = np.zeros(n, np.float32)
    face = np.zeros(n, np.int32)
    bary = np.zeros((n, 2), np.float32)
    flags = (1 if any_hit else 0) | (2 if wide else 0)
    check(lib().pt_trace_rays(scene.handle, fptr(rays), n, flags, fptr(t), i32ptr(face), fptr(bary)))
    return t, face, bary


def collapse_wide(nodes8: np.ndarray):
    """4-wide BVH (n, 16) float32 (child words and quantized planes as uint32
    bits) and its traversal stack bound from a binary BVH in the reference's
    node layout ((m, 8) float32).  Host only."""
    nodes8 = np.ascontiguousarray(nodes8, np.float32)
    nw, ms = C.c_uint32(0), C.c_uint32(0)
    check(lib().pt_bvh_collapse_wide(fptr(nodes8), nodes8.shape[0], None, C.byref(nw), C.byref(ms)))
    out = np.zeros((nw.value, 16), np.float32)
    check(lib().pt_bvh_collapse_wide(fptr(nodes8), nodes8.shape[0], fptr(out), C.byref(nw), C.byref(ms)))
    return out, int(ms.value)


def probe_copy(nbytes: int = 1 << 30, iters: int = 10) -> float:
    """Achieved GB/s (read + write) of a plain device copy kernel."""
    g = C.c_float(0.0)
    check(lib().pt_probe_copy(int(nbytes), int(iters), C.byref(g)))
    return float(g.value)


def probe_reciprocal() -> tuple[int, int]:
    """(mismatches, compared): the triangle test's reciprocal against IEEE
    division over every float of magnitude in [2^-126, 2^126)."""
    bad, n = C.c_uint64(0), C.c_uint64(0)
    check(lib().pt_probe_reciprocal(C.byref(bad), C.byref(n)))
    return int(bad.value), int(n.value)
