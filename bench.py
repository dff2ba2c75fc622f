"""Benchmark of the path-tracing hot path on one MI355X.

Workload: 64 displaced level-5 icospheres (1,310,720 triangles) + ground,
64 random Disney materials, 2048x1024 sky+sun environment, 1920x1080,
64 spp, depth 16, Owen-scrambled Sobol, BVH with 32 bins and <= 4 triangles
per leaf.  One step = one 64-spp frame through the integrator.

Prints one JSON line:
  value      traced rays (closest + shadow) per second, Mrays/s, whole job
  roofline   algorithmic bytes of the render kernel over its HIP-event time
             (32 B per node record tested + 36 B per triangle test + 48 B per
             ray) against the 8 TB/s HBM3E peak
  cpu_baseline  the oracle (a C++ restatement of the reference integrator)
             on a bounded sample of the same frame, all host cores

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--integrator path]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
WORKLOAD = "spheres-64x20480tri-1080p-64spp-depth16-sobol-env2048"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic_bytes(c: dict) -> int:
    rays = c["closest_rays"] + c["shadow_rays"]
    return 32 * c["node_visits"] + 36 * c["tri_tests"] + 48 * rays


def profiled_traffic(kernel: str, workload: str):
    """HBM bytes per launch of `kernel` on `workload` from the committed rocprofv3
    summaries (profiles/*.md, FETCH_SIZE + WRITE_SIZE runs scaled by the copy-kernel
    calibration), or None when no summary covers this kernel and workload."""
    import glob
    import re
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*.md"))):
        for block in re.findall(r"```json\n(.*?)\n```", open(path).read(), re.S):
            try:
                d = json.loads(block)
            except ValueError:
                continue
            if d.get("kernel") == kernel and d.get("workload") == workload and "traffic_bytes_per_launch" in d:
                return int(d["traffic_bytes_per_launch"])
    return None


def cpu_baseline(sm, target_s: float = 12.0) -> dict:
    """The oracle on a bounded sample: whole rows of the same frame, 1 spp."""
    import ctypes as C

    import oracle
    import pathtracingtutorial_amd as pt
    from pathtracingtutorial_amd import _capi
    v = _capi.HostSceneView()
    _capi.check(pt.lib().pt_scene_host_view(sm.handle, C.byref(v)))
    t0 = time.perf_counter()
    ora = oracle.OracleScene.from_view(v)
    build_s = time.perf_counter() - t0
    cam = v.camera
    info = sm.info()
    cores = os.cpu_count() or 1
    # calibrate on a 16-row strip at 1 spp, then render the whole frame at the
    # spp that gives ~target_s of work (frame 0 of the same sampler sequence)
    y0 = cam.height // 2
    t0 = time.perf_counter()
    _, c = ora.render(cam, 1, info.max_depth, info.sampler, 0, 0, rect=(0, cam.width, y0, y0 + 16), threads=cores)
    per_row = (time.perf_counter() - t0) / 16
    spp = int(max(1, min(info.sample_per_frame, target_s / max(per_row * cam.height, 1e-3))))
    t0 = time.perf_counter()
    _, c = ora.render(cam, spp, info.max_depth, info.sampler, 0, 0, threads=cores)
    dt = time.perf_counter() - t0
    rays = c["closest_rays"] + c["shadow_rays"]
    return {"value": rays / dt / 1e6, "unit": "Mrays/s", "cores": cores, "kind": "port",
            "sample": f"the full {cam.width}x{cam.height} frame at {spp} spp ({c['paths']} paths, {rays} rays, "
                      f"{dt:.1f} s); oracle BVH build {build_s:.1f} s"}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--integrator", default="path", choices=["path", "wavefront"])
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run the plain copy kernel (HBM calibration) and print its GB/s")
    a = ap.parse_args(argv)
    if a.copy_probe:
        import pathtracingtutorial_amd as pt
        nbytes = 1 << 30
        gbps = pt.probe_copy(nbytes, 20)
        print(json.dumps({"probe": "copy", "bytes_per_launch_read": nbytes, "bytes_per_launch_write": nbytes,
                          "launches": 21, "gbps": round(gbps, 1)}), flush=True)
        return 0
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported in this build "
              "(the path shards by sample index across GPUs, see DESIGN.md)", file=sys.stderr)
        return 2

    import pathtracingtutorial_amd as pt
    integ = pt.PathIntegrator() if a.integrator == "path" else pt.WavefrontIntegrator()
    sm = pt.SceneManager()
    t0 = time.perf_counter()
    sm.generate(1, a.seed)
    log(f"scene generated in {time.perf_counter() - t0:.2f} s")
    t0 = time.perf_counter()
    sm.buildAccel()  # the workload's options: 32 bins, <= 4 per leaf
    info = sm.info()
    log(f"accel: {info.num_faces} tris, {info.num_bvh_nodes} nodes, depth {info.bvh_depth}, "
        f"GPU build {info.bvh_build_ms:.2f} ms, total {time.perf_counter() - t0:.2f} s")
    sm.uploadToDevice()
    r = pt.Renderer(sm, integ)
    for i in range(a.warmup):
        r.renderBatch()
        log(f"warmup step {i}: {r.lastBatchMs():.1f} ms")
    # timed: frames 0..K-1 from a cleared film
    r.resetAccumulation()
    kernel_ms = []
    t0 = time.perf_counter()
    for i in range(a.steps):
        r.renderBatch()
        kernel_ms.append(r.lastBatchMs())
        log(f"step {i}: {kernel_ms[-1]:.1f} ms")
    wall = time.perf_counter() - t0
    # the same frames again with in-kernel counters, for exact ray and node counts
    r.resetAccumulation()
    sm.enableCounters(True)
    sm.resetCounters()
    for i in range(a.steps):
        r.renderBatch()
    c = sm.counters()
    sm.enableCounters(False)
    rays = c["closest_rays"] + c["shadow_rays"]
    ms_step = wall * 1000.0 / a.steps
    mrays = rays / wall / 1e6
    avg_kernel_s = (sum(kernel_ms) / len(kernel_ms)) / 1000.0
    bytes_per_launch = algorithmic_bytes(c) / a.steps
    achieved = bytes_per_launch / avg_kernel_s / 1e9
    img = r.downloadCurrentImage()
    log(f"counters: {c}; image mean {img[..., :3].mean():.4f}; finite {bool(np.isfinite(img).all())}")
    out = {
        "metric": "Mrays/s (1M tris, 1080p) wavefront vs plain at equal relMSE; backward Mrays/s",
        "value": round(mrays, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_step, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": WORKLOAD, "triangles": info.num_faces, "resolution": [info.width, info.height],
                   "spp": info.sample_per_frame, "max_depth": info.max_depth, "sampler": "sobol",
                   "bvh": {"bins": 32, "max_leaf": 4, "nodes": info.num_bvh_nodes, "depth": info.bvh_depth},
                   "integrator": integ.name(), "paths_per_step": info.width * info.height * info.sample_per_frame,
                   "rays_per_step": rays // a.steps,
                   "shadow_fraction": round(c["shadow_rays"] / max(1, rays), 4)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": profiled_traffic("k_render_plain", WORKLOAD),
                     "kernel": "k_render_plain", "kernel_ms": round(avg_kernel_s * 1000, 3),
                     "bytes_per_launch": int(bytes_per_launch),
                     "nodes_per_ray": round(c["node_visits"] / max(1, rays), 2),
                     "tris_per_ray": round(c["tri_tests"] / max(1, rays), 2)},
    }
    if not a.no_cpu_baseline:
        try:
            out["cpu_baseline"] = cpu_baseline(sm)
        except Exception as e:  # the baseline is reported, never required for the GPU number
            out["cpu_baseline"] = {"value": None, "unit": "Mrays/s", "error": str(e)}
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
