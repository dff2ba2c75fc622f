"""Benchmark of the path-tracing hot path on one MI355X.

Workload: 64 displaced level-5 icospheres (1,310,720 triangles) + ground,
64 random Disney materials, 2048x1024 sky+sun environment, 1920x1080,
64 spp, depth 16, Owen-scrambled Sobol, BVH with 32 bins and <= 4 triangles
per leaf.  One step = one 64-spp frame through the integrator.

Prints one JSON line:
  value      traced rays (closest + shadow) per second, Mrays/s, whole job, of
             the wavefront integrator (the plain megakernel's figures sit in
             "plain", rendered from the same random numbers: the same estimate)
  roofline   the dominant traversal kernel's algorithmic bytes per launch over
             its average HIP-event duration, against the 8 TB/s HBM3E peak:
             k_wf_extend (32 B per 32-byte unit of node data fetched: a binary
             node record is one, a 4-wide node two; + 36 B per triangle test
             + 52 B of ray/hit/queue I/O per ray); for --integrator path,
             k_render_plain (48 B per ray of I/O)
  cpu_baseline  the oracle (a C++ restatement of the reference integrator)
             on a bounded sample of the same frame, all host cores

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--integrator both|wavefront|path]
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


# Bytes the wavefront shading kernel moves per counted event (the workload has
# an environment light and no emitting triangles, so a light sample is the
# environment's): a shaded item reads its queue entry (4), path record (64)
# and hit (16); a continuing path's record is written back (64, counted with
# the continuation); an escape reads 4 bilinear
# environment texels (64) and the texel's pmf (4, the MIS weight); a surface
# hit reads the triangle's shading record (64: three vertex normals and uvs,
# the material index) and the 256-byte material record; a light sample reads
# the alias entry (8), the
# pmf (4) and 4 texels (64); a shadow-queue entry is 48 bytes written, a
# queued continuation 4 + its record 64, a radiance update 16 read + 16
# written.
SHADE_EVENT_BYTES = {"shade_items": 4 + 64 + 16, "shade_env_lookups": 64 + 4, "shade_hits": 64 + 256,
                     "shade_light_samples": 8 + 4 + 64, "shade_shadow_entries": 48, "shade_continuations": 4 + 64,
                     "shade_radiance_adds": 32}


def shade_bytes(c: dict) -> int:
    return sum(n * c[k] for k, n in SHADE_EVENT_BYTES.items())


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


def profiled_issue(kernel: str):
    """Issue and memory counters of `kernel` (its non-counting instantiation)
    from the committed rocprofv3 summary (profiles/wavefront_issue_counters.md):
    VALU issue share of all SIMD cycles, share of lanes active per VALU
    instruction, share of wave time waiting on memory, L2 hit rate; or None."""
    import glob
    import re
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*.md"))):
        for block in re.findall(r"```json\n(.*?)\n```", open(path).read(), re.S):
            try:
                d = json.loads(block)
            except ValueError:
                continue
            for name, v in d.items():
                if isinstance(v, dict) and name.startswith(kernel + "<") and "valu_issue_share" in v:
                    return {"kernel": name, **{k: v[k] for k in ("valu_issue_share", "active_lanes", "wait_share",
                                                                  "l2_hit")}}
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
    # the host share this process may use (the OpenMP-style thread cap where
    # one is set, else the affinity mask)
    cores = int(os.environ.get("OMP_NUM_THREADS") or 0) or len(os.sched_getaffinity(0)) or 1
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
                      f"{dt:.1f} s); oracle BVH build {build_s:.1f} s",
            "cpu": _cpu_model(), "build": oracle.lib().oracle_build_flags().decode()}


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def run_integrator(sm, integ, steps: int, warmup: int) -> dict:
    """Warmup, then `steps` timed frames from a cleared film, then the same
    frames again with in-kernel counters (exact ray / node / triangle counts)."""
    import pathtracingtutorial_amd as pt
    r = pt.Renderer(sm, integ)
    for i in range(warmup):
        r.renderBatch()
        log(f"{integ.name()} warmup step {i}: {r.lastBatchMs():.1f} ms")
    r.resetAccumulation()
    kernel_ms = []
    t0 = time.perf_counter()
    for i in range(steps):
        r.renderBatch()
        kernel_ms.append(r.lastBatchMs())
    wall = time.perf_counter() - t0
    log(f"{integ.name()} steps: " + ", ".join(f"{m:.1f}" for m in kernel_ms) + f" ms; wall {wall:.3f} s")
    img = r.downloadCurrentImage()
    r.resetAccumulation()
    sm.enableCounters(True)
    sm.resetCounters()
    for i in range(steps):
        r.renderBatch()
    c = sm.counters()
    sm.enableCounters(False)
    return {"wall": wall, "kernel_ms": kernel_ms, "counters": c, "image": img, "renderer": r}


def stage_times(sm, steps: int) -> dict:
    """One more frame of the wavefront with every launch bracketed by HIP events
    on the integrator's stream: per-stage GPU ms and launch counts."""
    import pathtracingtutorial_amd as pt
    cfg = sm.wavefrontConfig()
    sm.setWavefrontConfig(time_stages=1)
    r = pt.Renderer(sm, pt.WavefrontIntegrator())
    r.renderBatch()
    st = sm.wavefrontStats()
    sm.setWavefrontConfig(**cfg)
    return st


def backward_leg(steps: int = 3) -> dict:
    """Forward and path-replay backward of the differentiable Cornell box
    (256x256, 64 spp, depth 6): GPU ms of each pass and the rays the backward
    pass traces per second (the wavefront forward keeps each path's
    radiance, so the backward traces only the adjoint replay)."""
    import torch

    import pathtracingtutorial_amd as pt
    from pathtracingtutorial_amd.diff import DiffRenderer
    sm = pt.SceneManager()
    sm.generate(3, 3)
    sm.buildAccel()
    sm.uploadToDevice()
    dr = DiffRenderer(sm)
    info = sm.info()
    g = torch.ones((info.height, info.width, 4), dtype=torch.float32, device="cuda")
    dr.forward_image(0)
    dr.adjoint(0, g)
    f_ms, b_ms, reused = [], [], []
    for i in range(steps):
        dr.forward_image(i)
        f_ms.append(dr.last_forward_ms)
        dr.adjoint(i, g)
        b_ms.append(dr.last_backward_ms)
        reused.append(dr.last_backward_reused)
    sm.enableCounters(True)
    dr.forward_image(0)
    sm.resetCounters()
    dr.adjoint(0, g)
    c = sm.counters()
    sm.enableCounters(False)
    rays = c["closest_rays"] + c["shadow_rays"]
    fm, bm = float(np.mean(f_ms)), float(np.mean(b_ms))
    return {"workload": "diff-cornell-36tri-256x256-64spp-depth6-pcg", "value": round(rays / (bm / 1e3) / 1e6, 2),
            "unit": "Mrays/s", "backward_ms": round(bm, 3), "forward_ms": round(fm, 3),
            "backward_over_forward": round(bm / fm, 3), "rays_per_backward": rays,
            "params": "256x256 albedo texture, roughness + metallic of 8 materials, 512x256 env",
            "forward": "wavefront", "backward_seeded_from_forward": all(reused)}


def batched_leg(njobs: int = 64, seed: int = 1000) -> dict:
    """The batched workload: njobs seeded scenes (icospheres of 1e3..5e5
    triangles, 128^2..1024^2 pixels, 4..64 spp, depth 12, own 512x256 sky) in
    one wavefront schedule, against the same scenes one at a time through the
    plain megakernel.  GPU time from HIP events; rays from in-kernel counters
    of a counted re-run."""
    import pathtracingtutorial_amd as pt
    t0 = time.perf_counter()
    scenes = []
    for j in range(njobs):
        sm = pt.SceneManager()
        sm.generate(4, seed + j)
        sm.buildAccel()
        sm.uploadToDevice()
        scenes.append(sm)
    infos = [sm.info() for sm in scenes]
    log(f"batched workload: {njobs} scenes built in {time.perf_counter() - t0:.1f} s; "
        f"{sum(i.num_faces for i in infos)} triangles, {sum(i.width * i.height for i in infos)} pixels, "
        f"{sum(i.width * i.height * i.sample_per_frame for i in infos)} paths")
    # one at a time through the megakernel
    plain = [pt.Renderer(sm, pt.PathIntegrator()) for sm in scenes]
    for r in plain:
        r.renderBatch()  # warm-up
    for sm in scenes:
        sm.uploadToDevice()
    plain_ms = 0.0
    for r in plain:
        r.renderBatch()
        plain_ms += r.lastBatchMs()
    # batched wavefront (the first scene's engine and autotuned-by-default config)
    for sm in scenes:
        sm.uploadToDevice()
    wf = [pt.Renderer(sm, pt.WavefrontIntegrator()) for sm in scenes]
    pt.render_jobs(wf)  # warm-up + graph capture
    for sm in scenes:
        sm.uploadToDevice()
    batch_ms = pt.render_jobs(wf)
    # rays (the same frames, counted)
    for sm in scenes:
        sm.uploadToDevice()
    scenes[0].enableCounters(True)
    scenes[0].resetCounters()
    pt.render_jobs(wf)
    c = scenes[0].counters()
    scenes[0].enableCounters(False)
    rays = c["closest_rays"] + c["shadow_rays"]
    return {"workload": f"{njobs} seeded icosphere scenes (seeds {seed}..{seed + njobs - 1}), depth 12",
            "value": round(rays / (batch_ms / 1e3) / 1e6, 2), "unit": "Mrays/s",
            "batched_ms": round(batch_ms, 2), "sequential_megakernel_ms": round(plain_ms, 2),
            "speedup": round(plain_ms / batch_ms, 3),
            "sequential_megakernel_mrays": round(rays / (plain_ms / 1e3) / 1e6, 2), "rays": rays,
            "paths": c["paths"]}


def views_leg(steps: int = 2) -> dict:
    """Differentiable + batched: 16 cameras on a ring (radius 8, 35 deg) around
    the 2,048,000-triangle scene, 512x512 each, 32 spp forward in one wavefront
    schedule and 32 spp path-replay backward in one launch, depth 10;
    parameters: the 1024x1024 ground albedo texture and the 1024x512 env."""
    import torch

    import pathtracingtutorial_amd as pt
    from pathtracingtutorial_amd.diff import MultiViewRenderer, ring_cameras
    sm = pt.SceneManager()
    sm.generate(5, 1)
    sm.buildAccel()
    sm.uploadToDevice()
    mv = MultiViewRenderer(sm, ring_cameras(16))
    g = torch.full((16, 512, 512, 4), 1.0 / (16 * 512 * 512), dtype=torch.float32, device="cuda")
    want = (False, False, False, True, True)
    mv.forward_images(0)
    mv.adjoint(0, g, want)
    f_ms, b_ms, reused = [], [], []
    for i in range(steps):
        mv.forward_images(i)
        f_ms.append(mv.last_forward_ms)
        mv.adjoint(i, g, want)
        b_ms.append(mv.last_backward_ms)
        reused.append(mv.last_backward_reused)
    sm.enableCounters(True)
    sm.resetCounters()
    mv.forward_images(0)
    cf = sm.counters()
    sm.resetCounters()
    mv.adjoint(0, g, want)
    cb = sm.counters()
    sm.enableCounters(False)
    fm, bm = float(np.mean(f_ms)), float(np.mean(b_ms))
    rf, rb = cf["closest_rays"] + cf["shadow_rays"], cb["closest_rays"] + cb["shadow_rays"]
    return {"workload": "16 views x 512x512, 2048002 tris, 32 spp, depth 10, grads: 1024^2 texture + 1024x512 env",
            "forward_ms": round(fm, 2), "backward_ms": round(bm, 2), "backward_over_forward": round(bm / fm, 3),
            "forward_mrays": round(rf / (fm / 1e3) / 1e6, 2), "value": round(rb / (bm / 1e3) / 1e6, 2),
            "unit": "Mrays/s",
            # the forward keeps each path's radiance (16 B per path), so the
            # backward replays without a primal pass of its own
            "backward_seeded_from_forward": all(reused)}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--integrator", default="both", choices=["both", "path", "wavefront"])
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="skip the wavefront autotuner (use the default config)")
    ap.add_argument("--no-backward", action="store_true", help="skip the differentiable (backward) measurement")
    ap.add_argument("--no-batched", action="store_true", help="skip the 64-scene batched measurement")
    ap.add_argument("--tune-trials", type=int, default=24)
    ap.add_argument("--wavefront-config", default="",
                    help='JSON object of wavefront config fields to use instead of autotuning, '
                         'e.g. \'{"chunk_spp": 32, "extend_block": 64}\'')
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run the plain copy kernel (HBM calibration) and print its GB/s")
    ap.add_argument("--only", default="", choices=["", "batched", "backward", "views"],
                    help="only run this leg and print its JSON object (for profiling it on its own)")
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
              "(one process renders on one GPU; frames would shard by sample index across GPUs)", file=sys.stderr)
        return 2
    if a.only:
        if a.only in ("backward", "views"):
            import torch
            torch.cuda.init()  # before the library's runtime (see below)
        leg = {"batched": batched_leg, "backward": backward_leg, "views": views_leg}[a.only]
        print(json.dumps({a.only: leg()}), flush=True)
        return 0

    if not a.no_backward:
        # torch ships its own HIP runtime: it must initialise the GPU before
        # the library's runtime does, or it finds no device
        try:
            import torch
            torch.cuda.init()
        except Exception as e:
            log(f"torch GPU runtime unavailable ({e}); skipping the backward leg")
            a.no_backward = True
    import pathtracingtutorial_amd as pt
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
    runs = {}
    order = ["wavefront", "path"] if a.integrator == "both" else [a.integrator]
    tuned = None
    if a.wavefront_config:
        sm.setWavefrontConfig(**json.loads(a.wavefront_config))
    elif "wavefront" in order and not a.no_tune:
        t0 = time.perf_counter()
        tuned, ms_spp = sm.tuneWavefront(a.tune_trials)
        log(f"autotuned wavefront config in {time.perf_counter() - t0:.1f} s: {tuned} ({ms_spp:.2f} ms per spp)")
    for name in order:
        integ = pt.PathIntegrator() if name == "path" else pt.WavefrontIntegrator()
        sm.uploadToDevice()  # frame_count 0: both integrators render the same frames
        runs[name] = run_integrator(sm, integ, a.steps, a.warmup)
    head = order[0]
    R = runs[head]
    c = R["counters"]
    rays = c["closest_rays"] + c["shadow_rays"]
    ms_step = R["wall"] * 1000.0 / a.steps
    mrays = rays / R["wall"] / 1e6
    log(f"{head} counters: {c}")
    if head == "wavefront":
        st = stage_times(sm, a.steps)
        log(f"wavefront stages: {st}")
        ext = st["extend"]
        kernel_s = ext["ms"] / max(1, ext["launches"]) / 1000.0
        launches = ext["launches"] * a.steps  # the counted run renders `steps` frames
        closest_nodes = c["node_visits"] - c["shadow_node_visits"]
        closest_tris = c["tri_tests"] - c["shadow_tri_tests"]
        bytes_per_launch = (32 * closest_nodes + 36 * closest_tris + 52 * c["closest_rays"]) / max(1, launches)
        kernel = "k_wf_extend"
        per_ray = {"nodes_per_ray": round(closest_nodes / max(1, c["closest_rays"]), 2),
                   "tris_per_ray": round(closest_tris / max(1, c["closest_rays"]), 2)}
        extra = {"stages_ms_per_frame": {k: round(v["ms"], 2) for k, v in st.items()},
                 "launches_per_frame": ext["launches"]}
        sh = st["shade"]
        sh_s = sh["ms"] / max(1, sh["launches"]) / 1000.0
        sh_bytes = shade_bytes(c) / max(1, sh["launches"] * a.steps)
        sh_ach = sh_bytes / sh_s / 1e9
        # the shading stage: the hit paths' kernel and the escaped paths' kernel
        # (bytes, time and profiled traffic all cover both)
        shade_roofline = {"bound": "hbm", "kernel": "k_wf_shade", "kernels": ["k_wf_shade", "k_wf_shade_escape"],
                          "achieved": round(sh_ach, 1),
                          "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(sh_ach / HBM_PEAK_GBS, 4),
                          "traffic": profiled_traffic("k_wf_shade", WORKLOAD),
                          "issue": profiled_issue("k_wf_shade"), "kernel_ms": round(sh_s * 1000, 3),
                          "bytes_per_launch": int(sh_bytes),
                          "events_per_frame": {k: c[k] // a.steps for k in SHADE_EVENT_BYTES}}
    else:
        st = None
        kernel_s = (sum(R["kernel_ms"]) / len(R["kernel_ms"])) / 1000.0
        bytes_per_launch = algorithmic_bytes(c) / a.steps
        kernel = "k_render_plain"
        per_ray = {"nodes_per_ray": round(c["node_visits"] / max(1, rays), 2),
                   "tris_per_ray": round(c["tri_tests"] / max(1, rays), 2)}
        extra = {}
    achieved = bytes_per_launch / kernel_s / 1e9
    img = R["image"]
    log(f"image mean {img[..., :3].mean():.4f}; finite {bool(np.isfinite(img).all())}")
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
                   "integrator": "Wavefront" if head == "wavefront" else "Path",
                   "paths_per_step": info.width * info.height * info.sample_per_frame,
                   "rays_per_step": rays // a.steps,
                   "shadow_fraction": round(c["shadow_rays"] / max(1, rays), 4)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": profiled_traffic(kernel, WORKLOAD),
                     "issue": profiled_issue(kernel) if head == "wavefront" else None,
                     "kernel": kernel, "kernel_ms": round(kernel_s * 1000, 3),
                     "bytes_per_launch": int(bytes_per_launch), **per_ray, **extra},
    }
    if head == "wavefront":
        out["shade_roofline"] = shade_roofline
        out["config"]["wavefront"] = sm.wavefrontConfig()
        out["config"]["wavefront_autotuned"] = tuned is not None
    if "path" in runs and head != "path":
        P = runs["path"]
        pc = P["counters"]
        prays = pc["closest_rays"] + pc["shadow_rays"]
        p_mrays = prays / P["wall"] / 1e6
        p_kernel_s = (sum(P["kernel_ms"]) / len(P["kernel_ms"])) / 1000.0
        p_ach = algorithmic_bytes(pc) / a.steps / p_kernel_s / 1e9
        a_img, b_img = R["image"][..., :3].astype(np.float64), P["image"][..., :3].astype(np.float64)
        differ = float((np.abs(a_img - b_img) > 1e-4 * (1.0 + np.abs(b_img))).mean())
        differ_bits = int((a_img != b_img).sum())
        out["plain"] = {"value": round(p_mrays, 2), "ms_per_step": round(P["wall"] * 1000.0 / a.steps, 2),
                        "speedup_of_wavefront": round(mrays / p_mrays, 3),
                        "roofline": {"kernel": "k_render_plain", "achieved": round(p_ach, 1),
                                     "frac": round(p_ach / HBM_PEAK_GBS, 4),
                                     "traffic": profiled_traffic("k_render_plain", WORKLOAD)},
                        # same random numbers per path: the images are the same estimate
                        "image_channels_differing_1e-4": round(differ, 6),
                        "image_channels_differing_bits": differ_bits,
                        "rays_per_step": prays // a.steps}
    if not a.no_batched:
        try:
            out["batched"] = batched_leg()
            log(f"batched: {out['batched']}")
        except Exception as e:  # reported beside the forward number, never required for it
            out["batched"] = {"value": None, "error": str(e)}
    if not a.no_backward:
        try:
            out["backward"] = backward_leg()
            log(f"backward: {out['backward']}")
        except Exception as e:  # reported beside the forward number, never required for it
            out["backward"] = {"value": None, "error": str(e)}
        try:
            out["views"] = views_leg()
            log(f"views: {out['views']}")
        except Exception as e:
            out["views"] = {"value": None, "error": str(e)}
    if not a.no_cpu_baseline:
        try:
            out["cpu_baseline"] = cpu_baseline(sm)
        except Exception as e:  # the baseline is reported, never required for the GPU number
            out["cpu_baseline"] = {"value": None, "unit": "Mrays/s", "error": str(e)}
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
