"""Writes the golden vectors of tests/golden/ from the reference's own code.

Runs only where the reference tree is present: it builds oracle/_ref/libref.so
(the reference's host sources compiled as-is, see oracle/Makefile) and
records inputs and outputs of the reference functions on seeded inputs.
The fixtures are data only (numbers, hashes); no reference source travels.

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import ctypes as C
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
import pathtracingtutorial_amd as pt  # noqa: E402
from pathtracingtutorial_amd import _capi  # noqa: E402

REF = oracle.REF_TREE
_FP = C.POINTER(C.c_float)
_U32P = C.POINTER(C.c_uint32)
_I32P = C.POINTER(C.c_int32)


def f32p(a):
    return a.ctypes.data_as(_FP)


def sha(b: bytes) -> str:
    return hashlib.sha256(b).hexdigest()


def setup(ref):
    f = C.c_float
    u = C.c_uint32
    sig = {
        "ref_layout": (None, [_U32P]),
        "ref_default_material": (None, [C.c_void_p]),
        "ref_sobol_get1d": (f, [C.c_int, C.c_int, u, u]),
        "ref_sobol_hashmix": (u, [u]),
        "ref_sobol_owen": (u, [u, u]),
        "ref_sobol_dimseed": (u, [C.c_int, C.c_int, u]),
        "ref_reverse_bits": (u, [u]),
        "ref_scene_build": (C.c_void_p, [C.c_void_p, u, C.c_void_p, u]),
        "ref_scene_free": (None, [C.c_void_p]),
        "ref_bvh_num_nodes": (u, [C.c_void_p]),
        "ref_bvh_get": (None, [C.c_void_p, C.c_void_p, _U32P]),
        "ref_trace": (None, [C.c_void_p, _FP, C.c_int, C.c_int, _FP, _I32P, _FP]),
        "ref_env_sampler": (u, [_FP, C.c_int, C.c_int, _FP, _FP]),
        "ref_emissive_sampler": (u, [C.c_void_p, u, C.c_void_p, u, C.c_void_p, u, _FP, _FP, _U32P]),
        "ref_dir_to_uv": (None, [_FP, _FP]),
        "ref_uv_to_dir": (None, [_FP, _FP]),
        "ref_d_ggx": (f, [f, f]),
        "ref_d_ggx_aniso": (f, [f, f, f, f, f]),
        "ref_g1": (f, [f, f]),
        "ref_g2": (f, [f, f, f]),
        "ref_f_schlick": (f, [f, f]),
        "ref_f_schlick3": (None, [f, _FP, _FP]),
        "ref_f_dielectric": (f, [f, f]),
        "ref_sample_vndf": (None, [_FP, f, f, f, f, _FP]),
        "ref_pdf_vndf": (f, [_FP, _FP, f, f]),
        "ref_refract": (C.c_int, [_FP, _FP, f, _FP]),
        "ref_cosine_hemisphere": (None, [f, f, _FP]),
        "ref_power_heuristic": (f, [f, f]),
        "ref_frame": (None, [_FP, _FP, _FP, _FP]),
        "ref_load_obj": (C.c_int, [C.c_char_p, C.c_int, _FP, _FP, f, f, C.c_void_p, u, _U32P, C.c_void_p, u, _U32P]),
        "ref_load_image": (C.c_int, [C.c_char_p, _FP, C.c_size_t, _I32P, _I32P]),
        "ref_build_quad": (None, [_FP, _FP, _FP, _FP, C.c_int, C.c_void_p, C.c_void_p]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(ref, name)
        fn.restype = res
        fn.argtypes = args


def tri_soup(rng, n, spread=10.0, size=1.0):
    """n random triangles as (vertices bytes, faces bytes)."""
    verts = (_capi.Vertex * (3 * n))()
    faces = (_capi.Face * n)()
    c = rng.uniform(-spread, spread, (n, 3))
    for i in range(n):
        p = c[i] + rng.uniform(-size, size, (3, 3))
        nrm = np.cross(p[1] - p[0], p[2] - p[0])
        nrm = nrm / (np.linalg.norm(nrm) + 1e-30)
        for k in range(3):
            v = verts[3 * i + k]
            v.position = _capi.Vec3p.of(p[k])
            v.normal = _capi.Vec3p.of(nrm)
        faces[i] = _capi.Face(3 * i, 3 * i + 1, 3 * i + 2, 0)
    return bytes(verts), bytes(faces), 3 * n, n


def scene_arrays(sm: pt.SceneManager):
    v = _capi.HostSceneView()
    _capi.check(pt.lib().pt_scene_host_view(sm.handle, C.byref(v)))
    vb = C.string_at(v.vertices, 64 * v.num_vertices)
    fb = C.string_at(v.faces, 16 * v.num_faces)
    mb = C.string_at(v.materials, 256 * v.num_materials)
    return vb, fb, mb, v.num_vertices, v.num_faces, v.num_materials, v


def ref_bvh(ref, vb, nv, fb, nf):
    h = ref.ref_scene_build(vb, nv, fb, nf)
    n = ref.ref_bvh_num_nodes(h)
    nodes = np.zeros((n, 8), np.float32)
    order = np.zeros(nf, np.uint32)
    ref.ref_bvh_get(h, nodes.ctypes.data_as(C.c_void_p), order.ctypes.data_as(_U32P))
    return h, nodes, order


def seeded_rays(rng, n, lo, hi, toward=None):
    o = rng.uniform(lo, hi, (n, 3)).astype(np.float32)
    if toward is None:
        d = rng.normal(size=(n, 3))
    else:
        d = (toward + rng.normal(scale=0.3 * np.linalg.norm(np.asarray(hi) - np.asarray(lo)), size=(n, 3))) - o
    rays = np.zeros((n, 8), np.float32)
    rays[:, :3] = o
    rays[:, 3:6] = d.astype(np.float32)
    rays[:, 6] = 3.4028234e38
    return rays


def main():
    oracle.build(ref=True)
    ref = oracle.ref_lib()
    if ref is None:
        print("reference tree not present; nothing to do")
        return 1
    setup(ref)
    rng = np.random.default_rng(20261203)
    out = {}

    # ---- layout + default material
    lay = np.zeros(24, np.uint32)
    ref.ref_layout(lay.ctypes.data_as(_U32P))
    out["layout"] = lay.tolist()
    dm = C.create_string_buffer(256)
    ref.ref_default_material(dm)
    out["default_material_hex"] = dm.raw.hex()

    # ---- Sobol
    sob = []
    for _ in range(400):
        px, py = int(rng.integers(0, 4096)), int(rng.integers(0, 4096))
        idx, dim = int(rng.integers(0, 1 << 20)), int(rng.integers(0, 48))
        sob.append([px, py, idx, dim, float(np.float32(ref.ref_sobol_get1d(px, py, idx, dim)))])
    for px, py, idx in [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (1919, 1079, 63), (511, 511, 15)]:
        for dim in range(12):
            sob.append([px, py, idx, dim, float(np.float32(ref.ref_sobol_get1d(px, py, idx, dim)))])
    out["sobol_get1d"] = sob
    kat = []
    for x in [0, 1, 2, 0x9e3779b9, 0xdeadbeef, 0xffffffff, 12345678]:
        kat.append([x, ref.ref_sobol_hashmix(x), ref.ref_reverse_bits(x), ref.ref_sobol_owen(x, 0x1234567),
                    ref.ref_sobol_dimseed(x & 0xfff, (x >> 12) & 0xfff, x & 31)])
    out["sobol_kat"] = kat

    # ---- BVH on small meshes with the reference settings (full node arrays)
    bvh_cases = {}
    # the reference unit-test meshes (tests/test_bvh.cpp:27-81)
    def row_tris(k, pitch):
        verts = (_capi.Vertex * (3 * k))()
        faces = (_capi.Face * k)()
        for i in range(k):
            x = i * pitch
            pts = [(x, 0, 0), (x + 1, 0, 0), (x + 0.5, 1, 0)]
            for j in range(3):
                verts[3 * i + j].position = _capi.Vec3p.of(pts[j])
                verts[3 * i + j].normal = _capi.Vec3p.of((0, 0, 1))
            faces[i] = _capi.Face(3 * i, 3 * i + 1, 3 * i + 2, 0)
        return bytes(verts), bytes(faces), 3 * k, k
    cases = {"row1": row_tris(1, 3.0), "row5": row_tris(5, 3.0), "row100": row_tris(100, 1.0),
             "soup200": tri_soup(rng, 200), "soup3000": tri_soup(rng, 3000, 20.0, 0.7)}
    sm = pt.SceneManager()
    sm.generate(0)
    vb, fb, mb, nv, nf, nm, _ = scene_arrays(sm)
    cases["cornell"] = (vb, fb, nv, nf)
    for name, (vb, fb, nv, nf) in cases.items():
        h, nodes, order = ref_bvh(ref, vb, nv, fb, nf)
        u = nodes.view(np.uint32)
        entry = {"nv": nv, "nf": nf, "num_nodes": int(nodes.shape[0]), "nodes_sha": sha(nodes.tobytes()),
                 "order_sha": sha(order.tobytes())}
        if nodes.shape[0] <= 400:
            entry["nodes"] = [[float(r[0]), float(r[1]), float(r[2]), int(u[i, 3]), float(r[4]), float(r[5]),
                               float(r[6]), int(u[i, 7])] for i, r in enumerate(nodes)]
            entry["order"] = order.tolist()
        # leaf membership as sorted face sets, which is what a different partition order preserves
        leaves = []
        for i in range(nodes.shape[0]):
            if u[i, 3] & 0x80000000:
                first, cnt = int(u[i, 3] & 0x7FFFFFFF), int(u[i, 7])
                leaves.append(sorted(order[first:first + cnt].tolist()))
        entry["leaf_sets_sha"] = sha(json.dumps(leaves).encode())
        # seeded rays -> reference closest hit / any hit
        lo = np.frombuffer(vb, np.float32).reshape(-1, 16)[:, :3].min(0)
        hi = np.frombuffer(vb, np.float32).reshape(-1, 16)[:, :3].max(0)
        rays = seeded_rays(rng, 512, lo - 0.5, hi + 0.5, toward=(lo + hi) / 2)
        t = np.zeros(len(rays), np.float32)
        face = np.zeros(len(rays), np.int32)
        bary = np.zeros((len(rays), 2), np.float32)
        ref.ref_trace(h, f32p(rays), len(rays), 0, f32p(t), face.ctypes.data_as(_I32P), f32p(bary))
        rays_any = rays.copy()
        rays_any[:, 6] = rng.uniform(0.1, 1.0, len(rays)).astype(np.float32) * float(np.linalg.norm(hi - lo))
        ta = np.zeros(len(rays), np.float32)
        fa = np.zeros(len(rays), np.int32)
        ref.ref_trace(h, f32p(rays_any), len(rays), 1, f32p(ta), fa.ctypes.data_as(_I32P), f32p(bary))
        np.savez_compressed(os.path.join(HERE, f"trace_{name}.npz"), rays=rays, t=t, face=face, rays_any=rays_any,
                            occluded=ta)
        entry["hits"] = int((face >= 0).sum())
        ref.ref_scene_free(h)
        bvh_cases[name] = entry
        if name.startswith("soup"):
            np.savez_compressed(os.path.join(HERE, f"mesh_{name}.npz"), vertices=np.frombuffer(vb, np.uint8),
                                faces=np.frombuffer(fb, np.uint8))
    out["bvh"] = bvh_cases

    # ---- OBJ meshes through MeshLoader + the BVH over sphere2.obj (20,480 faces)
    objs = {}
    for name, t3, axis, ang, sc in [("sphere2.obj", (0, 0, 0), (0, 1, 0), 0.0, 1.0),
                                    ("Stanford Bunny.obj", (0.1, -0.2, 0.3), (0, 1, 0), 30.0, 2.0),
                                    ("teapot.obj", (0, 0, 0), (1, 1, 0), 45.0, 0.5),
                                    ("sphere.obj", (0, 0, 0), (0, 1, 0), 0.0, 1.0)]:
        path = os.path.join(REF, "models", name).encode()
        nv, nf = C.c_uint32(), C.c_uint32()
        t3a = np.array(t3, np.float32)
        axa = np.array(axis, np.float32)
        ref.ref_load_obj(path, 2, f32p(t3a), f32p(axa), ang, sc, None, 0, C.byref(nv), None, 0, C.byref(nf))
        vbuf = C.create_string_buffer(64 * nv.value)
        fbuf = C.create_string_buffer(16 * nf.value)
        ok = ref.ref_load_obj(path, 2, f32p(t3a), f32p(axa), ang, sc, vbuf, nv.value, C.byref(nv), fbuf, nf.value,
                              C.byref(nf))
        va = np.frombuffer(vbuf.raw, np.float32).reshape(-1, 16)
        objs[name] = {"ok": ok, "nv": nv.value, "nf": nf.value, "translate": list(t3), "axis": list(axis),
                      "angle": ang, "scale": sc,
                      "pos_sha": sha(np.ascontiguousarray(va[:, 0:3]).tobytes()),
                      "nrm_sha": sha(np.ascontiguousarray(va[:, 4:7]).tobytes()),
                      "uv_sha": sha(np.ascontiguousarray(va[:, 8:10]).tobytes()),
                      "faces_sha": sha(fbuf.raw),
                      "first_vertices": va[:4, [0, 1, 2, 4, 5, 6, 8, 9]].astype(float).tolist()}
        if name == "sphere2.obj":
            h, nodes, order = ref_bvh(ref, vbuf.raw, nv.value, fbuf.raw, nf.value)
            u = nodes.view(np.uint32)
            leaves = []
            for i in range(nodes.shape[0]):
                if u[i, 3] & 0x80000000:
                    first, cnt = int(u[i, 3] & 0x7FFFFFFF), int(u[i, 7])
                    leaves.append(sorted(order[first:first + cnt].tolist()))
            objs[name]["bvh"] = {"num_nodes": int(nodes.shape[0]), "nodes_sha": sha(nodes.tobytes()),
                                 "order_sha": sha(order.tobytes()), "leaf_sets_sha": sha(json.dumps(leaves).encode())}
            rays = seeded_rays(rng, 4096, np.array([-2, -2, -2]), np.array([2, 2, 2]), toward=np.zeros(3))
            t = np.zeros(len(rays), np.float32)
            face = np.zeros(len(rays), np.int32)
            bary = np.zeros((len(rays), 2), np.float32)
            ref.ref_trace(h, f32p(rays), len(rays), 0, f32p(t), face.ctypes.data_as(_I32P), f32p(bary))
            np.savez_compressed(os.path.join(HERE, "trace_sphere2.npz"), rays=rays, t=t, face=face,
                                positions=np.ascontiguousarray(va[:, 0:3]),
                                faces=np.frombuffer(fbuf.raw, np.uint32).reshape(-1, 4)[:, :3].copy())
            ref.ref_scene_free(h)
    out["obj"] = objs

    # ---- default.hdr through ImageIO::load (stb_image 2.30)
    w, h_ = C.c_int(), C.c_int()
    path = os.path.join(REF, "default.hdr").encode()
    ref.ref_load_image(path, None, 0, C.byref(w), C.byref(h_))
    img = np.zeros(w.value * h_.value * 4, np.float32)
    ref.ref_load_image(path, f32p(img), img.size, C.byref(w), C.byref(h_))
    img = img.reshape(h_.value, w.value, 4)
    samp = rng.integers(0, [h_.value, w.value], (32, 2))
    out["hdr_default"] = {"w": w.value, "h": h_.value, "sha": sha(img.tobytes()),
                          "samples": [[int(y), int(x), *img[y, x].astype(float).tolist()] for y, x in samp]}
    # env tables of a crop of it plus the reference test images (tests/test_sampling.cpp:70-110)
    envs = {"t2x1": np.array([[[1, 1, 1, 1], [3, 3, 3, 1]]], np.float32),
            "t4x1": np.array([[[1, 0, 0, 1], [2, 0, 0, 1], [3, 0, 0, 1], [4, 0, 0, 1]]], np.float32),
            "rand16x8": np.concatenate([rng.exponential(1.0, (8, 16, 3)), np.ones((8, 16, 1))], -1).astype(np.float32),
            "crop64x32": np.ascontiguousarray(img[300:332, 200:264])}
    env_out = {}
    for name, e in envs.items():
        e = np.ascontiguousarray(e, np.float32)
        n = e.shape[0] * e.shape[1]
        pmf = np.zeros(n, np.float32)
        al = np.zeros((n, 2), np.float32)
        k = ref.ref_env_sampler(f32p(e), e.shape[1], e.shape[0], f32p(pmf), f32p(al))
        env_out[name] = {"image": e.astype(float).tolist(), "n": int(k), "pmf": pmf[:k].astype(float).tolist(),
                         "prob": al[:k, 0].astype(float).tolist(),
                         "alias": al[:k].view(np.uint32)[:, 1].tolist()}
    out["env_tables"] = env_out

    # ---- emissive tables: the reference test case and the Cornell box
    def mats_bytes(ms):
        return b"".join(bytes(m) for m in ms)
    m0 = _capi.Material.from_buffer_copy(dm.raw)
    m1 = _capi.Material.from_buffer_copy(dm.raw)
    m2 = _capi.Material.from_buffer_copy(dm.raw)
    m1.emissive = _capi.Vec3p.of((2, 2, 2))
    m2.base_color = _capi.Vec3p.of((0.5, 0.5, 0.5))
    m2.emissive = _capi.Vec3p.of((1, 1, 1))
    verts = (_capi.Vertex * 9)()
    faces = (_capi.Face * 3)()
    tri = [((0, 0, 0), (1, 0, 0), (0, 1, 0)), ((0, 0, 0), (1, 0, 0), (0, 1, 0)), ((0, 0, 0), (2, 0, 0), (0, 1, 0))]
    for i, t3 in enumerate(tri):
        for k in range(3):
            verts[3 * i + k].position = _capi.Vec3p.of(t3[k])
        faces[i] = _capi.Face(3 * i, 3 * i + 1, 3 * i + 2, i)
    em_cases = {"test3": (bytes(verts), 9, bytes(faces), 3, mats_bytes([m0, m1, m2]), 3)}
    sm = pt.SceneManager()
    sm.generate(0)
    vb, fb, mb, nv, nf, nm, _ = scene_arrays(sm)
    em_cases["cornell"] = (vb, nv, fb, nf, mb, nm)
    em_out = {}
    for name, (vb, nv, fb, nf, mb, nm) in em_cases.items():
        tp = np.zeros(nf, np.float32)
        al = np.zeros((nf, 2), np.float32)
        ti = np.zeros(nf, np.uint32)
        k = ref.ref_emissive_sampler(vb, nv, fb, nf, mb, nm, f32p(tp), f32p(al), ti.ctypes.data_as(_U32P))
        em_out[name] = {"n": int(k), "tri_pmf": tp.astype(float).tolist(), "tri_index": ti[:k].tolist(),
                        "prob": al[:k, 0].astype(float).tolist(), "alias": al[:k].view(np.uint32)[:, 1].tolist()}
        if name == "test3":
            em_out[name]["vertices_hex"] = vb.hex()
            em_out[name]["faces_hex"] = fb.hex()
            em_out[name]["materials_hex"] = mb.hex()
    out["emissive_tables"] = em_out

    # ---- env mapping, microfacet, sampling, frame
    dirs = rng.normal(size=(256, 3)).astype(np.float32)
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    uvs = np.zeros((256, 2), np.float32)
    for i in range(256):
        ref.ref_dir_to_uv(f32p(dirs[i]), f32p(uvs[i]))
    uin = rng.uniform(0, 1, (256, 2)).astype(np.float32)
    dout = np.zeros((256, 3), np.float32)
    for i in range(256):
        ref.ref_uv_to_dir(f32p(uin[i]), f32p(dout[i]))
    out["envmap"] = {"dirs": dirs.astype(float).tolist(), "uv": uvs.astype(float).tolist(),
                     "uv_in": uin.astype(float).tolist(), "dir_out": dout.astype(float).tolist()}
    mf = []
    for _ in range(200):
        noh, hx, hy = float(np.float32(rng.uniform(0, 1))), float(np.float32(rng.uniform(-1, 1))), \
            float(np.float32(rng.uniform(-1, 1)))
        a, ax, ay = [float(np.float32(x)) for x in rng.uniform(0.001, 1, 3)]
        nov, nol, c = [float(np.float32(x)) for x in rng.uniform(0, 1, 3)]
        eta = float(np.float32(rng.uniform(0.5, 2.0)))
        f0 = np.float32(rng.uniform(0, 1, 3)).astype(np.float32)
        fs3 = np.zeros(3, np.float32)
        ref.ref_f_schlick3(c, f32p(f0), f32p(fs3))
        wo = rng.normal(size=3).astype(np.float32)
        wo[2] = abs(wo[2]) + 0.05
        wo /= np.linalg.norm(wo)
        wo = wo.astype(np.float32)
        u1, u2 = [float(np.float32(x)) for x in rng.uniform(0, 1, 2)]
        hv = np.zeros(3, np.float32)
        ref.ref_sample_vndf(f32p(wo), ax, ay, u1, u2, f32p(hv))
        pv = ref.ref_pdf_vndf(f32p(wo), f32p(hv), ax, ay)
        nrm = np.array([0, 0, 1], np.float32)
        wt = np.zeros(3, np.float32)
        rok = ref.ref_refract(f32p(wo), f32p(nrm), eta, f32p(wt))
        cos3 = np.zeros(3, np.float32)
        ref.ref_cosine_hemisphere(u1, u2, f32p(cos3))
        n3 = rng.normal(size=3).astype(np.float32)
        n3 /= np.linalg.norm(n3)
        n3 = n3.astype(np.float32)
        loc = np.zeros(3, np.float32)
        wor = np.zeros(3, np.float32)
        ref.ref_frame(f32p(n3), f32p(wo), f32p(loc), f32p(wor))
        mf.append({"in": [noh, hx, hy, a, ax, ay, nov, nol, c, eta, u1, u2], "f0": f0.astype(float).tolist(),
                   "wo": wo.astype(float).tolist(), "n": n3.astype(float).tolist(),
                   "d_ggx": ref.ref_d_ggx(noh, a), "d_aniso": ref.ref_d_ggx_aniso(noh, hx, hy, ax, ay),
                   "g1": ref.ref_g1(nov, a), "g2": ref.ref_g2(nol, nov, a), "fs": ref.ref_f_schlick(c, 0.04),
                   "fs3": fs3.astype(float).tolist(), "fd": ref.ref_f_dielectric(c, eta),
                   "vndf": hv.astype(float).tolist(), "pdf_vndf": pv, "refract_ok": rok,
                   "refract": wt.astype(float).tolist(), "cosine": cos3.astype(float).tolist(),
                   "power": ref.ref_power_heuristic(nov, nol), "local": loc.astype(float).tolist(),
                   "world": wor.astype(float).tolist()})
    out["microfacet"] = mf

    # ---- quad builder (geometry/shapes.cpp)
    qa = [np.array(p, np.float32) for p in ([0.3, -0.5, 0.2], [1.1, -0.4, 0.25], [1.0, 0.7, -0.3], [0.2, 0.6, -0.2])]
    qv = C.create_string_buffer(4 * 64)
    qf = C.create_string_buffer(2 * 16)
    ref.ref_build_quad(f32p(qa[0]), f32p(qa[1]), f32p(qa[2]), f32p(qa[3]), 7, qv, qf)
    out["quad"] = {"in": [q.astype(float).tolist() for q in qa], "vertices_hex": qv.raw.hex(), "faces_hex": qf.raw.hex()}

    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(out, f)
    print("wrote", os.path.join(HERE, "golden.json"), os.path.getsize(os.path.join(HERE, "golden.json")), "bytes")
    return 0


if __name__ == "__main__":
    sys.exit(main())
