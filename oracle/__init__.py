"""Oracle of the MI355X path tracer — TEST INFRASTRUCTURE ONLY.

A CPU restatement (liboracle.so, from oracle.cpp) of the reference renderer's
per-pixel integrator and the stages it calls, plus — where the reference tree
is present — the reference's own host sources compiled as-is
(_ref/libref.so, from ref_driver.cpp via the Makefile).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use
this package, and only as the checker or the timed CPU baseline.  The
product (pathtracingtutorial_amd) never imports it.

Parity pins of each oracle function are listed in oracle/README.md.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libref.so")
REF_TREE = "/root/reference/Part_2"

_P = C.c_void_p
_FP = C.POINTER(C.c_float)
_DP = C.POINTER(C.c_double)
_U32P = C.POINTER(C.c_uint32)
_I32P = C.POINTER(C.c_int32)
_U64P = C.POINTER(C.c_uint64)

_SIGS = {
    "oracle_scene_create": (_P, [_P, C.c_uint32, _P, C.c_uint32, _P, C.c_uint32, _FP, C.c_int, C.c_int, _P,
                                 C.c_uint32, C.c_int, C.c_int]),
    "oracle_scene_add_texture": (C.c_int, [_P, _FP, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
    "oracle_scene_destroy": (None, [_P]),
    "oracle_build_flags": (C.c_char_p, []),
    "oracle_bvh_num_nodes": (C.c_uint32, [_P]),
    "oracle_bvh_get": (None, [_P, _FP, _U32P]),
    "oracle_trace": (None, [_P, _FP, C.c_int, C.c_int, _FP, _I32P, _FP]),
    "oracle_light_tables": (None, [_P, _U32P, _U32P, _FP, _FP, _FP, _FP]),
    "oracle_alias_table": (C.c_uint32, [_FP, C.c_uint32, _FP, _U32P]),
    "oracle_sobol_u32": (C.c_uint32, [C.c_int, C.c_int, C.c_uint32, C.c_uint32, C.c_uint32]),
    "oracle_sobol_float": (C.c_float, [C.c_int, C.c_int, C.c_uint32, C.c_uint32, C.c_uint32]),
    "oracle_pcg": (None, [C.c_int, C.c_int, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int, _U32P, _FP]),
    "oracle_camera_frame": (None, [_P, _P]),
    "oracle_bsdf_eval": (None, [_P, _DP, _DP, _DP, C.c_int, C.c_int, _DP, _DP]),
    "oracle_bsdf_sample": (C.c_int, [_P, _DP, _DP, C.c_int, _FP, _DP, _DP, _DP, _I32P]),
    "oracle_dir_to_uv": (None, [_FP, _FP]),
    "oracle_uv_to_dir": (None, [_FP, _FP]),
    "oracle_env_pdf": (C.c_float, [_P, _FP]),
    "oracle_emissive_pdf": (C.c_double, [_P, C.c_int, C.c_double, _DP]),
    "oracle_sample_light": (C.c_int, [_P, C.c_int, _DP, _DP, C.c_int, C.c_int, C.c_uint32, C.c_uint32, _DP, _DP,
                                      _DP]),
    "oracle_tex2d": (None, [_FP, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float, _FP]),
    "oracle_render": (None, [_P, _P, C.c_int, C.c_int, C.c_int, C.c_uint32, C.c_uint32, C.c_int, C.c_int, C.c_int,
                             C.c_int, C.c_int, C.c_int, _FP, _U64P]),
    "oracle_microfacet": (None, [_FP, _FP, _FP, _FP, _FP]),
    "oracle_adjoint_fd": (None, [_P, _P, C.c_int, C.c_int, C.c_int, C.c_uint32, C.c_uint32, _FP, C.c_int, _I32P,
                                 _I32P, _I32P, _I32P, C.c_double, C.c_int, _DP, _DP]),
    "oracle_adjoint": (None, [_P, _P, C.c_int, C.c_int, C.c_int, C.c_uint32, C.c_uint32, _FP, C.c_int, _DP, _DP, _DP,
                              _DP]),
    "oracle_fetch_material": (None, [_P, _P, C.c_float, C.c_float, _FP]),
    "oracle_surface": (C.c_int, [_P, _FP, _FP, _FP]),
}

_lib = None
_ref = None


def build(ref: bool = True) -> None:
    """Compiles liboracle.so (and _ref/libref.so when the reference tree is present)."""
    targets = ["oracle"] + (["ref"] if ref else [])
    subprocess.run(["make", "-s", "-C", HERE, *targets], check=True)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_LIB):
            build(ref=False)
        h = C.CDLL(ORACLE_LIB)
        for name, (res, args) in _SIGS.items():
            fn = getattr(h, name)
            fn.restype = res
            fn.argtypes = args
        _lib = h
    return _lib


def ref_lib() -> C.CDLL | None:
    """The reference's own code (golden-vector generation only); None when absent."""
    global _ref
    if _ref is None and os.path.exists(REF_LIB):
        _ref = C.CDLL(REF_LIB)
    return _ref


def _f(a):
    return a.ctypes.data_as(_FP)


def _d(a):
    return a.ctypes.data_as(_DP)


def _u32(a):
    return a.ctypes.data_as(_U32P)


def _i32(a):
    return a.ctypes.data_as(_I32P)


class OracleScene:
    """Reference BVH + light tables over host arrays in the reference layouts.

    vertices: structured bytes (N*64), faces (N*16), materials (N*256); env (H,W,4) float32 or None.
    """

    def __init__(self, vertices: bytes, nv: int, faces: bytes, nf: int, materials: bytes, nm: int,
                 env: np.ndarray | None = None, lights: bytes = b"", nl: int = 0, bins: int = 24,
                 leaf_threshold: int = 4):
        L = lib()
        self._keep = [vertices, faces, materials, lights]
        envp, ew, eh = None, 0, 0
        if env is not None and env.size:
            env = np.ascontiguousarray(env, np.float32)
            self._keep.append(env)
            envp, eh, ew = _f(env), env.shape[0], env.shape[1]
        vb = C.create_string_buffer(vertices, len(vertices)) if vertices else None
        fb = C.create_string_buffer(faces, len(faces)) if faces else None
        mb = C.create_string_buffer(materials, len(materials))
        lb = C.create_string_buffer(lights, len(lights)) if lights else None
        self._keep += [vb, fb, mb, lb]
        self.h = L.oracle_scene_create(vb, nv, fb, nf, mb, nm, envp, ew, eh, lb, nl, bins, leaf_threshold)
        self.nf = nf
        self._nm = nm
        self._tex_shapes = []
        self._env_shape = (eh, ew, 4) if ew > 0 else None

    @classmethod
    def from_view(cls, view, bins: int = 24, leaf_threshold: int = 4) -> "OracleScene":
        """From a pathtracingtutorial_amd HostSceneView (the scene's host arrays)."""
        vb = C.string_at(view.vertices, 64 * view.num_vertices) if view.num_vertices else b""
        fb = C.string_at(view.faces, 16 * view.num_faces) if view.num_faces else b""
        mb = C.string_at(view.materials, 256 * view.num_materials)
        lb = C.string_at(view.lights, 112 * view.num_lights) if view.num_lights else b""
        env = None
        if view.env_width > 0 and view.env_rgba:
            n = view.env_width * view.env_height * 4
            env = np.ctypeslib.as_array(view.env_rgba, shape=(n,)).reshape(view.env_height, view.env_width, 4).copy()
        return cls(vb, view.num_vertices, fb, view.num_faces, mb, view.num_materials, env, lb, view.num_lights,
                   bins, leaf_threshold)

    def add_texture(self, rgba: np.ndarray, wrap_s: int = 10497, wrap_t: int = 10497, linear: int = 0) -> int:
        rgba = np.ascontiguousarray(rgba, np.float32)
        self._tex_shapes.append((rgba.shape[0], rgba.shape[1]))
        return lib().oracle_scene_add_texture(self.h, _f(rgba), rgba.shape[1], rgba.shape[0], wrap_s, wrap_t, linear)

    def bvh(self):
        n = lib().oracle_bvh_num_nodes(self.h)
        nodes = np.zeros((n, 8), np.float32)
        order = np.zeros(self.nf, np.uint32)
        lib().oracle_bvh_get(self.h, _f(nodes), _u32(order))
        return nodes, order

    def trace(self, rays: np.ndarray, any_hit: bool = False):
        rays = np.ascontiguousarray(rays, np.float32)
        n = rays.shape[0]
        t = np.zeros(n, np.float32)
        face = np.zeros(n, np.int32)
        bary = np.zeros((n, 2), np.float32)
        lib().oracle_trace(self.h, _f(rays), n, 1 if any_hit else 0, _f(t), _i32(face), _f(bary))
        return t, face, bary

    def light_tables(self):
        ne, nenv = C.c_uint32(), C.c_uint32()
        lib().oracle_light_tables(self.h, C.byref(ne), C.byref(nenv), None, None, None, None)
        tp = np.zeros(self.nf, np.float32)
        ta = np.zeros((ne.value, 2), np.float32)
        ep = np.zeros(nenv.value, np.float32)
        ea = np.zeros((nenv.value, 2), np.float32)
        lib().oracle_light_tables(self.h, None, None, _f(tp), _f(ta), _f(ep), _f(ea))
        return {"tri_pmf": tp, "tri_alias": ta, "env_pmf": ep, "env_alias": ea}

    def env_pdf(self, d) -> float:
        d = np.ascontiguousarray(d, np.float32)
        return float(lib().oracle_env_pdf(self.h, _f(d)))

    def emissive_pdf(self, prim: int, dist: float, d) -> float:
        d = np.ascontiguousarray(d, np.float64)
        return float(lib().oracle_emissive_pdf(self.h, int(prim), float(dist), _d(d)))

    def sample_light(self, env: bool, p, n, x: int, y: int, frame: int, sample: int):
        """(wi, L, pdf) of sampleEnvLight (env) or sampleAreaLight at p with normal n, or None."""
        p = np.ascontiguousarray(p, np.float64)
        n = np.ascontiguousarray(n, np.float64)
        wi, L, pdf = np.zeros(3), np.zeros(3), C.c_double(0.0)
        ok = lib().oracle_sample_light(self.h, 1 if env else 0, _d(p), _d(n), x, y, frame, sample, _d(wi), _d(L),
                                       C.byref(pdf))
        return (wi, L, pdf.value) if ok else None

    def render(self, camera, spf: int, max_depth: int, sampler: int, seed: int = 0, frame: int = 0,
               rect=None, threads: int = 0, fp64: bool = False):
        """Batch-mean RGB of pixels in rect=(x0, x1, y0, y1) (default: whole image), shape (h, w, 3)."""
        x0, x1, y0, y1 = rect or (0, camera.width, 0, camera.height)
        out = np.zeros((y1 - y0, x1 - x0, 3), np.float32)
        cnt = np.zeros(3, np.uint64)
        lib().oracle_render(self.h, C.byref(camera), spf, max_depth, sampler, seed & 0xFFFFFFFF, frame, x0, x1, y0,
                            y1, threads, 1 if fp64 else 0, _f(out), cnt.ctypes.data_as(_U64P))
        return out, {"closest_rays": int(cnt[0]), "shadow_rays": int(cnt[1]), "paths": int(cnt[2])}

    def adjoint_fd(self, camera, spf: int, max_depth: int, sampler: int, seed: int, frame: int,
                   weights: np.ndarray, params, h: float = 1e-3, threads: int = 0):
        """Central differences of sum(weights * batch image) for params = [(kind, a, b, c), ...]
        (kind 0 baseColor, 1 roughness, 2 metallic, 3 texture texel, 4 env texel) of the
        detached estimator in double precision; returns (gradients, value at the current parameters)."""
        weights = np.ascontiguousarray(weights, np.float32).reshape(-1)
        assert weights.size == camera.width * camera.height * 3
        p = np.ascontiguousarray(np.array(params, np.int32).reshape(-1, 4))
        k, a, b, c = (np.ascontiguousarray(p[:, i]) for i in range(4))
        grad = np.zeros(len(p), np.float64)
        value = C.c_double(0.0)
        lib().oracle_adjoint_fd(self.h, C.byref(camera), spf, max_depth, sampler, seed & 0xFFFFFFFF, frame,
                                _f(weights), len(p), _i32(k), _i32(a), _i32(b), _i32(c), h, threads,
                                grad.ctypes.data_as(_DP), C.byref(value))
        return grad, float(value.value)

    def adjoint(self, camera, spf: int, max_depth: int, sampler: int, seed: int, frame: int, weights: np.ndarray,
                threads: int = 0):
        """Reverse-mode FP64 gradient of sum(weights * batch image) with respect to every
        parameter (the paths and detached estimator of adjoint_fd): dict of base_color (M, 3),
        roughness (M,), metallic (M,), textures (list of (H, W, 4) arrays), env ((He, We, 4)
        or None) and value."""
        weights = np.ascontiguousarray(weights, np.float32).reshape(-1)
        assert weights.size == camera.width * camera.height * 3
        nm = self._nm
        gm = np.zeros((nm, 5), np.float64)
        ntex = sum(h * w * 4 for (h, w) in self._tex_shapes)
        gt = np.zeros(max(1, ntex), np.float64)
        ge = np.zeros(self._env_shape, np.float64) if self._env_shape else None
        value = C.c_double(0.0)
        lib().oracle_adjoint(self.h, C.byref(camera), spf, max_depth, sampler, seed & 0xFFFFFFFF, frame,
                             _f(weights), threads, gm.ctypes.data_as(_DP), gt.ctypes.data_as(_DP),
                             ge.ctypes.data_as(_DP) if ge is not None else None, C.byref(value))
        tex, off = [], 0
        for (h, w) in self._tex_shapes:
            tex.append(gt[off:off + h * w * 4].reshape(h, w, 4))
            off += h * w * 4
        return {"base_color": gm[:, :3], "roughness": gm[:, 3], "metallic": gm[:, 4], "textures": tex, "env": ge,
                "value": float(value.value)}

    def fetch_material(self, mat, u: float, v: float) -> dict:
        """fetchMaterial of `mat` with this scene's textures at (u, v), in float."""
        out = np.zeros(20, np.float32)
        lib().oracle_fetch_material(self.h, C.byref(mat), float(u), float(v), _f(out))
        names = ["base", "metallic", "roughness", "spec_trans", "ior", "clearcoat", "clearcoat_roughness", "sheen",
                 "sheen_tint", "subsurface", "anisotropic", "sheen_color", "emissive", "specular_factor"]
        vals = [out[0:3], out[3], out[4], out[5], out[6], out[7], out[8], out[9], out[10], out[11], out[12],
                out[13:16], out[16:19], out[19]]
        return dict(zip(names, vals))

    def surface(self, o, d):
        """The integrator's shading point at the closest hit of (o, d): dict of t, P, sn (interpolated
        normal), gn (on the ray's side), N (after the normal map), entering, face; None on a miss."""
        o = np.ascontiguousarray(o, np.float32)
        d = np.ascontiguousarray(d, np.float32)
        out = np.zeros(16, np.float32)
        if not lib().oracle_surface(self.h, _f(o), _f(d), _f(out)):
            return None
        return {"t": float(out[1]), "P": out[2:5].copy(), "sn": out[5:8].copy(), "gn": out[8:11].copy(),
                "N": out[11:14].copy(), "entering": bool(out[14]), "face": int(out[15])}

    def __del__(self):
        try:
            if getattr(self, "h", None):
                lib().oracle_scene_destroy(self.h)
        except Exception:
            pass


def sobol_u32(px, py, index, dim, seed=0) -> int:
    return int(lib().oracle_sobol_u32(px, py, index, dim, seed))


def sobol_float(px, py, index, dim, seed=0) -> float:
    return float(lib().oracle_sobol_float(px, py, index, dim, seed))


def pcg(x, y, frame, sample, n, seed=0):
    u = np.zeros(n, np.uint32)
    f = np.zeros(n, np.float32)
    lib().oracle_pcg(x, y, frame, sample, seed, n, _u32(u), _f(f))
    return u, f


def alias_table(w: np.ndarray):
    w = np.ascontiguousarray(w, np.float32)
    prob = np.zeros(len(w), np.float32)
    alias = np.zeros(len(w), np.uint32)
    n = lib().oracle_alias_table(_f(w), len(w), _f(prob), _u32(alias))
    return prob[:n], alias[:n]


def camera_frame(cam):
    from pathtracingtutorial_amd._capi import CameraFrame
    f = CameraFrame()
    lib().oracle_camera_frame(C.byref(cam), C.byref(f))
    return f


def bsdf_eval(mat, wo, wi, n, entering=True, fp64=False):
    a = [np.ascontiguousarray(x, np.float64) for x in (wo, wi, n)]
    f = np.zeros(3, np.float64)
    pdf = np.zeros(1, np.float64)
    lib().oracle_bsdf_eval(C.byref(mat), _d(a[0]), _d(a[1]), _d(a[2]), 1 if entering else 0, 1 if fp64 else 0,
                           _d(f), _d(pdf))
    return f, float(pdf[0])


def bsdf_sample(mat, wo, n, u, entering=True):
    wo = np.ascontiguousarray(wo, np.float64)
    n = np.ascontiguousarray(n, np.float64)
    u = np.ascontiguousarray(u, np.float32)
    wi = np.zeros(3, np.float64)
    w = np.zeros(3, np.float64)
    pdf = np.zeros(1, np.float64)
    delta = np.zeros(1, np.int32)
    ok = lib().oracle_bsdf_sample(C.byref(mat), _d(wo), _d(n), 1 if entering else 0, _f(u), _d(wi), _d(w), _d(pdf),
                                  _i32(delta))
    return bool(ok), wi, w, float(pdf[0]), bool(delta[0])


def dir_to_uv(d):
    d = np.ascontiguousarray(d, np.float32)
    uv = np.zeros(2, np.float32)
    lib().oracle_dir_to_uv(_f(d), _f(uv))
    return uv


def uv_to_dir(uv):
    uv = np.ascontiguousarray(uv, np.float32)
    d = np.zeros(3, np.float32)
    lib().oracle_uv_to_dir(_f(uv), _f(d))
    return d


def tex2d(rgba: np.ndarray, u: float, v: float, wrap_s=10497, wrap_t=10497):
    rgba = np.ascontiguousarray(rgba, np.float32)
    out = np.zeros(4, np.float32)
    lib().oracle_tex2d(_f(rgba), rgba.shape[1], rgba.shape[0], wrap_s, wrap_t, u, v, _f(out))
    return out


def microfacet(inputs, f0, wo, n) -> np.ndarray:
    """The oracle's microfacet / sampling / frame pieces (27 floats, see oracle.h)."""
    a = [np.ascontiguousarray(x, np.float32) for x in (inputs, f0, wo, n)]
    out = np.zeros(27, np.float32)
    lib().oracle_microfacet(_f(a[0]), _f(a[1]), _f(a[2]), _f(a[3]), _f(out))
    return out
