"""Differentiable rendering over the C ABI.

`render(dr, frame, base_color, roughness, metallic, texture, env)` is a
torch.autograd.Function: the forward pass writes the parameters into the
uploaded scene and renders one batch (`pt_render_batch`, the plain
integrator) with the random numbers of `frame`; the backward pass runs the
path-replay adjoint (`pt_render_adjoint`) of the same batch with
d loss / d image and returns d loss / d parameter.  Tensors live on the GPU
(torch's ROCm device 0, the device the library renders on) and are handed
to the library as device pointers; no CPU copy of images or textures.

torch carries its own copy of the HIP runtime, and it only finds the GPU if
it initialises before the library's runtime: importing this module calls
torch.cuda.init(), so import it before the first scene is built or
uploaded.

Parameters (any may be None, then it is left as it is and gets no gradient):
  base_color (M, 3), roughness (M,), metallic (M,): material factors
  texture    (H, W, 4): texels of one texture (`DiffRenderer.texture_id`)
  env        (He, We, 4): environment texels
The 4th channel of textures and images carries no gradient.  The env's
importance-sampling tables keep the distribution of the uploaded
environment (`rebuild_env_tables()` refreshes them); the estimator stays
unbiased either way.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _capi
from ._capi import check, lib

ENV = 0xFFFFFFFF

try:
    torch.cuda.init()
except RuntimeError:  # no GPU here: the module still imports (for inspection)
    pass


class DiffRenderer:
    """Owns a film for `scene` and the gradient buffers of its parameters.
    integrator: the forward's form (the wavefront by default; the megakernel,
    PT_INTEGRATOR_PATH, renders the same image).  keep_radiance: a wavefront
    forward keeps each path's radiance (16 bytes per path) so that the adjoint
    of the same batch skips its primal pass (pt_scene_set_keep_radiance)."""

    def __init__(self, scene, texture_id: int = 0, integrator: int = _capi.PT_INTEGRATOR_WAVEFRONT,
                 keep_radiance: bool = True):
        self.scene = scene
        self.integrator = int(integrator)
        check(lib().pt_scene_set_keep_radiance(scene.handle, 1 if keep_radiance else 0))
        info = scene.info()
        self.num_materials = info.num_materials
        self.width, self.height = info.width, info.height
        self.texture_id = texture_id if info.num_textures > texture_id >= 0 else -1
        self.env_shape = (info.env_height, info.env_width, 4) if info.env_width > 0 else None
        self._film = C.c_void_p()
        check(lib().pt_film_create(self.width, self.height, C.byref(self._film)))
        self.tex_offset, self.tex_count, self.pool_texels = 0, 0, 0
        for t in range(info.num_textures):
            off, cnt = C.c_uint64(), C.c_uint64()
            check(lib().pt_scene_texture_offset(scene.handle, t, C.byref(off), C.byref(cnt)))
            if t == self.texture_id:
                self.tex_offset, self.tex_count = off.value, cnt.value
            self.pool_texels = max(self.pool_texels, off.value + cnt.value)
        self.last_forward_ms = 0.0
        self.last_backward_ms = 0.0
        self.last_backward_reused = False

    # -- parameter transfer ------------------------------------------------------
    def push(self, base_color=None, roughness=None, metallic=None, texture=None, env=None) -> None:
        torch.cuda.synchronize()
        h = self.scene.handle

        def host(t, n):
            if t is None:
                return None
            a = np.ascontiguousarray(t.detach().float().cpu().numpy().reshape(-1), np.float32)
            assert a.size == n, (a.size, n)
            return a

        bc, ro, me = host(base_color, 3 * self.num_materials), host(roughness, self.num_materials), \
            host(metallic, self.num_materials)
        if bc is not None or ro is not None or me is not None:
            fp = lambda a: a.ctypes.data_as(C.POINTER(C.c_float)) if a is not None else None
            check(lib().pt_scene_update_materials(h, fp(bc), fp(ro), fp(me)))
        if texture is not None:
            assert self.texture_id >= 0, "the scene has no texture to differentiate"
            t = texture.detach().float().contiguous()
            assert t.is_cuda and t.numel() == 4 * self.tex_count
            check(lib().pt_scene_update_texture(h, self.texture_id, C.c_void_p(t.data_ptr()), t.numel(), 1))
        if env is not None:
            e = env.detach().float().contiguous()
            assert e.is_cuda and self.env_shape is not None and tuple(e.shape) == self.env_shape
            check(lib().pt_scene_update_texture(h, ENV, C.c_void_p(e.data_ptr()), e.numel(), 1))

    def rebuild_env_tables(self, on_device: bool = True) -> None:
        """Environment sampling tables for the current texels: built on the GPU
        (default; same distribution, another alias layout) or on the host
        (the reference's Vose table)."""
        if on_device:
            check(lib().pt_scene_build_env_tables_device(self.scene.handle))
        else:
            check(lib().pt_scene_rebuild_env_tables(self.scene.handle))

    # -- the two passes ------------------------------------------------------------
    def forward_image(self, frame: int) -> torch.Tensor:
        check(lib().pt_scene_set_frame(self.scene.handle, int(frame)))
        ms = C.c_float(0.0)
        check(lib().pt_render_batch(self.scene.handle, self._film, self.integrator, C.byref(ms)))
        self.last_forward_ms = ms.value
        img = torch.empty((self.height, self.width, 4), dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        check(lib().pt_film_copy_device(self._film, 1, C.c_void_p(img.data_ptr())))
        return img

    def adjoint(self, frame: int, image_grad: torch.Tensor, want=(True, True, True, True, True),
                replay: bool = False):
        """d loss / d (base_color, roughness, metallic, texture, env) given d loss / d image."""
        g = image_grad.detach().float().contiguous()
        assert g.is_cuda and tuple(g.shape) == (self.height, self.width, 4)
        M = self.num_materials
        z = lambda *s: torch.zeros(s, dtype=torch.float32, device="cuda")
        d_base = z(M, 3) if want[0] else None
        d_rough = z(M) if want[1] else None
        d_metal = z(M) if want[2] else None
        d_pool = z(max(1, self.pool_texels), 4) if (want[3] and self.texture_id >= 0) else None
        d_env = z(*self.env_shape) if (want[4] and self.env_shape is not None) else None
        rep = z(self.height, self.width, 4) if replay else None
        ptr = lambda t: C.c_void_p(t.data_ptr()) if t is not None else None
        grads = _capi.AdjointGrads(ptr(d_base), ptr(d_rough), ptr(d_metal), ptr(d_pool), ptr(d_env))
        torch.cuda.synchronize()
        ms = C.c_float(0.0)
        check(lib().pt_render_adjoint(self.scene.handle, int(frame), C.c_void_p(g.data_ptr()), C.byref(grads),
                                      ptr(rep), C.byref(ms)))
        self.last_backward_ms = ms.value
        reused = C.c_int(0)
        check(lib().pt_scene_get_keep_radiance(self.scene.handle, None, C.byref(reused)))
        self.last_backward_reused = bool(reused.value)
        d_tex = None
        if d_pool is not None:
            d_tex = d_pool[self.tex_offset:self.tex_offset + self.tex_count]
        out = {"base_color": d_base, "roughness": d_rough, "metallic": d_metal, "texture": d_tex, "env": d_env}
        if replay:
            out["replay"] = rep
        return out

    def render(self, frame: int, base_color=None, roughness=None, metallic=None, texture=None, env=None):
        """Differentiable batch image (H, W, 4), row 0 = bottom image row."""
        return _Render.apply(self, int(frame), base_color, roughness, metallic, texture, env)

    def __del__(self):
        try:
            if self._film:
                lib().pt_film_destroy(self._film)
        except Exception:
            pass


class _Render(torch.autograd.Function):
    @staticmethod
    def forward(ctx, dr, frame, base_color, roughness, metallic, texture, env):
        dr.push(base_color, roughness, metallic, texture, env)
        ctx.dr, ctx.frame = dr, frame
        ctx.shapes = [None if t is None else tuple(t.shape) for t in (base_color, roughness, metallic, texture, env)]
        ctx.want = [t is not None and t.requires_grad for t in (base_color, roughness, metallic, texture, env)]
        return dr.forward_image(frame)

    @staticmethod
    def backward(ctx, grad):
        if not any(ctx.want):
            return (None,) * 7
        g = ctx.dr.adjoint(ctx.frame, grad, want=ctx.want)
        outs = []
        for key, shape, w in zip(("base_color", "roughness", "metallic", "texture", "env"), ctx.shapes, ctx.want):
            t = g[key] if w else None
            outs.append(t.reshape(shape) if t is not None else None)
        return (None, None, *outs)


class MultiViewRenderer:
    """Several cameras of one scene: forward in one wavefront schedule
    (pt_render_views), backward in one adjoint launch (pt_render_adjoint_views).
    All cameras must have the same resolution (the views are stacked).
    keep_radiance: the forward keeps each path's radiance (16 bytes per path)
    so that the adjoint of the same batch skips its primal pass
    (pt_scene_set_keep_radiance)."""

    def __init__(self, scene, cameras, texture_id: int = 0, keep_radiance: bool = True):
        self.base = DiffRenderer(scene, texture_id)
        self.scene = scene
        check(lib().pt_scene_set_keep_radiance(scene.handle, 1 if keep_radiance else 0))
        n = len(cameras)
        assert n > 0 and len({(c.width, c.height) for c in cameras}) == 1
        self.n = n
        self.cams = (_capi.Camera * n)(*cameras)
        self.width, self.height = cameras[0].width, cameras[0].height
        self.films = []
        for c in cameras:
            f = C.c_void_p()
            check(lib().pt_film_create(c.width, c.height, C.byref(f)))
            self.films.append(f)
        self._film_arr = (C.c_void_p * n)(*[f.value for f in self.films])
        self.last_forward_ms = 0.0
        self.last_backward_ms = 0.0
        self.last_backward_reused = False

    def push(self, *args, **kw):
        self.base.push(*args, **kw)

    def _frames(self, frame):
        return (C.c_uint32 * self.n)(*[int(frame) * self.n + i for i in range(self.n)])

    def forward_images(self, frame: int) -> torch.Tensor:
        ms = C.c_float(0.0)
        check(lib().pt_render_views(self.scene.handle, self.n, self.cams, self._film_arr, self._frames(frame),
                                    C.byref(ms)))
        self.last_forward_ms = ms.value
        out = torch.empty((self.n, self.height, self.width, 4), dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        for i, f in enumerate(self.films):
            check(lib().pt_film_copy_device(f, 1, C.c_void_p(out[i].data_ptr())))
        return out

    def adjoint(self, frame: int, image_grads: torch.Tensor, want=(True, True, True, True, True)):
        g = image_grads.detach().float().contiguous()
        assert g.is_cuda and tuple(g.shape) == (self.n, self.height, self.width, 4)
        b = self.base
        M = b.num_materials
        z = lambda *s: torch.zeros(s, dtype=torch.float32, device="cuda")
        d_base = z(M, 3) if want[0] else None
        d_rough = z(M) if want[1] else None
        d_metal = z(M) if want[2] else None
        d_pool = z(max(1, b.pool_texels), 4) if (want[3] and b.texture_id >= 0) else None
        d_env = z(*b.env_shape) if (want[4] and b.env_shape is not None) else None
        ptr = lambda t: C.c_void_p(t.data_ptr()) if t is not None else None
        grads = _capi.AdjointGrads(ptr(d_base), ptr(d_rough), ptr(d_metal), ptr(d_pool), ptr(d_env))
        gp = (C.c_void_p * self.n)(*[g[i].data_ptr() for i in range(self.n)])
        torch.cuda.synchronize()
        ms = C.c_float(0.0)
        check(lib().pt_render_adjoint_views(self.scene.handle, self.n, self.cams, self._frames(frame), gp,
                                            C.byref(grads), None, C.byref(ms)))
        self.last_backward_ms = ms.value
        reused = C.c_int(0)
        check(lib().pt_scene_get_keep_radiance(self.scene.handle, None, C.byref(reused)))
        self.last_backward_reused = bool(reused.value)
        d_tex = d_pool[b.tex_offset:b.tex_offset + b.tex_count] if d_pool is not None else None
        return {"base_color": d_base, "roughness": d_rough, "metallic": d_metal, "texture": d_tex, "env": d_env}

    def render(self, frame: int, base_color=None, roughness=None, metallic=None, texture=None, env=None):
        """Differentiable (views, H, W, 4) images of batch `frame` (view i uses frame * views + i)."""
        return _RenderViews.apply(self, int(frame), base_color, roughness, metallic, texture, env)

    def __del__(self):
        try:
            for f in self.films:
                if f:
                    lib().pt_film_destroy(f)
        except Exception:
            pass


class _RenderViews(torch.autograd.Function):
    @staticmethod
    def forward(ctx, mv, frame, base_color, roughness, metallic, texture, env):
        mv.push(base_color, roughness, metallic, texture, env)
        ctx.mv, ctx.frame = mv, frame
        ctx.shapes = [None if t is None else tuple(t.shape) for t in (base_color, roughness, metallic, texture, env)]
        ctx.want = [t is not None and t.requires_grad for t in (base_color, roughness, metallic, texture, env)]
        return mv.forward_images(frame)

    @staticmethod
    def backward(ctx, grad):
        if not any(ctx.want):
            return (None,) * 7
        g = ctx.mv.adjoint(ctx.frame, grad, want=ctx.want)
        outs = []
        for key, shape, w in zip(("base_color", "roughness", "metallic", "texture", "env"), ctx.shapes, ctx.want):
            t = g[key] if w else None
            outs.append(t.reshape(shape) if t is not None else None)
        return (None, None, *outs)


def ring_cameras(n: int = 16, radius: float = 8.0, height: float = 1.5, fov: float = 35.0, size: int = 512):
    """n cameras on a horizontal ring around the origin, looking at it."""
    import math
    cams = []
    for i in range(n):
        a = 2.0 * math.pi * i / n
        c = _capi.Camera()
        c.position.x, c.position.y, c.position.z = radius * math.cos(a), height, radius * math.sin(a)
        c.look_at.x = c.look_at.y = c.look_at.z = 0.0
        c.up.x, c.up.y, c.up.z = 0.0, 1.0, 0.0
        c.fov = fov
        c.width = c.height = size
        c.aperture = 0.0
        c.focal_distance = radius
        cams.append(c)
    return cams
