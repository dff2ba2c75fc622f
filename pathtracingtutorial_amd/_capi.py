"""ctypes binding of libptamd.so (the C ABI declared in include/pt_amd.h).

The library is loaded from the package directory.  There is no fallback: if
the shared library is missing, importing the renderer raises, naming the
build command.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libptamd.so")

PT_OK = 0
STATUS_NAMES = {
    0: "PT_OK", 1: "PT_ERR_INVALID_ARG", 2: "PT_ERR_IO", 3: "PT_ERR_PARSE", 4: "PT_ERR_DEVICE",
    5: "PT_ERR_STATE", 6: "PT_ERR_UNSUPPORTED", 7: "PT_ERR_OUT_OF_MEMORY", 8: "PT_ERR_LIMIT",
}
PT_INTEGRATOR_PATH = 0
PT_INTEGRATOR_WAVEFRONT = 1


class PtError(RuntimeError):
    """A failed C-ABI call: .status is the pt_status code."""

    def __init__(self, status: int, msg: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


class Vec3p(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float), ("_pad", C.c_float)]

    @classmethod
    def of(cls, v) -> "Vec3p":
        return cls(float(v[0]), float(v[1]), float(v[2]), 0.0)

    def tolist(self):
        return [self.x, self.y, self.z]


class Vertex(C.Structure):
    _fields_ = [("position", Vec3p), ("normal", Vec3p), ("u", C.c_float), ("v", C.c_float),
                ("_pad", C.c_float * 2), ("tangent", C.c_float * 4)]


class Face(C.Structure):
    _fields_ = [("v0", C.c_uint32), ("v1", C.c_uint32), ("v2", C.c_uint32), ("material_id", C.c_int32)]


class Material(C.Structure):
    _fields_ = [
        ("base_color", Vec3p), ("metallic", C.c_float),
        ("roughness", C.c_float), ("specular_tint", C.c_float), ("anisotropic", C.c_float), ("sheen", C.c_float),
        ("sheen_tint", C.c_float), ("clearcoat", C.c_float), ("clearcoat_roughness", C.c_float),
        ("spec_trans", C.c_float),
        ("ior", C.c_float), ("subsurface", C.c_float), ("specular_factor", C.c_float),
        ("specular_color_r", C.c_float),
        ("specular_color_g", C.c_float), ("specular_color_b", C.c_float), ("anisotropy_rotation", C.c_float),
        ("normal_scale", C.c_float),
        ("_pad_e", C.c_float * 3),
        ("emissive", Vec3p), ("emissive_strength", C.c_float), ("alpha_cutoff", C.c_float),
        ("alpha_mode", C.c_int32), ("base_alpha", C.c_float), ("double_sided", C.c_int32),
        ("sheen_color_r", C.c_float), ("sheen_color_g", C.c_float), ("sheen_color_b", C.c_float),
        ("occlusion_strength", C.c_float),
        ("thickness_factor", C.c_float), ("attenuation_distance", C.c_float), ("att_color_r", C.c_float),
        ("att_color_g", C.c_float), ("att_color_b", C.c_float),
        ("uv_offset_x", C.c_float), ("uv_offset_y", C.c_float), ("uv_scale_x", C.c_float),
        ("uv_scale_y", C.c_float), ("uv_rotation", C.c_float), ("_pad1", C.c_float), ("_pad2", C.c_float),
        ("base_color_tex_id", C.c_int32), ("metallic_roughness_tex_id", C.c_int32), ("normal_tex_id", C.c_int32),
        ("emissive_tex_id", C.c_int32),
        ("occlusion_tex_id", C.c_int32), ("clearcoat_tex_id", C.c_int32), ("clearcoat_roughness_tex_id", C.c_int32),
        ("clearcoat_normal_tex_id", C.c_int32),
        ("transmission_tex_id", C.c_int32), ("sheen_color_tex_id", C.c_int32), ("sheen_roughness_tex_id", C.c_int32),
        ("_pad3", C.c_int32),
        ("_pad_tail", C.c_int32 * 3),
    ]


class PunctualLight(C.Structure):
    _fields_ = [("position", Vec3p), ("range", C.c_float), ("_pad0", C.c_float * 3), ("color", Vec3p),
                ("intensity", C.c_float), ("_pad1", C.c_float * 3), ("direction", Vec3p),
                ("inner_cone", C.c_float), ("outer_cone", C.c_float), ("type", C.c_int32),
                ("_pad2", C.c_float * 2), ("_tail", C.c_float * 3)]


class Camera(C.Structure):
    _fields_ = [("position", Vec3p), ("look_at", Vec3p), ("up", Vec3p), ("fov", C.c_float),
                ("aperture", C.c_float), ("focal_distance", C.c_float), ("width", C.c_int32),
                ("height", C.c_int32), ("_pad", C.c_float * 3)]


class CameraFrame(C.Structure):
    _fields_ = [("position", C.c_float * 3), ("forward", C.c_float * 3), ("right", C.c_float * 3),
                ("up", C.c_float * 3), ("tan_half_fov", C.c_float), ("aspect", C.c_float),
                ("aperture", C.c_float), ("focal_distance", C.c_float)]


class TextureDesc(C.Structure):
    _fields_ = [("rgba", C.POINTER(C.c_float)), ("width", C.c_int32), ("height", C.c_int32),
                ("wrap_s", C.c_int32), ("wrap_t", C.c_int32), ("linear", C.c_int32), ("_pad", C.c_int32)]


class RenderSettings(C.Structure):
    _fields_ = [("sample_per_frame", C.c_int32), ("target_samples", C.c_int32), ("max_depth", C.c_int32),
                ("sampler", C.c_int32), ("seed", C.c_uint32), ("_pad", C.c_int32 * 3)]


class HostSceneDesc(C.Structure):
    _fields_ = [("vertices", C.POINTER(Vertex)), ("num_vertices", C.c_uint32),
                ("faces", C.POINTER(Face)), ("num_faces", C.c_uint32),
                ("materials", C.POINTER(Material)), ("num_materials", C.c_uint32),
                ("textures", C.POINTER(TextureDesc)), ("num_textures", C.c_uint32),
                ("lights", C.POINTER(PunctualLight)), ("num_lights", C.c_uint32),
                ("camera", C.POINTER(Camera)),
                ("env_rgba", C.POINTER(C.c_float)), ("env_width", C.c_int32), ("env_height", C.c_int32),
                ("settings", C.POINTER(RenderSettings))]


class HostSceneView(C.Structure):
    _fields_ = [("vertices", C.POINTER(Vertex)), ("num_vertices", C.c_uint32),
                ("faces", C.POINTER(Face)), ("num_faces", C.c_uint32),
                ("materials", C.POINTER(Material)), ("num_materials", C.c_uint32),
                ("lights", C.POINTER(PunctualLight)), ("num_lights", C.c_uint32),
                ("env_rgba", C.POINTER(C.c_float)), ("env_width", C.c_int32), ("env_height", C.c_int32),
                ("num_textures", C.c_uint32),
                ("camera", Camera), ("settings", RenderSettings)]


class BvhOptions(C.Structure):
    _fields_ = [("bins", C.c_int32), ("leaf_threshold", C.c_int32), ("max_leaf", C.c_int32),
                ("builder", C.c_int32)]


class SceneInfo(C.Structure):
    _fields_ = [("num_vertices", C.c_uint32), ("num_faces", C.c_uint32), ("num_materials", C.c_uint32),
                ("num_textures", C.c_uint32), ("num_lights", C.c_uint32), ("num_emissive_triangles", C.c_uint32),
                ("num_bvh_nodes", C.c_uint32), ("bvh_depth", C.c_uint32), ("env_width", C.c_int32),
                ("env_height", C.c_int32), ("width", C.c_int32), ("height", C.c_int32),
                ("sample_per_frame", C.c_int32), ("target_samples", C.c_int32), ("max_depth", C.c_int32),
                ("sampler", C.c_int32), ("seed", C.c_uint32), ("frame_count", C.c_uint32),
                ("bvh_build_ms", C.c_float), ("max_leaf_size", C.c_uint32)]


class Counters(C.Structure):
    _fields_ = [("closest_rays", C.c_uint64), ("shadow_rays", C.c_uint64), ("node_visits", C.c_uint64),
                ("tri_tests", C.c_uint64), ("paths", C.c_uint64), ("stack_overflows", C.c_uint64),
                ("shadow_node_visits", C.c_uint64), ("shadow_tri_tests", C.c_uint64),
                ("shade_items", C.c_uint64), ("shade_env_lookups", C.c_uint64), ("shade_hits", C.c_uint64),
                ("shade_light_samples", C.c_uint64), ("shade_shadow_entries", C.c_uint64),
                ("shade_continuations", C.c_uint64), ("shade_radiance_adds", C.c_uint64)]


class WavefrontConfig(C.Structure):
    _fields_ = [("chunk_spp", C.c_int32), ("extend_block", C.c_int32), ("shade_block", C.c_int32),
                ("waves_per_simd", C.c_int32), ("sort_by_material", C.c_int32), ("use_graph", C.c_int32),
                ("time_stages", C.c_int32), ("shade_waves", C.c_int32), ("stack_lds", C.c_int32),
                ("grab_extend", C.c_int32), ("grab_shade", C.c_int32), ("grab_shadow", C.c_int32),
                ("bvh_width", C.c_int32)]


class WavefrontStats(C.Structure):
    _fields_ = [("stage_ms", C.c_float * 6), ("launches", C.c_uint32 * 6)]


WAVEFRONT_STAGES = ("generate", "extend", "sort", "shade", "shadow", "control")


class Job(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("film", C.c_void_p)]


class AdjointGrads(C.Structure):
    _fields_ = [("base_color", C.c_void_p), ("roughness", C.c_void_p), ("metallic", C.c_void_p),
                ("textures", C.c_void_p), ("env", C.c_void_p)]


# Every exported symbol with its signature (kept in sync with include/pt_amd.h).
_P = C.c_void_p
_PP = C.POINTER(C.c_void_p)
_U32P = C.POINTER(C.c_uint32)
_I32P = C.POINTER(C.c_int32)
_FP = C.POINTER(C.c_float)
SIGNATURES = {
    "pt_abi_version": (C.c_int, []),
    "pt_last_error": (C.c_char_p, []),
    "pt_init": (C.c_int, [C.c_int]),
    "pt_struct_sizes": (C.c_int, [_U32P, C.c_int]),
    "pt_scene_load": (C.c_int, [C.c_char_p, _PP]),
    "pt_scene_create": (C.c_int, [C.POINTER(HostSceneDesc), _PP]),
    "pt_scene_generate": (C.c_int, [C.c_int, C.c_uint64, _PP]),
    "pt_scene_destroy": (None, [_P]),
    "pt_scene_get_info": (C.c_int, [_P, C.POINTER(SceneInfo)]),
    "pt_scene_build_accel": (C.c_int, [_P, C.POINTER(BvhOptions)]),
    "pt_scene_upload": (C.c_int, [_P]),
    "pt_scene_set_spp": (C.c_int, [_P, C.c_int]),
    "pt_scene_set_sample_per_frame": (C.c_int, [_P, C.c_int]),
    "pt_scene_set_max_depth": (C.c_int, [_P, C.c_int]),
    "pt_scene_set_sampler": (C.c_int, [_P, C.c_int]),
    "pt_scene_set_seed": (C.c_int, [_P, C.c_uint32]),
    "pt_scene_set_camera": (C.c_int, [_P, C.POINTER(Camera)]),
    "pt_scene_get_camera": (C.c_int, [_P, C.POINTER(Camera)]),
    "pt_scene_camera_frame": (C.c_int, [_P, C.POINTER(CameraFrame)]),
    "pt_scene_host_view": (C.c_int, [_P, C.POINTER(HostSceneView)]),
    "pt_scene_texture": (C.c_int, [_P, C.c_uint32, C.POINTER(TextureDesc)]),
    "pt_scene_set_texture_data": (C.c_int, [_P, C.c_uint32, _FP, C.c_size_t]),
    "pt_scene_get_bvh": (C.c_int, [_P, C.c_int, _FP, _U32P, _U32P, _U32P]),
    "pt_scene_get_light_tables": (C.c_int, [_P, _FP, _U32P, _FP, _U32P, _FP, _FP, _U32P]),
    "pt_bvh_collapse_wide": (C.c_int, [_FP, C.c_uint32, _FP, _U32P, _U32P]),
    "pt_scene_get_wide_bvh": (C.c_int, [C.c_void_p, _FP, _U32P, _U32P]),
    "pt_film_create": (C.c_int, [C.c_int, C.c_int, _PP]),
    "pt_film_destroy": (None, [_P]),
    "pt_film_reset": (C.c_int, [_P, _P]),
    "pt_film_download": (C.c_int, [_P, C.c_int, _FP]),
    "pt_film_device_ptr": (C.c_int, [_P, C.c_int, _PP]),
    "pt_film_copy_device": (C.c_int, [_P, C.c_int, C.c_void_p]),
    "pt_image_write": (C.c_int, [C.c_char_p, _FP, C.c_int, C.c_int]),
    "pt_scene_output_path": (C.c_int, [_P, C.c_char_p, C.c_size_t]),
    "pt_image_read_hdr": (C.c_int, [C.c_char_p, _FP, _I32P, _I32P]),
    "pt_image_read": (C.c_int, [C.c_char_p, _FP, _I32P, _I32P]),
    "pt_render_batch": (C.c_int, [_P, _P, C.c_int, _FP]),
    "pt_render_all": (C.c_int, [_P, _P, C.c_int, _FP]),
    "pt_render_jobs": (C.c_int, [C.POINTER(Job), C.c_int, _FP]),
    "pt_scene_set_wavefront_config": (C.c_int, [_P, C.POINTER(WavefrontConfig)]),
    "pt_scene_get_wavefront_config": (C.c_int, [_P, C.POINTER(WavefrontConfig)]),
    "pt_scene_tune_wavefront": (C.c_int, [_P, C.c_int, C.POINTER(WavefrontConfig), _FP]),
    "pt_scene_wavefront_stats": (C.c_int, [_P, C.POINTER(WavefrontStats)]),
    "pt_render_adjoint": (C.c_int, [_P, C.c_uint32, C.c_void_p, C.POINTER(AdjointGrads), C.c_void_p, _FP]),
    "pt_scene_set_frame": (C.c_int, [_P, C.c_uint32]),
    "pt_scene_set_adjoint_integrator": (C.c_int, [_P, C.c_int]),
    "pt_scene_get_adjoint_integrator": (C.c_int, [_P, C.POINTER(C.c_int)]),
    "pt_render_views": (C.c_int, [_P, C.c_int, C.POINTER(Camera), C.POINTER(C.c_void_p), _U32P, _FP]),
    "pt_render_adjoint_views": (C.c_int, [_P, C.c_int, C.POINTER(Camera), _U32P, C.POINTER(C.c_void_p),
                                          C.POINTER(AdjointGrads), C.POINTER(C.c_void_p), _FP]),
    "pt_scene_texture_offset": (C.c_int, [_P, C.c_uint32, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "pt_scene_update_materials": (C.c_int, [_P, _FP, _FP, _FP]),
    "pt_scene_update_texture": (C.c_int, [_P, C.c_uint32, C.c_void_p, C.c_size_t, C.c_int]),
    "pt_scene_rebuild_env_tables": (C.c_int, [_P]),
    "pt_scene_build_env_tables_device": (C.c_int, [_P]),
    "pt_scene_update_emission": (C.c_int, [_P, _FP, _FP]),
    "pt_scene_rebuild_emissive_tables": (C.c_int, [_P]),
    "pt_scene_build_emissive_tables_device": (C.c_int, [_P]),
    "pt_alias_build_device": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p]),
    "pt_alias_build_device_noalloc": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p]),
    "pt_counters_enable": (C.c_int, [_P, C.c_int]),
    "pt_counters_get": (C.c_int, [_P, C.POINTER(Counters)]),
    "pt_counters_reset": (C.c_int, [_P]),
    "pt_probe_sampler": (C.c_int, [C.c_int, C.c_uint32, _I32P, C.c_int, C.c_int, _U32P, _FP]),
    "pt_trace_rays": (C.c_int, [_P, _FP, C.c_int, C.c_int, _FP, _I32P, _FP]),
    "pt_probe_copy": (C.c_int, [C.c_size_t, C.c_int, _FP]),
    "pt_probe_reciprocal": (C.c_int, [C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "pt_scene_set_keep_radiance": (C.c_int, [_P, C.c_int]),
    "pt_scene_get_keep_radiance": (C.c_int, [_P, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "pt_probe_bsdf": (C.c_int, [C.c_void_p, C.c_int, _FP, _I32P, _FP, _FP]),
}

_lib = None


def lib() -> C.CDLL:
    """The loaded libptamd.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m pathtracingtutorial_amd.build`")
        h = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(h, name)
            fn.restype = res
            fn.argtypes = args
        _lib = h
    return _lib


def check(status: int) -> None:
    if status != PT_OK:
        msg = lib().pt_last_error()
        raise PtError(status, msg.decode() if msg else "")


def fptr(a: np.ndarray):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(_FP)


def u32ptr(a: np.ndarray):
    assert a.dtype == np.uint32 and a.flags.c_contiguous
    return a.ctypes.data_as(_U32P)


def i32ptr(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_I32P)
