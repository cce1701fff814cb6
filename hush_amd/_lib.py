"""ctypes declarations of the C ABI in include/*.h (libhush.so, built in-tree by hush_amd/build.py).

Loading fails loudly when the library is missing: there is no CPU fallback for the denoiser.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("HUSH_LIB") or os.path.join(HERE, "libhush.so")
SCENE_LIB_PATH = os.path.join(HERE, "libhush_scene.so")

# every symbol include/hush.h declares
EXPORTS = [
    "hush_hidden_of", "hush_create", "hush_destroy", "hush_resize", "hush_reset", "hush_set_exposure",
    "hush_get_exposure", "hush_get_size", "hush_hidden", "hush_frame", "hush_set_precision", "hush_get_precision",
    "hush_set_fp8_scales", "hush_get_fp8_scales",
    "hush_set_ieee_math", "hush_get_ieee_math", "hush_set_compensated", "hush_get_compensated", "hush_mlp_forward",
    "hush_denoise", "hush_encode", "hush_output", "hush_last_error", "hush_set_timing", "hush_pass_times",
    "hush_debug_surface", "hush_debug_set_state", "hush_debug_run_pass", "hush_set_debug_mask",
    "hush_debug_wide_gathers", "hush_debug_math_selftest",
    # include/svgf.h
    "svgf_default_options", "svgf_create", "svgf_destroy", "svgf_resize", "svgf_reset", "svgf_get_size", "svgf_frame",
    "svgf_get_options", "svgf_encode", "svgf_output", "svgf_last_error", "svgf_set_timing", "svgf_pass_times",
    "svgf_debug_surface", "svgf_debug_set_state", "svgf_debug_run_pass",
    # include/hush_train.h
    "hush_train_create", "hush_train_destroy", "hush_train_last_error", "hush_train_prepare", "hush_train_forward",
    "hush_train_loss", "hush_train_backward", "hush_train_commit", "hush_train_a0", "hush_train_set_a0_fixed",
    "hush_train_set_forward_precision", "hush_train_mlp_out", "hush_loss_noalloc",
    # include/hush_scene.h
    "hush_scene_create", "hush_scene_destroy", "hush_scene_last_error", "hush_scene_render",
    # include/post.h
    "post_default_options", "post_create", "post_destroy", "post_resize", "post_encode", "post_last_error",
    "post_debug_surface",
]


class hush_surface(ctypes.Structure):
    _fields_ = [("ptr", ctypes.c_void_p), ("pitch", ctypes.c_size_t), ("width", ctypes.c_int32),
                ("height", ctypes.c_int32), ("format", ctypes.c_int32)]


class hush_inputs(ctypes.Structure):
    _fields_ = [("radiance", hush_surface), ("normal_depth", hush_surface), ("motion", hush_surface),
                ("material", hush_surface), ("emission", hush_surface)]


class hush_scene_frame(ctypes.Structure):
    _fields_ = [("cams", ctypes.c_void_p), ("prims", ctypes.c_void_p), ("off", ctypes.c_void_p), ("lst", ctypes.c_void_p),
                ("n_prims", ctypes.c_int32), ("n_list", ctypes.c_int32), ("seed", ctypes.c_uint64), ("frame", ctypes.c_int32),
                ("width", ctypes.c_int32), ("height", ctypes.c_int32), ("x0", ctypes.c_int32), ("y0", ctypes.c_int32),
                ("w", ctypes.c_int32), ("h", ctypes.c_int32)]


class svgf_options(ctypes.Structure):
    _fields_ = [("sigma_l", ctypes.c_float), ("sigma_n", ctypes.c_float), ("sigma_z", ctypes.c_float),
                ("alpha_color", ctypes.c_float), ("alpha_moments", ctypes.c_float), ("taa_alpha", ctypes.c_float),
                ("taa", ctypes.c_int32), ("iterations", ctypes.c_int32)]


class svgf_inputs(ctypes.Structure):
    _fields_ = [("radiance", hush_surface), ("normal_depth", hush_surface), ("motion", hush_surface),
                ("material", hush_surface), ("emission", hush_surface), ("prev_normal_depth", hush_surface),
                ("prev_motion", hush_surface)]


class post_options(ctypes.Structure):
    _fields_ = [("bloom", ctypes.c_float), ("vignette", ctypes.c_float)]


class post_args(ctypes.Structure):
    _fields_ = [("left", hush_surface), ("right", hush_surface), ("mid", hush_surface), ("split", ctypes.c_float),
                ("split2", ctypes.c_float), ("same", ctypes.c_int32), ("exposure", ctypes.c_float),
                ("line", ctypes.c_float), ("target", ctypes.c_void_p), ("target_pitch", ctypes.c_size_t),
                ("width", ctypes.c_int32), ("height", ctypes.c_int32), ("format", ctypes.c_int32)]


_lib = None


def load() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run `python -c 'import __graft_entry__ as g; g.build()'` "
                          "or `python hush_amd/build.py`")
    L = ctypes.CDLL(LIB_PATH)
    P, i32, f32, sz = ctypes.c_void_p, ctypes.c_int32, ctypes.c_float, ctypes.c_size_t
    S = ctypes.POINTER(hush_surface)
    I = ctypes.POINTER(hush_inputs)
    L.hush_hidden_of.argtypes = [sz]; L.hush_hidden_of.restype = i32
    L.hush_create.argtypes = [i32, i32, i32, f32, P, sz, i32, ctypes.POINTER(P)]; L.hush_create.restype = i32
    L.hush_destroy.argtypes = [P]; L.hush_destroy.restype = None
    L.hush_resize.argtypes = [P, i32, i32]; L.hush_resize.restype = i32
    L.hush_reset.argtypes = [P]; L.hush_reset.restype = i32
    L.hush_set_exposure.argtypes = [P, f32]; L.hush_set_exposure.restype = i32
    L.hush_get_exposure.argtypes = [P]; L.hush_get_exposure.restype = f32
    L.hush_get_size.argtypes = [P, ctypes.POINTER(i32), ctypes.POINTER(i32)]; L.hush_get_size.restype = i32
    L.hush_hidden.argtypes = [P]; L.hush_hidden.restype = i32
    L.hush_frame.argtypes = [P]; L.hush_frame.restype = ctypes.c_int64
    L.hush_set_precision.argtypes = [P, i32]; L.hush_set_precision.restype = i32
    L.hush_get_precision.argtypes = [P]; L.hush_get_precision.restype = i32
    L.hush_set_fp8_scales.argtypes = [P, ctypes.POINTER(f32)]; L.hush_set_fp8_scales.restype = i32
    L.hush_get_fp8_scales.argtypes = [P, ctypes.POINTER(f32)]; L.hush_get_fp8_scales.restype = i32
    L.hush_set_ieee_math.argtypes = [P, i32]; L.hush_set_ieee_math.restype = i32
    L.hush_get_ieee_math.argtypes = [P]; L.hush_get_ieee_math.restype = i32
    L.hush_set_compensated.argtypes = [P, i32]; L.hush_set_compensated.restype = i32
    L.hush_get_compensated.argtypes = [P]; L.hush_get_compensated.restype = i32
    L.hush_mlp_forward.argtypes = [P, i32, P, P, ctypes.c_int64, P]; L.hush_mlp_forward.restype = i32
    L.hush_denoise.argtypes = [P, I, P, S]; L.hush_denoise.restype = i32
    L.hush_encode.argtypes = [P, I, P, S]; L.hush_encode.restype = i32
    L.hush_output.argtypes = [P, S]; L.hush_output.restype = i32
    L.hush_last_error.argtypes = [P]; L.hush_last_error.restype = ctypes.c_char_p
    L.hush_set_timing.argtypes = [P, i32]; L.hush_set_timing.restype = i32
    L.hush_pass_times.argtypes = [P, ctypes.POINTER(f32), ctypes.POINTER(ctypes.c_char_p)]; L.hush_pass_times.restype = i32
    L.hush_debug_surface.argtypes = [P, i32, S]; L.hush_debug_surface.restype = i32
    L.hush_debug_set_state.argtypes = [P, ctypes.c_int64, i32]; L.hush_debug_set_state.restype = i32
    L.hush_debug_run_pass.argtypes = [P, i32, I, P]; L.hush_debug_run_pass.restype = i32
    L.hush_set_debug_mask.argtypes = [P, i32]; L.hush_set_debug_mask.restype = i32
    L.hush_debug_wide_gathers.argtypes = [P, i32]; L.hush_debug_wide_gathers.restype = i32
    L.hush_debug_math_selftest.argtypes = [i32, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint32)]
    L.hush_debug_math_selftest.restype = i32
    O = ctypes.POINTER(svgf_options)
    SI = ctypes.POINTER(svgf_inputs)
    L.svgf_default_options.argtypes = [O]; L.svgf_default_options.restype = None
    L.svgf_create.argtypes = [i32, i32, i32, O, ctypes.POINTER(P)]; L.svgf_create.restype = i32
    L.svgf_destroy.argtypes = [P]; L.svgf_destroy.restype = None
    L.svgf_resize.argtypes = [P, i32, i32]; L.svgf_resize.restype = i32
    L.svgf_reset.argtypes = [P]; L.svgf_reset.restype = i32
    L.svgf_get_size.argtypes = [P, ctypes.POINTER(i32), ctypes.POINTER(i32)]; L.svgf_get_size.restype = i32
    L.svgf_frame.argtypes = [P]; L.svgf_frame.restype = ctypes.c_int64
    L.svgf_get_options.argtypes = [P, O]; L.svgf_get_options.restype = i32
    L.svgf_encode.argtypes = [P, SI, P, S]; L.svgf_encode.restype = i32
    L.svgf_output.argtypes = [P, S]; L.svgf_output.restype = i32
    L.svgf_last_error.argtypes = [P]; L.svgf_last_error.restype = ctypes.c_char_p
    L.svgf_set_timing.argtypes = [P, i32]; L.svgf_set_timing.restype = i32
    L.svgf_pass_times.argtypes = [P, ctypes.POINTER(f32), ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(i32)]
    L.svgf_pass_times.restype = i32
    L.svgf_debug_surface.argtypes = [P, i32, S]; L.svgf_debug_surface.restype = i32
    L.svgf_debug_set_state.argtypes = [P, ctypes.c_int64, i32]; L.svgf_debug_set_state.restype = i32
    L.svgf_debug_run_pass.argtypes = [P, i32, SI, P]; L.svgf_debug_run_pass.restype = i32
    L.hush_scene_create.argtypes = [i32, ctypes.POINTER(P)]; L.hush_scene_create.restype = i32
    L.hush_scene_destroy.argtypes = [P]; L.hush_scene_destroy.restype = None
    L.hush_scene_last_error.argtypes = [P]; L.hush_scene_last_error.restype = ctypes.c_char_p
    L.hush_scene_render.argtypes = [P, ctypes.POINTER(hush_scene_frame), I, P, sz, P]; L.hush_scene_render.restype = i32
    PO = ctypes.POINTER(post_options)
    L.post_default_options.argtypes = [PO]; L.post_default_options.restype = None
    L.post_create.argtypes = [i32, PO, ctypes.POINTER(P)]; L.post_create.restype = i32
    L.post_destroy.argtypes = [P]; L.post_destroy.restype = None
    L.post_resize.argtypes = [P, i32, i32]; L.post_resize.restype = i32
    L.post_encode.argtypes = [P, ctypes.POINTER(post_args), P]; L.post_encode.restype = i32
    L.post_last_error.argtypes = [P]; L.post_last_error.restype = ctypes.c_char_p
    L.post_debug_surface.argtypes = [P, i32, S]; L.post_debug_surface.restype = i32
    _lib = L
    return L
