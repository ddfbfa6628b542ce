"""ctypes binding of libffields.so (the C ABI declared in include/ffields.h).

The library is loaded from this package directory.  There is no fallback: if
the library is missing or fails to load, every call raises.  Structures below
mirror the C structs field for field.
"""

import ctypes
import os

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
# FFIELDS_LIB selects another build of the same library (the CPU AddressSanitizer
# build of the argument checks, tests/asan_build.py); unset, the in-tree one
LIB_PATH = os.environ.get("FFIELDS_LIB") or os.path.join(_HERE, "libffields.so")

FF_MAX_LEVELS = 8

FF_OK, FF_ERR_ARG, FF_ERR_UNSUPPORTED, FF_ERR_WORKSPACE, FF_ERR_LAUNCH, FF_ERR_SCRATCH_LIMIT = range(6)
MAP = {"sawtooth": 0, "triangle": 1, "sinc": 2, "x": 3}
INTERP = {"bilinear": 0, "nearest": 1}
BASIS_NONE, BASIS_GRID = 0, 1
COEFF_NONE, COEFF_GRID, COEFF_VM = 0, 1, 2
F32, BF16 = 0, 1
LOSS = {"mse": 0, "l1": 1}

c_i32 = ctypes.c_int32
c_f32 = ctypes.c_float
c_vp = ctypes.c_void_p


class FieldDesc(ctypes.Structure):
    _fields_ = [
        ("in_dim", c_i32), ("x_dim", c_i32), ("n_levels", c_i32), ("n_channels", c_i32),
        ("mapping", c_i32), ("basis_kind", c_i32), ("basis_interp", c_i32),
        ("coeff_kind", c_i32), ("coeff_interp", c_i32), ("coeff_ndim", c_i32),
        ("table_dtype", c_i32),
        ("basis_dims", c_i32 * FF_MAX_LEVELS), ("basis_reso", c_i32 * FF_MAX_LEVELS),
        ("coeff_reso", c_i32 * 3),
        ("map_origin", c_f32 * 3), ("map_scale", c_f32 * FF_MAX_LEVELS),
        ("coeff_origin", c_f32 * 3), ("coeff_inv", c_f32 * 3),
        ("basis", c_vp * FF_MAX_LEVELS), ("coeff", c_vp * 6),
    ]


class MlpDesc(ctypes.Structure):
    _fields_ = [("n_layers", c_i32), ("in_dim", c_i32), ("hidden", c_i32), ("out_dim", c_i32),
                ("w0", c_vp), ("b0", c_vp), ("w1", c_vp)]


class Grads(ctypes.Structure):
    _fields_ = [("basis", c_vp * FF_MAX_LEVELS), ("coeff", c_vp * 6),
                ("w0", c_vp), ("b0", c_vp), ("w1", c_vp), ("replicas", c_i32),
                ("replica_stride", ctypes.c_int64)]


class AdamTensor(ctypes.Structure):
    _fields_ = [("param", c_vp), ("grad", c_vp), ("exp_avg", c_vp), ("exp_avg_sq", c_vp),
                ("shadow_bf16", c_vp), ("n", ctypes.c_int64), ("lr", c_f32),
                ("grad_replicas", c_i32), ("grad_stride", ctypes.c_int64)]


class RaySampler(ctypes.Structure):
    _fields_ = [("aabb_min", c_f32 * 3), ("aabb_max", c_f32 * 3), ("step_size", c_f32),
                ("n_samples", c_i32), ("distance_scale", c_f32), ("occ", c_vp),
                ("occ_reso", c_i32 * 3), ("occ_origin", c_f32 * 3), ("occ_inv", c_f32 * 3)]


class Samples(ctypes.Structure):
    _fields_ = [("xyz", c_vp), ("z", c_vp), ("dist", c_vp), ("ray", c_vp), ("idx", c_vp),
                ("ray_off", c_vp), ("cap", ctypes.c_int64), ("coding", c_vp)]


COLOR_SH, COLOR_FEATURES = 0, 1


class RenderHead(ctypes.Structure):
    _fields_ = [("density_act", c_i32), ("density_shift", c_f32), ("color", c_i32), ("feat_stride", c_i32),
                ("feat", c_vp), ("d_feat", c_vp), ("app_pos", c_vp)]


class ColorMlpDesc(ctypes.Structure):
    _fields_ = [("in_feat", c_i32), ("view_pe", c_i32), ("fea_pe", c_i32), ("hidden", c_i32), ("n_layers", c_i32),
                ("w0", c_vp), ("b0", c_vp), ("w1", c_vp), ("b1", c_vp), ("w2", c_vp)]


class ColorRows(ctypes.Structure):
    _fields_ = [("feat", c_vp), ("feat_stride", ctypes.c_int64), ("dirs", c_vp), ("dir_stride", ctypes.c_int64),
                ("dir_index", c_vp), ("rows", c_vp), ("count_dev", c_vp), ("n_max", ctypes.c_int64)]


class ColorMlpGrads(ctypes.Structure):
    _fields_ = [("w0", c_vp), ("b0", c_vp), ("w1", c_vp), ("b1", c_vp), ("w2", c_vp)]


class ProposalDesc(ctypes.Structure):
    _fields_ = [("reso", c_i32 * 3), ("table_dtype", c_i32), ("grid", c_vp), ("w0", c_vp), ("b0", c_vp),
                ("w1", c_vp), ("aabb_min", c_f32 * 3), ("aabb_max", c_f32 * 3), ("density_shift", c_f32),
                ("distance_scale", c_f32), ("near", c_f32), ("n_fine", c_i32)]


class ProposalGrads(ctypes.Structure):
    _fields_ = [("grid", c_vp), ("w0", c_vp), ("b0", c_vp), ("w1", c_vp), ("replicas", c_i32),
                ("replica_stride", ctypes.c_int64)]


class Cameras(ctypes.Structure):
    _fields_ = [("c2w", c_vp), ("n_cams", c_i32), ("H", c_i32), ("W", c_i32), ("fx", c_f32), ("fy", c_f32),
                ("cx", c_f32), ("cy", c_f32), ("blender", c_i32), ("normalize", c_i32)]


_lib = None


def _declare(lib):
    P = ctypes.POINTER
    I64 = ctypes.c_int64
    sz = ctypes.c_size_t
    sig = {
        "ff_status_str": (ctypes.c_char_p, [ctypes.c_int]),
        "ff_version": (ctypes.c_int, []),
        "ff_get_coding": (ctypes.c_int, [P(FieldDesc), c_vp, I64, c_vp, c_vp, c_vp, c_vp]),
        "ff_get_coding_bwd": (ctypes.c_int, [P(FieldDesc), c_vp, I64, c_vp, c_vp, c_vp, P(Grads), c_vp]),
        "ff_linear_mat_fwd": (ctypes.c_int, [P(MlpDesc), c_vp, I64, c_vp, c_vp]),
        "ff_linear_mat_bwd_workspace_bytes": (sz, [P(MlpDesc), I64]),
        "ff_linear_mat_bwd": (ctypes.c_int, [P(MlpDesc), c_vp, I64, c_vp, c_vp, P(Grads), c_vp, sz, c_vp]),
        "ff_field_fwd": (ctypes.c_int, [P(FieldDesc), P(MlpDesc), c_vp, I64, c_vp, c_vp]),
        "ff_field_fused_supported": (ctypes.c_int, [P(FieldDesc), P(MlpDesc)]),
        "ff_field_bwd_workspace_bytes": (sz, [P(FieldDesc), P(MlpDesc), I64]),
        "ff_field_bwd": (ctypes.c_int, [P(FieldDesc), P(MlpDesc), c_vp, I64, c_vp, P(Grads), c_vp, sz,
                                        c_vp]),
        "ff_field_train_regression": (ctypes.c_int, [P(FieldDesc), P(MlpDesc), c_vp, c_vp, I64,
                                                     ctypes.c_int, c_f32, P(Grads), c_vp, c_vp, sz,
                                                     c_vp]),
        "ff_adam_step": (ctypes.c_int, [c_vp, c_vp, c_vp, c_vp, c_vp, I64, c_f32, c_f32, c_f32, c_f32,
                                        I64, ctypes.c_int, c_vp]),
        "ff_adam_multi": (ctypes.c_int, [P(AdamTensor), ctypes.c_int, c_f32, c_f32, c_f32, I64,
                                         ctypes.c_int, c_vp]),
        "ff_cast_bf16": (ctypes.c_int, [c_vp, c_vp, I64, c_vp]),
        "ff_reduce_replicas": (ctypes.c_int, [c_vp, c_i32, I64, I64, c_vp]),
        "ff_sample_rays": (ctypes.c_int, [P(RaySampler), c_vp, c_vp, I64, P(Samples), c_vp]),
        "ff_render_field_fwd": (ctypes.c_int, [P(FieldDesc), P(MlpDesc), P(RenderHead), c_vp, I64,
                                               P(Samples), c_vp, c_vp, c_vp]),
        "ff_render_field_bwd_workspace_bytes": (sz, [P(FieldDesc), P(MlpDesc), I64]),
        "ff_render_field_bwd": (ctypes.c_int, [P(FieldDesc), P(MlpDesc), P(RenderHead), c_vp, I64,
                                               P(Samples), c_vp, P(Grads), c_vp, sz, c_vp]),
        "ff_render_field_bwd_noalloc_limit_bytes": (sz, []),
        "ff_render_field_bwd_noalloc": (ctypes.c_int, [P(FieldDesc), P(MlpDesc), P(RenderHead), c_vp, I64,
                                                       P(Samples), c_vp, P(Grads), c_vp]),
        "ff_render_alpha": (ctypes.c_int, [P(FieldDesc), P(MlpDesc), P(RenderHead), P(RaySampler), c_vp,
                                           I64, c_f32, c_vp, c_vp]),
        "ff_select_app": (ctypes.c_int, [P(Samples), I64, c_vp, c_f32, c_vp, c_vp, c_vp, c_vp]),
        "ff_color_mlp_workspace_bytes": (sz, [P(ColorMlpDesc), I64, ctypes.c_int]),
        "ff_color_mlp_fwd": (ctypes.c_int, [P(ColorMlpDesc), P(ColorRows), c_vp, I64, c_vp, sz, c_vp]),
        "ff_color_mlp_bwd": (ctypes.c_int, [P(ColorMlpDesc), P(ColorRows), c_vp, I64, c_vp, I64, P(ColorMlpGrads),
                                            c_vp, sz, c_vp]),
        "ff_dropout": (ctypes.c_int, [c_vp, I64, c_f32, c_vp, ctypes.c_uint64, ctypes.c_uint64, c_vp, c_vp, c_vp]),
        "ff_occupancy_pool": (ctypes.c_int, [c_vp, c_i32, c_i32, c_i32, c_f32, c_f32, c_vp, c_vp]),
        "ff_proposal_sample": (ctypes.c_int, [P(ProposalDesc), c_vp, c_vp, c_vp, I64, P(Samples), c_vp]),
        "ff_proposal_bwd_workspace_bytes": (ctypes.c_size_t, [I64]),
        "ff_proposal_bwd": (ctypes.c_int, [P(ProposalDesc), c_vp, c_vp, I64, P(Samples), c_vp, c_f32,
                                           P(ProposalGrads), c_vp, c_vp, ctypes.c_size_t, c_vp]),
        "ff_sh_rgb": (ctypes.c_int, [c_vp, I64, c_vp, I64, I64, c_vp, c_vp]),
        "ff_sh_rgb_bwd": (ctypes.c_int, [c_vp, I64, c_vp, I64, I64, c_vp, c_vp, c_vp]),
        "ff_atomic_rate_probe": (ctypes.c_int, [c_vp, I64, I64, c_i32, c_vp]),
        "ff_camera_rays": (ctypes.c_int, [P(Cameras), c_vp, c_vp, I64, c_vp, c_vp]),
        "ff_gather_rays": (ctypes.c_int, [c_vp, c_vp, I64, c_vp, I64, c_vp, c_vp, c_vp]),
        "ff_sq_err_sum": (ctypes.c_int, [c_vp, c_vp, I64, c_vp, c_vp]),
        "ff_remove_small_components": (ctypes.c_int, [c_vp, c_i32, c_i32, c_i32, ctypes.c_double, c_vp, c_vp,
                                                       c_vp, c_vp]),
        "ff_mt_count": (ctypes.c_int, [c_vp, I64, I64, I64, ctypes.c_double, c_i32, c_vp, c_vp]),
        "ff_mt_emit": (ctypes.c_int, [c_vp, I64, I64, I64, ctypes.c_double, c_i32, c_vp, I64, c_vp, c_vp, c_vp]),
        "ff_mt_verts": (ctypes.c_int, [c_vp, I64, I64, I64, ctypes.c_double, c_vp, I64, c_vp, c_vp]),
        "ff_composite_fwd": (ctypes.c_int, [P(Samples), I64, c_vp, c_vp, ctypes.c_int, c_f32, c_vp,
                                            c_vp, c_vp, c_vp, c_vp, c_vp]),
        "ff_composite_bwd": (ctypes.c_int, [P(Samples), I64, c_vp, c_vp, ctypes.c_int, c_f32, c_vp,
                                            c_vp, c_vp, c_f32, c_vp, c_vp, c_vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args


def exported_symbols():
    """Names the C ABI header declares (checked against the library in tests)."""
    import re
    hdr = os.path.join(os.path.dirname(os.path.dirname(_HERE)), "include")
    names = []
    for f in sorted(os.listdir(hdr)):
        if f.endswith(".h"):
            txt = open(os.path.join(hdr, f)).read()
            names += re.findall(r"^\s*(?:const\s+)?[\w]+\*?\s+\**(ff_\w+)\s*\(", txt, re.M)
    return names


def lib():
    """The loaded library; raises if it cannot be loaded (no fallback)."""
    global _lib, _load_error
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libffields.so not built ({LIB_PATH}); run __graft_entry__.build()")
        l = ctypes.CDLL(LIB_PATH)
        _declare(l)
        _lib = l
    return _lib


def check(status, what=""):
    if status != FF_OK:
        msg = lib().ff_status_str(status).decode()
        raise RuntimeError(f"ffields: {what} failed: {msg} (status {status})")


_raw_stream = getattr(torch._C, "_cuda_getCurrentRawStream", None)


def stream_of(t=None):
    """The current stream of t's device (of the current device if t is None)
    as a hipStream_t.  torch's raw-stream query returns the same handle as
    torch.cuda.current_stream(dev).cuda_stream without building a Stream
    object (a few microseconds per call on the training step's hot path)."""
    if t is None:
        idx = torch.cuda.current_device()
    elif t.device.type == "cuda":
        idx = t.device.index if t.device.index is not None else torch.cuda.current_device()
    else:
        return ctypes.c_void_p(torch.cuda.current_stream(t.device).cuda_stream)  # raises: not a device tensor
    if _raw_stream is not None:
        return ctypes.c_void_p(_raw_stream(idx))
    return ctypes.c_void_p(torch.cuda.current_stream(idx).cuda_stream)


def ptr(t):
    return ctypes.c_void_p(0 if t is None else t.data_ptr())
