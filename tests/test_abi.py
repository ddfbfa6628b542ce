"""CPU tests of the C ABI: the library loads, exports every symbol declared in
include/*.h, and rejects bad arguments with a status before touching a GPU.

The argument-validation tests hand the entry points made-up device addresses
(0x1000, ...).  They run only where no GPU is visible: if a validation check
regressed, the call would otherwise launch kernels on those addresses."""
import ctypes

import pytest

import ffields._lib as L


def _gpu_visible():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


fake_pointers = pytest.mark.skipif(_gpu_visible(), reason="passes made-up device pointers; CPU-only check")


def test_library_exports_every_declared_symbol():
    lib = L.lib()
    names = L.exported_symbols()
    assert len(names) >= 10
    for n in names:
        assert hasattr(lib, n), n
    assert lib.ff_version() >= 1


def test_status_strings():
    lib = L.lib()
    assert lib.ff_status_str(0) == b"ok"
    assert b"invalid" in lib.ff_status_str(1)
    assert b"unknown" in lib.ff_status_str(99)


def _valid_desc():
    d = L.FieldDesc()
    d.in_dim = d.x_dim = 3
    d.n_levels = 1
    d.n_channels = 4
    d.mapping = 0
    d.basis_kind = L.BASIS_GRID
    d.coeff_kind = L.COEFF_GRID
    d.coeff_ndim = 3
    d.basis_dims[0] = 4
    d.basis_reso[0] = 8
    d.coeff_reso[0] = d.coeff_reso[1] = d.coeff_reso[2] = 4
    d.basis[0] = 0x1000
    d.coeff[0] = 0x2000
    return d


@fake_pointers
@pytest.mark.parametrize("mutate", [
    lambda d: setattr(d, "in_dim", 4),
    lambda d: setattr(d, "n_channels", 5),
    lambda d: setattr(d, "x_dim", 2),
    lambda d: setattr(d, "mapping", 9),
    lambda d: setattr(d, "basis_interp", 7),
    lambda d: setattr(d, "table_dtype", 3),
    lambda d: d.basis.__setitem__(0, 0),
    lambda d: d.coeff.__setitem__(0, 0),
    lambda d: d.basis_reso.__setitem__(0, 1),
    lambda d: setattr(d, "coeff_kind", L.COEFF_VM),      # VM needs six tables
])
def test_bad_field_desc_rejected(mutate):
    d = _valid_desc()
    mutate(d)
    st = L.lib().ff_get_coding(ctypes.byref(d), ctypes.c_void_p(0x3000), 10, None, None, None, None)
    assert st == L.FF_ERR_ARG


@fake_pointers
def test_null_arguments_rejected():
    lib = L.lib()
    d = _valid_desc()
    assert lib.ff_get_coding(None, ctypes.c_void_p(0x3000), 10, None, None, None, None) == L.FF_ERR_ARG
    assert lib.ff_get_coding(ctypes.byref(d), None, 10, None, None, None, None) == L.FF_ERR_ARG
    assert lib.ff_get_coding(ctypes.byref(d), None, -1, None, None, None, None) == L.FF_ERR_ARG
    m = L.MlpDesc()
    assert lib.ff_linear_mat_fwd(ctypes.byref(m), None, 0, None, None) == L.FF_ERR_ARG
    m.n_layers, m.in_dim, m.hidden, m.out_dim, m.w0 = 2, 4, 8, 40, 0x10
    assert lib.ff_linear_mat_fwd(ctypes.byref(m), None, 0, None, None) == L.FF_ERR_ARG  # out_dim > 32
    g = L.Grads()
    assert lib.ff_get_coding_bwd(ctypes.byref(d), ctypes.c_void_p(0x3000), 10, None, None, None,
                                 ctypes.byref(g), None) == L.FF_ERR_ARG             # null grad tables
    assert lib.ff_adam_multi(None, 1, 0.9, 0.99, 1e-8, 1, 0, None) == L.FF_ERR_ARG
    t = L.AdamTensor(0, 0, 0, 0, 0, 5, 0.1)
    assert lib.ff_adam_multi(ctypes.byref(t), 1, 0.9, 0.99, 1e-8, 1, 0, None) == L.FF_ERR_ARG
    assert lib.ff_adam_multi(ctypes.byref(t), 1, 0.9, 0.99, 1e-8, 0, 0, None) == L.FF_ERR_ARG


@fake_pointers
def test_zero_size_is_a_noop():
    d = _valid_desc()
    assert L.lib().ff_get_coding(ctypes.byref(d), None, 0, None, None, None, None) == L.FF_OK


@fake_pointers
def test_fused_support_query():
    d = _valid_desc()
    m = L.MlpDesc()
    m.n_layers, m.in_dim, m.hidden, m.out_dim, m.w0, m.b0, m.w1 = 2, 4, 64, 3, 0x10, 0x20, 0x30
    assert L.lib().ff_field_fused_supported(ctypes.byref(d), ctypes.byref(m)) == 0   # 4 channels: generic


@fake_pointers
def test_proposal_bwd_workspace_checked():
    lib = L.lib()
    nb = lib.ff_proposal_bwd_workspace_bytes(4096)
    assert nb >= (4096 // 16 + 32) * 320 * 4          # at least one partial row per wave plus scratch
    assert lib.ff_proposal_bwd_workspace_bytes(0) == 0
    p = L.ProposalDesc()
    p.reso[0] = p.reso[1] = p.reso[2] = 8
    p.table_dtype = L.F32
    p.grid = p.w0 = p.b0 = p.w1 = 0x1000
    for k in range(3):
        p.aabb_min[k], p.aabb_max[k] = -1.0, 1.0
    p.n_fine = 48
    s = L.Samples()
    s.xyz = s.z = s.dist = s.ray = s.idx = s.ray_off = 0x2000
    s.cap = 48 * 16
    g = L.ProposalGrads()
    g.grid = g.w0 = g.b0 = g.w1 = 0x3000
    g.replicas = 1
    st = lib.ff_proposal_bwd(ctypes.byref(p), ctypes.c_void_p(0x4000), None, 16, ctypes.byref(s),
                             ctypes.c_void_p(0x5000), 1.0, ctypes.byref(g), None, ctypes.c_void_p(0x6000), 16, None)
    assert st == L.FF_ERR_WORKSPACE
    p.n_fine = 65
    st = lib.ff_proposal_bwd(ctypes.byref(p), ctypes.c_void_p(0x4000), None, 16, ctypes.byref(s),
                             ctypes.c_void_p(0x5000), 1.0, ctypes.byref(g), None, ctypes.c_void_p(0x6000), 1 << 30,
                             None)
    assert st == L.FF_ERR_ARG


@fake_pointers
def test_sampler_limits_rejected():
    """The sampler holds at most 2,048 samples per ray (32 rounds of 64) and
    the proposal at most 64 fine samples: larger requests fail with
    FF_ERR_ARG before any work is queued."""
    lib = L.lib()
    s = L.RaySampler()
    for k in range(3):
        s.aabb_min[k], s.aabb_max[k] = -1.0, 1.0
    s.step_size = 0.01
    s.distance_scale = 25.0
    out = L.Samples()
    out.xyz = out.z = out.dist = out.ray = out.idx = out.ray_off = 0x1000
    out.cap = 16 * 4096
    for ns, want in ((2049, L.FF_ERR_ARG), (0, L.FF_ERR_ARG)):
        s.n_samples = ns
        st = lib.ff_sample_rays(ctypes.byref(s), ctypes.c_void_p(0x2000), None, 16, ctypes.byref(out), None)
        assert st == want, (ns, st)
    s.n_samples = 2048
    out.cap = 16 * 2048 - 1                       # capacity one short of n_rays * n_samples
    st = lib.ff_sample_rays(ctypes.byref(s), ctypes.c_void_p(0x2000), None, 16, ctypes.byref(out), None)
    assert st == L.FF_ERR_ARG


@fake_pointers
def test_ray_entry_points_validate_arguments():
    """ff_camera_rays / ff_gather_rays: empty requests return ok with NULL
    pointers (an empty tensor has none); inconsistent pointers and sizes are
    rejected before any launch."""
    lib = L.lib()
    P = ctypes.c_void_p
    c = L.Cameras(P(0x1000), 2, 4, 5, 3.0, 3.0, 2.5, 2.0, 0, 1)
    assert lib.ff_camera_rays(ctypes.byref(c), None, None, 0, None, None) == L.FF_OK
    assert lib.ff_camera_rays(ctypes.byref(c), P(0x2000), None, 8, P(0x3000), None) == L.FF_ERR_ARG  # pix missing
    assert lib.ff_camera_rays(ctypes.byref(c), None, None, 39, P(0x3000), None) == L.FF_ERR_ARG      # != 2*4*5
    assert lib.ff_camera_rays(ctypes.byref(c), P(0x2000), P(0x2800), 8, None, None) == L.FF_ERR_ARG  # no output
    assert lib.ff_camera_rays(None, None, None, 0, None, None) == L.FF_ERR_ARG
    for bad in (lambda x: setattr(x, "n_cams", 0), lambda x: setattr(x, "W", 0), lambda x: setattr(x, "fy", 0.0),
                lambda x: setattr(x, "c2w", None)):
        c2 = L.Cameras(P(0x1000), 2, 4, 5, 3.0, 3.0, 2.5, 2.0, 0, 1)
        bad(c2)
        assert lib.ff_camera_rays(ctypes.byref(c2), None, None, 40, P(0x3000), None) == L.FF_ERR_ARG
    assert lib.ff_gather_rays(P(0x1000), P(0x1800), 10, None, 0, None, None, None) == L.FF_OK
    assert lib.ff_gather_rays(P(0x1000), P(0x1800), 10, P(0x2000), 4, P(0x3000), None, None) == L.FF_ERR_ARG
    assert lib.ff_gather_rays(P(0x1000), None, 10, None, 4, P(0x3000), None, None) == L.FF_ERR_ARG      # no idx
    assert lib.ff_gather_rays(P(0x1000), None, -1, P(0x2000), 4, P(0x3000), None, None) == L.FF_ERR_ARG
    assert lib.ff_gather_rays(P(0x1000), None, 10, P(0x2000), -4, P(0x3000), None, None) == L.FF_ERR_ARG


def test_mesh_entry_points_validate_arguments():
    """ff_mt_count / ff_mt_emit / ff_mt_verts: volumes under 2 points on an
    axis, over 2^20, or whose edge keys would overflow int64, and missing
    buffers are rejected before any launch; no distinct keys is a no-op."""
    lib = L.lib()
    P = ctypes.c_void_p
    lev = ctypes.c_double(0.5)
    assert lib.ff_mt_count(None, 4, 4, 4, lev, 1, P(0x2000), None) == L.FF_ERR_ARG
    assert lib.ff_mt_count(P(0x1000), 1, 4, 4, lev, 1, P(0x2000), None) == L.FF_ERR_ARG
    assert lib.ff_mt_count(P(0x1000), 4, 4, (1 << 20) + 1, lev, 1, P(0x2000), None) == L.FF_ERR_ARG
    assert lib.ff_mt_count(P(0x1000), 1 << 20, 1 << 20, 4, lev, 1, P(0x2000), None) == L.FF_ERR_ARG  # keys overflow
    assert lib.ff_mt_count(P(0x1000), 4, 4, 4, lev, 1, None, None) == L.FF_ERR_ARG
    assert lib.ff_mt_emit(P(0x1000), 4, 4, 4, lev, 1, None, 0, None, P(0x3000), None) == L.FF_ERR_ARG
    assert lib.ff_mt_emit(P(0x1000), 4, 4, 4, lev, 1, None, 0, P(0x2000), None, None) == L.FF_ERR_ARG
    assert lib.ff_mt_emit(P(0x1000), 4, 4, 4, lev, 1, P(0x4000), 28, P(0x2000), P(0x3000), None) == L.FF_ERR_ARG
    assert lib.ff_mt_emit(P(0x1000), 4, 4, 4, lev, 1, P(0x4000), -1, P(0x2000), P(0x3000), None) == L.FF_ERR_ARG
    assert lib.ff_mt_emit(P(0x1000), 4, 4, 4, lev, 1, P(0x4000), 0, None, None, None) == L.FF_OK
    assert lib.ff_mt_verts(P(0x1000), 4, 4, 4, lev, None, 0, None, None) == L.FF_OK
    assert lib.ff_mt_verts(P(0x1000), 4, 4, 4, lev, P(0x2000), -1, P(0x3000), None) == L.FF_ERR_ARG
    assert lib.ff_mt_verts(P(0x1000), 4, 4, 4, lev, None, 5, P(0x3000), None) == L.FF_ERR_ARG
    assert lib.ff_mt_verts(P(0x1000), 0, 4, 4, lev, P(0x2000), 5, P(0x3000), None) == L.FF_ERR_ARG


def test_component_filter_validates_arguments():
    """ff_remove_small_components: an empty volume is a no-op with NULL
    buffers; negative sizes, more than 2^31 - 1 voxels, missing buffers and a
    NaN min_size are rejected before any launch."""
    lib = L.lib()
    P = ctypes.c_void_p
    bufs = (P(0x2000), P(0x3000), P(0x4000))
    assert lib.ff_remove_small_components(None, 0, 5, 5, 3.0, None, None, None, None) == L.FF_OK
    assert lib.ff_remove_small_components(P(0x1000), -1, 5, 5, 3.0, *bufs, None) == L.FF_ERR_ARG
    assert lib.ff_remove_small_components(P(0x1000), 2048, 1024, 1024, 3.0, *bufs, None) == L.FF_ERR_ARG
    assert lib.ff_remove_small_components(None, 4, 5, 5, 3.0, *bufs, None) == L.FF_ERR_ARG
    assert lib.ff_remove_small_components(P(0x1000), 4, 5, 5, 3.0, P(0x2000), None, P(0x4000), None) == L.FF_ERR_ARG
    assert lib.ff_remove_small_components(P(0x1000), 4, 5, 5, float("nan"), *bufs, None) == L.FF_ERR_ARG


def test_sq_err_sum_validates_arguments():
    lib = L.lib()
    P = ctypes.c_void_p
    assert lib.ff_sq_err_sum(None, None, 0, P(0x3000), None) == L.FF_OK
    assert lib.ff_sq_err_sum(P(0x1000), P(0x2000), 4, None, None) == L.FF_ERR_ARG
    assert lib.ff_sq_err_sum(P(0x1000), None, 4, P(0x3000), None) == L.FF_ERR_ARG
    assert lib.ff_sq_err_sum(P(0x1000), P(0x2000), -1, P(0x3000), None) == L.FF_ERR_ARG
