"""ff_render_field_bwd_noalloc (static-scratch variant of ff_render_field_bwd)
against ff_render_field_bwd, through the C ABI, on two calls in a row, the
second with a smaller input.  Decoder gradients come from the same kernels and
the same fixed-order reduction, so they must agree bit for bit; table
gradients are FP32 atomic sums, whose order may differ, so they are compared
at 1e-5 relative (floor 1e-6 of the largest magnitude)."""
import ctypes

import pytest
import torch

import ffields._lib as L
from helpers import assert_close
from test_gpu_render import make_model, make_rays
from ffields.render import _samples_struct

pytestmark = pytest.mark.gpu


def _grad_bufs(m):
    b, c = m._tables()
    tabs = [torch.zeros_like(t, memory_format=torch.preserve_format) for t in b + c]
    dec = [torch.zeros_like(p) for p in m.linear_mat.params_list()]
    g = L.Grads()
    for l in range(len(b)):
        g.basis[l] = tabs[l].data_ptr()
    for k in range(len(c)):
        g.coeff[k] = tabs[len(b) + k].data_ptr()
    g.w0, g.b0, g.w1 = dec[0].data_ptr(), dec[1].data_ptr(), dec[2].data_ptr()
    g.replicas, g.replica_stride = 1, 0
    return tabs, dec, g


def _one_case(m, n_rays, seed):
    rays = make_rays(n_rays, seed=seed).cuda()
    target = torch.rand((n_rays, 3), generator=torch.Generator().manual_seed(seed + 1)).cuda()
    jit = torch.rand(n_rays, generator=torch.Generator().manual_seed(seed + 2)).cuda()
    buf = m.render_buffers(n_rays, m.nSamples)
    s = m.render_samples(rays, m.nSamples, jit, buf)
    st = L.stream_of(rays)
    rgb_map = torch.empty((n_rays, 3), device="cuda")
    rgb_raw = torch.empty((n_rays, 3), device="cuda")
    thres = float(m.cfg.renderer.rayMarch_weight_thres)
    m._composite(s, n_rays, buf, True, rgb_map, rgb_raw, stream=st)
    L.check(L.lib().ff_composite_bwd(ctypes.byref(s), n_rays, L.ptr(buf["sigma"]), L.ptr(buf["rgb"]), 1, thres,
                                     L.ptr(rgb_raw), None, L.ptr(target), 1.0, None, L.ptr(buf["d_sample"]), st),
            "ff_composite_bwd")
    d, mm = m._field_structs()
    h = m._head()
    # with FP32 tables the backward overwrites the stored coding rows with
    # their gradients: give the second call the rows the forward wrote
    rows = None if buf.get("coding") is None else buf["coding"].clone()
    # the existing entry point, with a caller-owned workspace
    tabs_a, dec_a, ga = _grad_bufs(m)
    nb = L.lib().ff_render_field_bwd_workspace_bytes(ctypes.byref(d), ctypes.byref(mm), s.cap)
    assert 0 < nb <= L.lib().ff_render_field_bwd_noalloc_limit_bytes()
    ws = torch.empty((nb + 3) // 4, device="cuda")
    L.check(L.lib().ff_render_field_bwd(ctypes.byref(d), ctypes.byref(mm), ctypes.byref(h), L.ptr(rays), n_rays,
                                        ctypes.byref(s), L.ptr(buf["d_sample"]), ctypes.byref(ga), L.ptr(ws), nb,
                                        st), "ff_render_field_bwd")
    # the allocation-free one
    if rows is not None:
        buf["coding"].copy_(rows)
    tabs_b, dec_b, gb = _grad_bufs(m)
    L.check(L.lib().ff_render_field_bwd_noalloc(ctypes.byref(d), ctypes.byref(mm), ctypes.byref(h), L.ptr(rays),
                                                n_rays, ctypes.byref(s), L.ptr(buf["d_sample"]),
                                                ctypes.byref(gb), st), "ff_render_field_bwd_noalloc")
    torch.cuda.synchronize()
    n = int(buf["ray_off"][-1])
    assert n > 0
    for k, (a, b) in enumerate(zip(dec_a, dec_b)):
        assert torch.equal(a, b), f"decoder gradient {k} differs"
        assert a.abs().max().item() > 0
    for k, (a, b) in enumerate(zip(tabs_a, tabs_b)):
        assert_close(b.contiguous(), a.contiguous(), 1e-5, f"table gradient {k}", atol_frac=1e-6)


@pytest.mark.parametrize("dtype", ["fp32", "bf16"])
def test_noalloc_matches_workspace_entry_point(dtype):
    print(f"noalloc render field backward, tables {dtype}", flush=True)
    m = make_model(seed=3)
    if dtype == "bf16":
        m.table_dtype = "bf16"
        m.refresh_shadows()
        m._build_spec()
    _one_case(m, 512, seed=11)     # first call
    _one_case(m, 96, seed=12)      # second call, smaller input, same static scratch


def test_noalloc_limit_and_argument_checks():
    """The static scratch holds the workspace of the largest persistent grid
    this GPU gets, and bad arguments are refused before anything is queued."""
    m = make_model(seed=4)
    d, mm = m._field_structs()
    h = m._head()
    lim = L.lib().ff_render_field_bwd_noalloc_limit_bytes()
    assert lim == 16 << 20
    need = L.lib().ff_render_field_bwd_workspace_bytes(ctypes.byref(d), ctypes.byref(mm), 1 << 40)
    assert 0 < need <= lim
    g = L.Grads()
    assert L.lib().ff_render_field_bwd_noalloc(None, ctypes.byref(mm), ctypes.byref(h), None, 64, None, None,
                                               ctypes.byref(g), None) == L.FF_ERR_ARG
    buf = m.render_buffers(64, m.nSamples)
    s = _samples_struct(buf)
    st = L.lib().ff_render_field_bwd_noalloc(ctypes.byref(d), ctypes.byref(mm), ctypes.byref(h), None, 64,
                                             ctypes.byref(s), None, ctypes.byref(g), None)
    assert st == L.FF_ERR_ARG     # no rays, no sample gradients, no gradient buffers
