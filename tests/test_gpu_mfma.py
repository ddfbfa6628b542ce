"""BF16 matrix-core decoder of the NeRF field kernels (BF16 tables) against a
CPU emulation of its arithmetic: features, decoder weights (b0 as an extra
input column), hidden activations, output gradients and hidden gradients
rounded to BF16 (round to nearest even), products accumulated in FP32.

Bars: outputs and gradients within 2e-3 of the largest magnitude plus 2e-3
relative.  The emulation accumulates in a different order than the matrix
cores, so a value sitting on a BF16 rounding boundary can round the other way
on one side (one BF16 step, 2^-8 relative, of one operand); a layout or
indexing mistake gives errors of the size of the values themselves.  The
median relative error is also required to stay below 1e-4, which only holds
if almost every product is formed from the same BF16 operands."""
import ctypes
import sys

import pytest
import torch
import torch.nn.functional as F

from helpers import assert_close, oracle_of
from ffields import FactorFields
from ffields.render import _samples_struct
import ffields._lib as L
from oracle.render import sh_rgb

from test_gpu_render import make_rays, nerf_small_cfg

pytestmark = pytest.mark.gpu


def bf(t):
    return t.to(torch.bfloat16).float()


def make_bf16_model(seed=0):
    from helpers import randomize
    m = FactorFields(nerf_small_cfg(), "cuda", table_dtype="bf16")
    randomize(m, seed, scale=0.5)
    with torch.no_grad():
        m.linear_mat.backbone[0].weight.mul_(20.0)
    m.refresh_shadows()
    of = oracle_of(m, bf16_tables=True)
    g = torch.Generator().manual_seed(seed + 7)
    a = m.aabb.cpu()
    x = a[0] + (a[1] - a[0]) * torch.rand((4096, 3), generator=g)
    h = F.relu(F.linear(of.get_coding(x)[0], of.mlp[0][0], of.mlp[0][1]))
    o0 = F.linear(h, of.mlp[1][0])[:, 0]
    m.cfg.renderer.density_shift = float(-torch.quantile(o0, 0.6))
    return m


def emulate_forward(m, feats):
    """Decoder on BF16 operands: returns (o [N, 28], pre [N, 64], hb [N, 64], xb [N, 19], W0a_b, W1_b)."""
    lin = m.linear_mat.backbone
    w0 = lin[0].weight.detach().cpu()
    b0 = lin[0].bias.detach().cpu()
    w1 = lin[1].weight.detach().cpu()
    W0a = bf(torch.cat([w0, b0[:, None]], 1))                 # [64, 19]
    W1 = bf(w1)                                               # [28, 64]
    xb = bf(torch.cat([feats, torch.ones((feats.shape[0], 1))], 1))
    pre = (xb.double() @ W0a.double().T).float()
    hb = bf(F.relu(pre))
    o = (hb.double() @ W1.double().T).float()
    return o, pre, hb, xb, W0a, W1


def head(m, o, dirs):
    sig = F.softplus(o[:, 0] + m.cfg.renderer.density_shift)
    rgb = sh_rgb(dirs, o[:, 1:28])
    return sig, rgb


def sample_batch(m, B=300, seed=5):
    rays = make_rays(B, seed=seed)
    jit = torch.rand(B, generator=torch.Generator().manual_seed(seed + 1))
    buf = m.render_buffers(B, m.nSamples)
    s = m.render_samples(rays.cuda(), m.nSamples, jit.cuda(), buf)
    torch.cuda.synchronize()
    return rays, buf, s


def test_mfma_decoder_forward():
    print("[case] mfma decoder forward", file=sys.stderr, flush=True)
    m = make_bf16_model()
    rays, buf, _ = sample_batch(m)
    M = int(buf["ray_off"][-1])
    assert M > 1000
    of = oracle_of(m, bf16_tables=True)
    x = buf["xyz"][:M].cpu()
    feats = of.get_coding(x)[0]
    o, *_ = emulate_forward(m, feats)
    dirs = rays[buf["ray"][:M].cpu().long(), 3:6]
    sig, rgb = head(m, o, dirs)
    got_s, got_c = buf["sigma"][:M].cpu(), buf["rgb"][:M].cpu()
    assert_close(got_s, sig, 2e-3, "sigma", atol_frac=2e-3)
    assert_close(got_c, rgb, 2e-3, "rgb", atol_frac=2e-3)
    rel = ((got_c - rgb).abs() / rgb.abs().clamp(min=1e-6)).flatten()
    assert rel.median().item() < 1e-4, rel.median().item()


@pytest.mark.parametrize("stored", [True, False])
def test_mfma_decoder_backward(stored):
    """stored: the backward reads the coding rows the forward wrote into the
    render buffers; otherwise (coding NULL) it gathers the tables again."""
    print(f"[case] mfma decoder backward stored={stored}", file=sys.stderr, flush=True)
    m = make_bf16_model(seed=3)
    rays, buf, s = sample_batch(m, B=260, seed=21)
    assert s.coding
    if not stored:
        s.coding = None
    M = int(buf["ray_off"][-1])
    g = torch.Generator().manual_seed(22)
    d_sample = torch.randn((s.cap, 4), generator=g) * 0.1
    m.zero_grad(set_to_none=True)
    m.grad_replicas = 1
    d, mm = m._field_structs()
    h = m._head()
    gr = m._train_grads()
    nb = L.lib().ff_render_field_bwd_workspace_bytes(ctypes.byref(d), ctypes.byref(mm), s.cap)
    ws = torch.empty((nb + 3) // 4, device="cuda")
    rays_d, ds_d = rays.cuda(), d_sample.cuda()
    L.check(L.lib().ff_render_field_bwd(ctypes.byref(d), ctypes.byref(mm), ctypes.byref(h), L.ptr(rays_d),
                                        rays.shape[0], ctypes.byref(s), L.ptr(ds_d), ctypes.byref(gr),
                                        L.ptr(ws), nb, L.stream_of(ws)), "ff_render_field_bwd")
    torch.cuda.synchronize()
    # emulation
    of = oracle_of(m, bf16_tables=True).requires_grad_(True)
    x = buf["xyz"][:M].cpu()
    feats = of.get_coding(x)[0]
    o, pre, hb, xb, W0a, W1 = emulate_forward(m, feats.detach())
    o = o.requires_grad_(True)
    dirs = rays[buf["ray"][:M].cpu().long(), 3:6]
    sig, rgb = head(m, o, dirs)
    ds = d_sample[:M]
    dout = torch.autograd.grad((sig * ds[:, 0]).sum() + (rgb * ds[:, 1:4]).sum(), o)[0]
    db = bf(dout)
    dh = (db.double() @ W1.double()).float() * (pre > 0).float()
    dhb = bf(dh)
    df = (dhb.double() @ W0a.double()).float()[:, :18]
    dW1 = (db.double().T @ hb.double()).float()
    dW0a = (dhb.double().T @ xb.double()).float()
    feats.backward(df)
    lin = m.linear_mat.backbone
    assert_close(lin[1].weight.grad, dW1, 2e-3, "dW1", atol_frac=2e-3)
    assert_close(lin[0].weight.grad, dW0a[:, :18], 2e-3, "dW0", atol_frac=2e-3)
    assert_close(lin[0].bias.grad, dW0a[:, 18], 2e-3, "db0", atol_frac=2e-3)
    for k, (t, r) in enumerate(zip(list(m.basises) + list(m.coeffs), of.basises + of.coeffs)):
        assert_close(t.grad.contiguous(), r.grad, 2e-3, f"table {k}", atol_frac=2e-3)


def test_mfma_training_step_reduces_loss():
    """A few BF16 training steps through the fused pass lower the loss on a
    fixed batch (end-to-end sanity of the matrix-core forward/backward)."""
    print("[case] mfma training steps", file=sys.stderr, flush=True)
    from ffields import FusedAdam
    m = make_bf16_model(seed=5)
    B = 512
    rays = make_rays(B, seed=31).cuda()
    gt = torch.rand((B, 3), generator=torch.Generator().manual_seed(32)).cuda() * 0.5
    jit = torch.full((B,), 0.5, device="cuda")
    opt = FusedAdam(m.get_optparam_groups(0.001, 0.02), betas=(0.9, 0.99), shadow_of=m.shadow_lookup)
    losses = []
    for _ in range(30):
        ls = m.render_train_step(rays, gt, white_bg=True, jitter=jit)
        losses.append(ls.item())
        opt.step(zero_grad=True)
    assert losses[-1] < 0.7 * losses[0], losses


@pytest.mark.parametrize("dtype", ["fp32", "bf16"])
def test_render_error_vs_fp64_reference(dtype):
    """Error of each form's render against an FP64 run of the reference path
    (oracle renderer in float64 on the same FP32 master tables and weights,
    no occupancy grid, injected jitter).  FP32 form: within 1e-4 of the FP64
    render; BF16 tables + BF16 matrix-core decoder: mean absolute error below
    1e-2 (the measured values are printed)."""
    from oracle import OracleRenderer
    print(f"[case] fp64 render {dtype}", file=sys.stderr, flush=True)
    m = make_bf16_model(seed=7) if dtype == "bf16" else None
    if m is None:
        from helpers import randomize
        m = FactorFields(nerf_small_cfg(), "cuda")
        randomize(m, 7, scale=0.5)
        with torch.no_grad():
            m.linear_mat.backbone[0].weight.mul_(20.0)
        ref = make_bf16_model(seed=7)                      # same density shift derivation
        m.cfg.renderer.density_shift = ref.cfg.renderer.density_shift
    of = oracle_of(m, dtype=torch.float64)
    B = 512
    rays = make_rays(B, seed=41)
    jit = torch.rand(B, generator=torch.Generator().manual_seed(42))
    with torch.no_grad():
        rgb, _, _ = m(rays.cuda(), white_bg=True, is_train=False, jitter=jit.cuda())
    orr = OracleRenderer(of, None, None, torch.tensor(m._step, dtype=torch.float64), m.nSamples)
    ref_rgb, _, aux = orr.forward(rays.double(), jitter=jit.double(), white_bg=True)
    assert float(aux["acc"].max()) > 0.5
    err = (rgb.double().cpu() - ref_rgb).abs()
    print(f"[fp64 render] {dtype}: max abs error {err.max().item():.3e}, mean {err.mean().item():.3e}",
          file=sys.stderr, flush=True)
    if dtype == "fp32":
        assert err.max().item() < 1e-4
    else:
        assert err.mean().item() < 1e-2
