"""GPU parity of the reference's NeRF colour module MLPRender_Fea
(models/FactorFields.py:162-203, positional_encoding :74-79), the app-sample
selection that feeds it (:878-885), linear_mat's input dropout (:150-151) and
upsample_volume_grid (:701-708), against the oracle (oracle/render.py
mlp_render_fea, OracleRenderer with render_layers; oracle/field.py
linear_mat with an injected mask).

Bars: FP32 forward within 1e-5 relative (floor 1e-6 of the largest
magnitude), gradients within 1e-4 relative (floor 1e-5); app-sample lists and
dropout masks bit-exact."""
import sys

import pytest
import torch
import torch.nn.functional as F

from helpers import assert_close, assert_close_fp64, oracle_of, randomize
from test_gpu_render import make_rays, nerf_small_cfg
from ffields import FactorFields
from ffields.render import AlphaGridMask, MLPRender_Fea
from oracle import OracleRenderer, mlp_render_fea, raw2alpha, sample_point, alpha_grid_sample

pytestmark = pytest.mark.gpu


def color_module(view_pe=6, seed=0, scale=1.0):
    torch.manual_seed(seed)
    rm = MLPRender_Fea(31, num_layers=3, hidden_dim=128, viewpe=view_pe, feape=2).cuda()
    if scale != 1.0:
        with torch.no_grad():
            for p in rm.parameters():
                p.mul_(scale)
    return rm


def layers_of(rm, dtype=torch.float32, grad=False):
    out = []
    for lin in rm.mlp:
        w = lin.weight.detach().cpu().to(dtype).requires_grad_(grad)
        b = None if lin.bias is None else lin.bias.detach().cpu().to(dtype).requires_grad_(grad)
        out.append((w, b))
    return out


def rand_dirs(n, g):
    d = torch.randn((n, 3), generator=g)
    return d / d.norm(dim=-1, keepdim=True)


@pytest.mark.parametrize("view_pe", [6, 2])
def test_color_module_matches_oracle(view_pe):
    """renderModule(viewdirs, features): rgb, feature gradients and the five
    weight gradients against the FP32 oracle; rows not a multiple of the
    32-sample tile."""
    print(f"[case] color module view_pe={view_pe}", file=sys.stderr, flush=True)
    rm = color_module(view_pe, seed=1, scale=1.5)
    g = torch.Generator().manual_seed(2)
    N = 1000
    feat = torch.randn((N, 31), generator=g) * 2.0
    dirs = rand_dirs(N, g)
    R = torch.randn((N, 3), generator=g)
    fc = feat.cuda().requires_grad_(True)
    rgb = rm(dirs.cuda(), fc)
    (rgb * R.cuda()).sum().backward()
    lay = layers_of(rm, grad=True)
    fo = feat.clone().requires_grad_(True)
    ref = mlp_render_fea(dirs, fo, lay, viewpe=view_pe, feape=2)
    (ref * R).sum().backward()
    assert_close(rgb, ref, 1e-5, "rgb")
    assert_close(fc.grad, fo.grad, 1e-4, "d features", atol_frac=1e-5)
    for k, ((w, b), lin) in enumerate(zip(lay, rm.mlp)):
        assert_close(lin.weight.grad, w.grad, 1e-4, f"dW{k}", atol_frac=1e-5)
        if b is not None:
            assert_close(lin.bias.grad, b.grad, 1e-4, f"db{k}", atol_frac=1e-5)
    # error against an FP64 run of the reference path
    ref64 = mlp_render_fea(dirs.double(), feat.double(), layers_of(rm, torch.float64), viewpe=view_pe)
    err = (rgb.detach().cpu().double() - ref64).abs().max().item()
    print(f"[fp64] MLPRender_Fea rgb max abs error {err:.2e}", file=sys.stderr, flush=True)
    assert err < 1e-5


def test_color_module_tiny_and_empty():
    rm = color_module(6, seed=3)
    for N in (1, 31, 33):
        g = torch.Generator().manual_seed(N)
        feat, dirs = torch.randn((N, 31), generator=g), rand_dirs(N, g)
        rgb = rm(dirs.cuda(), feat.cuda())
        assert_close(rgb, mlp_render_fea(dirs, feat, layers_of(rm)), 1e-5, f"rgb N={N}")
    out = rm(torch.zeros((0, 3), device="cuda"), torch.zeros((0, 31), device="cuda"))
    assert out.shape == (0, 3)


def fea_cfg(view_pe=6):
    cfg = nerf_small_cfg()
    cfg.model.out_dim = 32
    cfg.renderer.update(shadingMode="MLP_Fea", num_layers=3, hidden_dim=128, view_pe=view_pe, fea_pe=2)
    return cfg


def make_fea_model(seed=0, occupancy=True, dtype="fp32"):
    m = FactorFields(fea_cfg(), "cuda", table_dtype=dtype)
    randomize(m, seed, scale=0.5)
    with torch.no_grad():
        m.linear_mat.backbone[0].weight.mul_(20.0)          # density varies in space
        m.linear_mat.backbone[1].weight[1:].mul_(0.005)     # features of order one
    of = oracle_of(m)
    g = torch.Generator().manual_seed(seed + 7)
    a = m.aabb.cpu()
    o0 = of.field(a[0] + (a[1] - a[0]) * torch.rand((4096, 3), generator=g))[:, 0]
    m.cfg.renderer.density_shift = float(-torch.quantile(o0, 0.6))
    if occupancy:
        g = torch.Generator().manual_seed(seed + 100)
        vol = (torch.rand((29, 41, 37), generator=g) < 0.6).float()
        aabb = torch.tensor([[-1.1, -0.9, -1.3], [1.2, 1.0, 0.8]])
        m.alphaMask = AlphaGridMask(torch.device("cuda"), aabb, vol)
    return m


def fea_oracle(m, grad=False):
    of = oracle_of(m).requires_grad_(grad)
    am = m.alphaMask
    return OracleRenderer(of, None if am is None else am.alpha_volume.cpu()[0, 0],
                          None if am is None else am.aabb.cpu(), torch.tensor(m._step, dtype=torch.float32),
                          m.nSamples, render_layers=layers_of(m.renderModule, grad=grad),
                          viewpe=m.renderModule.viewpe, feape=m.renderModule.feape)


def fea_oracle_fp64(m, grad=False):
    of = oracle_of(m, dtype=torch.float64)
    am = m.alphaMask
    return OracleRenderer(of, None if am is None else am.alpha_volume.cpu()[0, 0].double(),
                          None if am is None else am.aabb.cpu().double(),
                          torch.tensor(m._step, dtype=torch.float64), m.nSamples,
                          render_layers=layers_of(m.renderModule, torch.float64, grad=grad),
                          viewpe=m.renderModule.viewpe, feape=m.renderModule.feape)


def test_construct_from_nerf_yaml():
    """FactorFields with the values of configs/defaults.yaml + configs/nerf.yaml
    (reconstruction mode, MLP_Fea with view_pe 6, fea_pe 2, 3 layers of 128,
    out_dim 32) constructs, with the reference's renderModule keys."""
    from test_gpu_nerf_loop import nerf_yaml_cfg
    m = FactorFields(nerf_yaml_cfg(), "cuda")
    keys = [k for k in m.state_dict() if k.startswith("renderModule.")]
    assert keys == ["renderModule.mlp.0.weight", "renderModule.mlp.0.bias", "renderModule.mlp.1.weight",
                    "renderModule.mlp.1.bias", "renderModule.mlp.2.weight"]
    assert tuple(m.renderModule.mlp[0].weight.shape) == (128, 194)
    assert m.renderModule.in_mlpC == 194


def test_app_selection_matches_oracle():
    """The app samples (weight > rayMarch_weight_thres among the valid ones)
    equal the oracle's ray_valid_new mask, in ray-major order."""
    m = make_fea_model(seed=1)
    orr = fea_oracle(m)
    B = 300
    rays = make_rays(B, seed=5)
    jit = torch.rand(B, generator=torch.Generator().manual_seed(6))
    buf = m.render_buffers(B, m.nSamples)
    m.render_samples(rays.cuda(), m.nSamples, jit.cuda(), buf)
    torch.cuda.synchronize()
    _, _, aux = orr.forward(rays, jitter=jit, white_bg=True)
    M = int(buf["ray_off"][-1])
    A = int(buf["app_off"][-1])
    valid, app = aux["ray_valid"], aux["app_mask"]
    assert 0 < A < M
    # position of each valid sample in the packed list
    pos = torch.full(valid.shape, -1, dtype=torch.long)
    pos[valid] = torch.arange(M)
    want_rows = pos[app]
    assert A == len(want_rows)
    assert torch.equal(buf["app_rows"][:A].cpu().long(), want_rows)
    want_pos = torch.full((M,), -1, dtype=torch.long)
    want_pos[want_rows] = torch.arange(A)
    assert torch.equal(buf["app_pos"][:M].cpu().long(), want_pos)
    assert torch.equal((buf["app_off"][1:] - buf["app_off"][:-1]).cpu(), app.sum(-1).int())


@pytest.mark.parametrize("white_bg", [True, False])
def test_render_mlp_fea_matches_oracle(white_bg):
    """FactorFields.forward with the MLP_Fea colour module: rgb_map, depth and
    the gradients of every table, the decoder and the colour module."""
    print(f"[case] render MLP_Fea white_bg={white_bg}", file=sys.stderr, flush=True)
    m = make_fea_model(seed=2)
    orr = fea_oracle(m, grad=True)
    of = orr.field
    B = 400
    rays = make_rays(B, seed=9)
    jit = torch.rand(B, generator=torch.Generator().manual_seed(10))
    gt = torch.rand((B, 3), generator=torch.Generator().manual_seed(11))
    rgb, depth, _ = m(rays.cuda(), white_bg=white_bg, is_train=False, jitter=jit.cuda())
    ((rgb - gt.cuda()) ** 2).mean().backward()
    ref_rgb, ref_depth, aux = orr.forward(rays, jitter=jit, white_bg=white_bg)
    ((ref_rgb - gt) ** 2).mean().backward()
    assert int(aux["app_mask"].sum()) > 0
    ref64 = fea_oracle_fp64(m).forward(rays.double(), jitter=jit.double(), white_bg=white_bg)[0]
    a, b = rgb.detach().cpu().double(), ref_rgb.detach().double()
    print(f"[fp64] MLP_Fea render max abs rgb error: kernels {(a - ref64).abs().max():.2e}, "
          f"FP32 oracle {(b - ref64).abs().max():.2e}", file=sys.stderr, flush=True)
    bad = ((a - b).abs() > 1e-5 * b.abs() + 1e-6).nonzero()
    for r, c in bad[:12].tolist():
        print(f"[diff] ray {r} ch {c}: kernels {a[r, c]:.9f} oracle32 {b[r, c]:.9f} oracle64 {ref64[r, c]:.9f} "
              f"app {int(aux['app_mask'][r].sum())} valid {int(aux['ray_valid'][r].sum())}", file=sys.stderr)
    assert_close(rgb, ref_rgb, 1e-5, "rgb_map")
    assert_close(depth, ref_depth, 1e-5, "depth")
    for k, (t, o) in enumerate(zip(list(m.basises) + list(m.coeffs), of.basises + of.coeffs)):
        assert_close(t.grad.contiguous(), o.grad, 1e-4, f"table {k}", atol_frac=1e-5)
    for l, (w, b) in enumerate(of.mlp):
        lin = m.linear_mat.backbone[l]
        assert_close(lin.weight.grad, w.grad, 1e-4, f"decoder W{l}", atol_frac=1e-5)
        if b is not None:
            assert_close(lin.bias.grad, b.grad, 1e-4, f"decoder b{l}", atol_frac=1e-5)
    # the colour weights' gradients are sums over the app samples whose terms
    # cancel in places: checked against FP64 autograd, at the bar or as close
    # as the FP32 oracle gets
    o64 = fea_oracle_fp64(m, grad=True)
    ((o64.forward(rays.double(), jitter=jit.double(), white_bg=white_bg)[0] - gt.double()) ** 2).mean().backward()
    for l, ((w, b), (w64, b64)) in enumerate(zip(orr.render_layers, o64.render_layers)):
        lin = m.renderModule.mlp[l]
        assert_close_fp64(lin.weight.grad, w.grad, w64.grad, 1e-4, f"colour W{l}", atol_frac=1e-5)
        if b is not None:
            assert_close_fp64(lin.bias.grad, b.grad, b64.grad, 1e-4, f"colour b{l}", atol_frac=1e-5)


@pytest.mark.parametrize("dtype", ["fp32", "bf16"])
def test_fused_train_step_mlp_fea_matches_autograd(dtype):
    """render_train_step (no autograd) gives the autograd path's loss and
    gradients, colour module included, for FP32 and BF16 tables."""
    m = make_fea_model(seed=4, dtype=dtype)
    B = 640
    rays = make_rays(B, seed=12).cuda()
    jit = torch.rand(B, generator=torch.Generator().manual_seed(13)).cuda()
    gt = torch.rand((B, 3), generator=torch.Generator().manual_seed(14)).cuda()
    rgb, _, _ = m(rays, white_bg=True, is_train=False, jitter=jit)
    loss = ((rgb - gt) ** 2).mean()
    loss.backward()
    ref = {n: p.grad.clone() for n, p in m.named_parameters()}
    m.zero_grad(set_to_none=True)
    ls = m.render_train_step(rays, gt, white_bg=True, jitter=jit)
    assert abs(ls.item() / (3 * B) - loss.item()) <= 1e-5 * loss.item()
    for n, p in m.named_parameters():
        assert_close(p.grad, ref[n], 1e-5, n, atol_frac=1e-6)


def test_bf16_mlp_fea_render_error_vs_fp32_oracle():
    """BF16 tables + matrix-core decoder feeding the FP32 colour module,
    against the FP32 oracle (the BF16 form's stated error)."""
    m = make_fea_model(seed=5, dtype="bf16")
    orr = fea_oracle(m)
    B = 512
    rays = make_rays(B, seed=21)
    jit = torch.rand(B, generator=torch.Generator().manual_seed(22))
    with torch.no_grad():
        rgb, _, _ = m(rays.cuda(), white_bg=True, is_train=False, jitter=jit.cuda())
    ref, _, _ = orr.forward(rays, jitter=jit, white_bg=True)
    err = (rgb.cpu() - ref).abs()
    print(f"[bf16] MLP_Fea render max abs rgb error {err.max().item():.2e} mean {err.mean().item():.2e}",
          file=sys.stderr, flush=True)
    assert err.max().item() < 5e-2 and err.mean().item() < 5e-3


# ---------------------------------------------------------------- dropout
def test_dropout_injected_mask_matches_oracle():
    """linear_mat(x, is_train=True) with with_dropout and an injected mask:
    outputs and gradients (input and weights) equal the oracle's
    F.dropout-then-MLP with the same mask."""
    from helpers import small_cfg
    cfg = small_cfg(mode="image", D=2, out_dim=3)
    cfg.model.with_dropout = True
    m = FactorFields(cfg, "cuda")
    randomize(m, 3)
    g = torch.Generator().manual_seed(4)
    x = torch.randn((777, 18), generator=g)
    keep = torch.rand((777, 18), generator=g) < 0.9
    xc = x.cuda().requires_grad_(True)
    out = m.linear_mat(xc, is_train=True, keep=keep.cuda())
    R = torch.randn((777, 3), generator=g)
    (out * R.cuda()).sum().backward()
    of = oracle_of(m).requires_grad_(True)
    xo = x.clone().requires_grad_(True)
    ref = of.linear_mat(xo, keep=keep)
    (ref * R).sum().backward()
    assert_close(out, ref, 1e-5, "dropout forward")
    assert_close(xc.grad, xo.grad, 1e-4, "dropout d input", atol_frac=1e-5)
    for l, (w, b) in enumerate(of.mlp):
        assert_close(m.linear_mat.backbone[l].weight.grad, w.grad, 1e-4, f"W{l}", atol_frac=1e-5)
    # dropped inputs do not reach the output: the same call without them
    x2 = x * keep
    with torch.no_grad():
        out2 = m.linear_mat(x2.cuda() * (1.0 / 0.9), is_train=False)
    assert_close(out.detach(), out2, 1e-5, "dropout == masked input scaled")


def test_dropout_device_rng():
    """Masks drawn on the device: keep rate 1 - p, a fresh mask per call, the
    same mask in the backward, exact scaling by 1 / (1 - p)."""
    from ffields.ops import dropout
    n = 1 << 20
    x = torch.ones(n, device="cuda", requires_grad=True)
    y = dropout(x, 0.1, 12345, None)
    kept = (y != 0)
    rate = kept.float().mean().item()
    assert abs(rate - 0.9) < 0.002, rate
    assert torch.all(y[kept] == torch.tensor(1.0 / 0.9, dtype=torch.float32).cuda())
    y.backward(torch.ones_like(y) * 2.0)
    assert torch.equal(x.grad != 0, kept)
    y2 = dropout(x.detach(), 0.1, 12345, None)
    assert torch.equal(y2, y.detach())                  # same seed, same mask
    y3 = dropout(x.detach(), 0.1, 12346, None)
    agree = ((y3 != 0) == kept).float().mean().item()
    assert 0.80 < agree < 0.84                          # independent masks agree on 0.9^2 + 0.1^2 = 82 %


# ------------------------------------------------------ upsample_volume_grid
def test_upsample_volume_grid_updates_render_params():
    """models/FactorFields.py:693-708: stepSize = mean(aabbSize / (gridSize - 1))
    * step_ratio and nSamples = int(|aabbSize| / stepSize) + 1 for the new grid."""
    m = make_fea_model(seed=6, occupancy=False)
    m.upsample_volume_grid([150, 120, 90])
    size = (m.aabb[1] - m.aabb[0]).cpu()
    step = torch.mean(size / (torch.tensor([150, 120, 90]) - 1)) * 0.5
    assert abs(m.stepSize.item() - step.item()) <= 1e-7 * step.item()
    assert m.nSamples == int((torch.sqrt(torch.sum(size ** 2)) / step).item()) + 1
    assert m.gridSize.tolist() == [150, 120, 90]
