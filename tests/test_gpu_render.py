"""GPU parity of the render path against the oracle renderer
(models/FactorFields.py:586-602, :91-110, :843-898 restated in oracle/render.py).

Bars: sample counts, candidate indices, positions, t values and occupancy
masks bit-exact (integer/mask work, and the FP32 arithmetic is replicated
operation by operation); sigma/rgb/rgb_map/depth within 1e-5 relative
(absolute floor 1e-6 of the max); gradients within 1e-4 relative (floor
1e-5 of the max).  Jitter is injected so both sides see the same draws."""
import sys

import pytest
import torch

from helpers import assert_close, oracle_of, randomize
from ffields import FactorFields
from ffields.config import AttrDict
from ffields.render import AlphaGridMask
from oracle import OracleRenderer, alpha_grid_sample, sample_point
from oracle.render import sh_rgb

pytestmark = pytest.mark.gpu


def nerf_small_cfg(aabb=None):
    aabb = aabb or [[-1.2, -1.0, -1.4], [1.3, 1.1, 0.9]]
    return AttrDict.wrap(dict(
        defaults=dict(mode="reconstruction"),
        model=dict(basis_dims=[4, 4, 4, 2, 2, 2], basis_resos=[9, 11, 13, 7, 6, 5],
                   freq_bands=[1.0, 1.7, 2.9, 3.3, 4.1, 5.0], coeff_reso=8, coeff_type="grid",
                   basis_type="grid", coef_mode="bilinear", basis_mode="bilinear",
                   basis_mapping="sawtooth", param_derivation="literal", out_dim=28, hidden_dim=64,
                   num_layers=2, coef_init=1.0, total_params=0, T_basis=0, T_coeff=0,
                   with_dropout=False),
        dataset=dict(aabb=aabb, scene_reso=1024, is_unbound=False),
        renderer=dict(fea2denseAct="softplus", density_shift=-10, distance_scale=25.0, step_ratio=0.5,
                      max_samples=512, rayMarch_weight_thres=1e-3, alphaMask_thres=0.02,
                      shadingMode="SH"),
        training=dict(volume_resoInit=40, volume_resoFinal=40)))


def make_rays(n, seed=0, radius=4.0):
    g = torch.Generator().manual_seed(seed)
    o = torch.randn((n, 3), generator=g)
    o = o / o.norm(dim=-1, keepdim=True) * radius
    tgt = (torch.rand((n, 3), generator=g) * 2 - 1) * 1.2
    d = tgt - o
    d = d / d.norm(dim=-1, keepdim=True)
    d[:7, 1] = 0.0                      # exercise the d == 0 branch of sample_point
    d[:7] = d[:7] / d[:7].norm(dim=-1, keepdim=True)
    return torch.cat([o, d], -1)


def make_model(seed=0, occupancy=True):
    m = FactorFields(nerf_small_cfg(), "cuda")
    randomize(m, seed, scale=0.5)
    with torch.no_grad():
        m.linear_mat.backbone[0].weight.mul_(20.0)    # density varies in space
    # centre the density shift so that about 40% of the box is occupied
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


def oracle_renderer(m):
    of = oracle_of(m)
    am = m.alphaMask
    return OracleRenderer(of, None if am is None else am.alpha_volume.cpu()[0, 0],
                          None if am is None else am.aabb.cpu(),
                          torch.tensor(m._step, dtype=torch.float32), m.nSamples)


@pytest.mark.parametrize("occupancy", [True, False])
@pytest.mark.parametrize("train", [True, False])
def test_sampler_bit_exact(occupancy, train):
    print(f"[case] sampler occ={occupancy} train={train}", file=sys.stderr, flush=True)
    m = make_model(occupancy=occupancy)
    B = 777
    rays = make_rays(B)
    jit = torch.rand(B, generator=torch.Generator().manual_seed(3)) if train else None
    Ns = m.nSamples
    buf = m.render_buffers(B, Ns)
    m.render_samples(rays.cuda(), Ns, None if jit is None else jit.cuda(), buf)
    torch.cuda.synchronize()
    off = buf["ray_off"].cpu()
    M = int(off[-1])
    xyz, z_vals, inner = sample_point(rays[:, :3], rays[:, 3:6], m.aabb.cpu(),
                                      torch.tensor(m._step, dtype=torch.float32), Ns, jit)
    valid = inner.clone()
    if occupancy:
        am = m.alphaMask
        valid[inner.clone()] = alpha_grid_sample(am.alpha_volume.cpu()[0, 0], am.aabb.cpu(), xyz[inner]) > 0.5
    counts = valid.sum(-1)
    assert torch.equal(off[1:] - off[:-1], counts.int())
    rid, cid = torch.nonzero(valid, as_tuple=True)
    assert M == len(rid)
    assert torch.equal(buf["ray"][:M].cpu().long(), rid)
    assert torch.equal(buf["idx"][:M].cpu().long(), cid)
    assert torch.equal(buf["xyz"][:M].cpu(), xyz[valid])
    assert torch.equal(buf["z"][:M].cpu(), z_vals[valid])
    dists = torch.cat((z_vals[:, 1:] - z_vals[:, :-1], torch.zeros_like(z_vals[:, :1])), dim=-1) * 25.0
    assert torch.equal(buf["dist"][:M].cpu(), dists[valid])
    if occupancy:
        assert 0 < M < int(inner.sum())                 # the mask removed something


def test_field_head_matches_oracle():
    m = make_model()
    B = 300
    rays = make_rays(B, seed=5)
    jit = torch.rand(B, generator=torch.Generator().manual_seed(6))
    buf = m.render_buffers(B, m.nSamples)
    m.render_samples(rays.cuda(), m.nSamples, jit.cuda(), buf)
    torch.cuda.synchronize()
    M = int(buf["ray_off"][-1])
    of = oracle_of(m)
    x = buf["xyz"][:M].cpu()
    feat = of.field(x)
    sig = of.basis2density(feat[:, 0])
    dirs = rays[buf["ray"][:M].cpu().long(), 3:6]
    rgb = sh_rgb(dirs, feat[:, 1:28])
    assert_close(buf["sigma"][:M], sig, 1e-5, "sigma")
    assert_close(buf["rgb"][:M], rgb, 1e-5, "rgb")


@pytest.mark.parametrize("white_bg", [True, False])
def test_render_forward_matches_oracle(white_bg):
    m = make_model()
    orr = oracle_renderer(m)
    B = 512
    rays = make_rays(B, seed=7)
    jit = torch.rand(B, generator=torch.Generator().manual_seed(8))
    rgb, depth, _ = m(rays.cuda(), white_bg=white_bg, is_train=False, jitter=jit.cuda())
    ref_rgb, ref_depth, aux = orr.forward(rays, jitter=jit, white_bg=white_bg)
    assert float(aux["acc"].max()) > 0.5              # the test scene is not empty
    assert float(aux["acc"].min()) < 0.5
    assert_close(rgb, ref_rgb, 1e-5, "rgb_map")
    assert_close(depth, ref_depth, 1e-5, "depth")


@pytest.mark.parametrize("white_bg", [True, False])
def test_render_backward_matches_oracle(white_bg):
    m = make_model(seed=2)
    orr = oracle_renderer(m)
    of = orr.field.requires_grad_(True)
    B = 400
    rays = make_rays(B, seed=9)
    jit = torch.rand(B, generator=torch.Generator().manual_seed(10))
    gt = torch.rand((B, 3), generator=torch.Generator().manual_seed(11))
    rgb, _, _ = m(rays.cuda(), white_bg=white_bg, is_train=False, jitter=jit.cuda())
    ((rgb - gt.cuda()) ** 2).mean().backward()
    ref_rgb, _, _ = orr.forward(rays, jitter=jit, white_bg=white_bg)
    ((ref_rgb - gt) ** 2).mean().backward()
    for k, (t, o) in enumerate(zip(list(m.basises) + list(m.coeffs), of.basises + of.coeffs)):
        assert_close(t.grad.contiguous(), o.grad, 1e-4, f"table {k}", atol_frac=1e-5)
    for l, (w, b) in enumerate(of.mlp):
        lin = m.linear_mat.backbone[l]
        assert_close(lin.weight.grad, w.grad, 1e-4, f"W{l}", atol_frac=1e-5)
        if b is not None:
            assert_close(lin.bias.grad, b.grad, 1e-4, f"b{l}", atol_frac=1e-5)


def test_fused_train_step_matches_autograd():
    m = make_model(seed=4)
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


def test_empty_rays_and_misses():
    m = make_model()
    # rays pointing away from the box produce no samples; the pass must still run
    rays = make_rays(64, seed=15)
    rays[:, 3:6] = -rays[:, 3:6]
    rgb, depth, _ = m(rays.cuda(), white_bg=True, is_train=False)
    assert torch.allclose(rgb, torch.ones_like(rgb))
    assert torch.allclose(depth, torch.zeros_like(depth))


def test_sh_render_module_matches_oracle():
    """SHRender on its own (ff_sh_rgb / ff_sh_rgb_bwd through the
    ffields::sh_render custom op) against the oracle's sh_rgb
    (models/sh.py:34-85 + sigmoid): rgb to 1e-6, feature gradients to 1e-5."""
    print("[case] SHRender", file=sys.stderr, flush=True)
    from ffields.render import SHRender
    g = torch.Generator().manual_seed(31)
    d = torch.randn((3000, 3), generator=g)
    d = d / d.norm(dim=-1, keepdim=True)
    f = torch.randn((3000, 27), generator=g)
    r = torch.rand((3000, 3), generator=g)
    fc = f.cuda().requires_grad_(True)
    out = SHRender()(d.cuda(), fc)
    (out * r.cuda()).sum().backward()
    fr = f.clone().requires_grad_(True)
    ref = sh_rgb(d, fr)
    (ref * r).sum().backward()
    assert_close(out, ref, 1e-6, "SH rgb")
    assert_close(fc.grad, fr.grad, 1e-5, "SH feature grad", atol_frac=1e-6)
    # batched leading dims
    out2 = SHRender()(d.cuda().reshape(30, 100, 3), fc.detach().reshape(30, 100, 27))
    assert torch.equal(out2.reshape(-1, 3), out.detach())
