"""Full-size cases of the benchmarked workloads, through properties that do
not depend on the batch size.

* Every ray and every point is computed independently, so a subset evaluated
  on its own equals the same entries of the full batch bit for bit (packed
  sample offsets, capacities and tile assignment change with the batch; the
  values must not).  The subset is also checked against the oracle.
* The dictionary's table gradients do not depend on the order of its 2^22
  pixels, up to FP32 summation order (atomics and per-wave partial sums):
  the bench's image-and-cell order against draw order.  The same holds for
  the SDF and image regressions' gradients and for the proposal + BF16 NeRF
  training pass (its random draws injected), whose training render permutes
  bit for bit with the rays.

Shapes: NeRF 4,096 rays of the 800x800 synthetic cameras, 3-D grid basis
[4,4,4,2,2,2]@[32..128] (reference parameter derivation), 32^3x18
coefficient grid, 128^3 occupancy grid, uniform sampler at the reference's
step; SDF 2^20 points, VM coefficient 3x128^2 + 3x128; image 65,536 pixels
of a 512x512 image, 64x64x18 coefficient grid; dictionary 1,024
images of 256x256, basis [32,32,32,16,16,16], 1,024 x 32x32x144 BF16
coefficient grid, decoder 144-64-3, 2^22 pixels."""
import sys

import pytest
import torch

from ffields import FactorFields
from ffields import workloads as W
from ffields.render import AlphaGridMask
from helpers import assert_close, oracle_of, randomize
from test_gpu_dict import emulate
from test_gpu_render import oracle_renderer

pytestmark = pytest.mark.gpu


def case(name):
    print(f"[case] {name}", file=sys.stderr, flush=True)


def assert_close_gpu(a, b, rtol, what, atol_frac):
    """assert_close's bar, evaluated on the device (tables of 10^8 entries)."""
    assert a.shape == b.shape, (what, a.shape, b.shape)
    scale = float(b.abs().max()) or 1.0
    err = (a - b).abs()
    bad = err > rtol * b.abs() + atol_frac * scale
    assert not bool(bad.any()), (what, float(err.max()), scale, int(bad.sum()))


def nerf_model(dtype="fp32"):
    m = FactorFields(W.nerf_cfg(), "cuda", table_dtype=dtype)
    randomize(m, 3, scale=0.5)
    with torch.no_grad():
        m.linear_mat.backbone[0].weight.mul_(20.0)    # density varies in space
    of = oracle_of(m)
    g = torch.Generator().manual_seed(4)
    a = m.aabb.cpu()
    o0 = of.field(a[0] + (a[1] - a[0]) * torch.rand((4096, 3), generator=g))[:, 0]
    m.cfg.renderer.density_shift = float(-torch.quantile(o0, 0.6))
    vol = (torch.rand((128, 128, 128), generator=g) < 0.6).float()
    m.alphaMask = AlphaGridMask(torch.device("cuda"), a, vol)
    return m


@pytest.mark.parametrize("dtype", ["fp32", "bf16"])
def test_nerf_render_full_batch(dtype):
    """FP32 form: the subset within 1e-5 of the FP32 oracle.  BF16 tables and
    matrix-core decoder: mean absolute rgb error against the same FP32
    oracle below 1e-2, the bar of the small-size render test."""
    case(f"nerf full batch render ({dtype}): 4096 rays, 128^3 occupancy")
    m = nerf_model(dtype)
    dev = torch.device("cuda")
    scene = W.NerfScene(device=dev)
    rays = scene.random_rays(4096, torch.Generator(device=dev).manual_seed(5))
    jit = torch.rand(4096, generator=torch.Generator(device=dev).manual_seed(6), device=dev)
    with torch.no_grad():
        rgb, depth, _ = m(rays, white_bg=True, is_train=False, jitter=jit)
        idx = torch.randperm(4096, generator=torch.Generator().manual_seed(7))[:48].to(dev)
        rgb_s, depth_s, _ = m(rays[idx].contiguous(), white_bg=True, is_train=False, jitter=jit[idx].contiguous())
    assert torch.equal(rgb[idx], rgb_s) and torch.equal(depth[idx], depth_s)
    orr = oracle_renderer(m)
    ref_rgb, ref_depth, aux = orr.forward(rays[idx].cpu(), jitter=jit[idx].cpu(), white_bg=True)
    assert float(aux["acc"].max()) > 0.5                # the rays see the field
    if dtype == "fp32":
        assert_close(rgb_s, ref_rgb, 1e-5, "rgb_map")
        assert_close(depth_s, ref_depth, 1e-5, "depth")
    else:
        err = (rgb_s.double().cpu() - ref_rgb.double()).abs()
        print(f"[bf16 full render] max abs error {err.max().item():.3e}, mean {err.mean().item():.3e}",
              file=sys.stderr, flush=True)
        assert err.mean().item() < 1e-2


def dict_model():
    m = FactorFields(W.dict_cfg(1024, 256), "cuda", table_dtype="bf16")
    randomize(m, 8)
    assert m.fused_supported()
    return m


def dict_batch(n, seed):
    dev = torch.device("cuda")
    gen = torch.Generator(device=dev).manual_seed(seed)
    img, p = W.image_set_batch(n, 1024, 256, gen, dev, order=W.cell_order(256, 32, dev))
    return torch.stack([(p % 256).float() + 0.5, (p // 256).float() + 0.5, img.float() + 0.5], -1).contiguous()


def test_dict_forward_full_batch():
    case("dict full batch forward: 2^22 pixels, 1024 images")
    m = dict_model()
    n = 1 << 22
    x = dict_batch(n, 9)
    with torch.no_grad():
        out = m.field(x)
        idx = torch.randperm(n, generator=torch.Generator().manual_seed(10))[:2048].to(x.device)
        sub = m.field(x[idx].contiguous())
    assert torch.equal(out[idx], sub)
    ref, *_ = emulate(m, oracle_of(m, bf16_tables=True).get_coding(x[idx].cpu())[0])
    assert_close(sub, ref, 2e-3, "out", atol_frac=2e-3)


def test_dict_gradients_full_batch_order():
    case("dict full batch gradients: cell order vs draw order")
    m = dict_model()
    n = 1 << 22
    x = dict_batch(n, 11)
    t = torch.rand((n, 3), generator=torch.Generator(device=x.device).manual_seed(12), device=x.device)
    perm = torch.randperm(n, generator=torch.Generator(device=x.device).manual_seed(13), device=x.device)
    params = list(m.basises) + list(m.coeffs) + list(m.linear_mat.parameters())
    grads, losses = [], []
    for xi, ti in [(x, t), (x[perm].contiguous(), t[perm].contiguous())]:
        m.zero_grad(set_to_none=True)
        losses.append(float(m.regression_step(xi, ti, loss="mse")))
        grads.append([p.grad.detach().clone() for p in params])
    assert abs(losses[0] - losses[1]) <= 1e-5 * losses[0]
    for k, (a, b) in enumerate(zip(*grads)):
        assert float(b.abs().max()) > 0, f"param {k} has no gradient"
        assert_close_gpu(a, b, 1e-4, f"param {k}", atol_frac=1e-5)


def test_sdf_full_batch():
    """SDF regression (VM coefficient 3x128^2 + 3x128, 18 channels, MLP
    18-64-1, L1) at 2^20 points: subset invariance and the oracle for the
    forward; gradients independent of the batch order."""
    case("sdf full batch: 2^20 points, VM coefficient")
    m = FactorFields(W.sdf_cfg(), "cuda")
    randomize(m, 14, scale=0.5)
    assert m.fused_supported()
    dev = torch.device("cuda")
    n = 1 << 20
    g = torch.Generator(device=dev).manual_seed(15)
    x = (torch.rand((n, 3), generator=g, device=dev) * 2 - 1).contiguous()
    t = (torch.rand((n, 1), generator=g, device=dev) - 0.5).contiguous()
    with torch.no_grad():
        out = m.field(x)
        idx = torch.randperm(n, generator=torch.Generator().manual_seed(16))[:4096].to(dev)
        sub = m.field(x[idx].contiguous())
    assert torch.equal(out[idx], sub)
    assert_close(sub, oracle_of(m).field(x[idx].cpu()), 1e-5, "sdf out")
    perm = torch.randperm(n, generator=torch.Generator(device=dev).manual_seed(17), device=dev)
    params = list(m.basises) + list(m.coeffs) + list(m.linear_mat.parameters())
    grads = []
    for xi, ti in [(x, t), (x[perm].contiguous(), t[perm].contiguous())]:
        m.zero_grad(set_to_none=True)
        m.regression_step(xi, ti, loss="l1")
        grads.append([p.grad.detach().clone() for p in params])
    for k, (a, b) in enumerate(zip(*grads)):
        assert float(b.abs().max()) > 0, f"param {k} has no gradient"
        assert_close_gpu(a, b, 1e-4, f"param {k}", atol_frac=1e-5)


def test_proposal_train_step_full_batch():
    """The benchmarked training pass (BF16 tables, learned proposal with a
    64^3x8 grid placing 48 samples on each of 4,096 rays) with its random
    draws injected: permuting the rays, with their draws and targets,
    permutes the training render bit for bit and leaves every gradient (field
    tables, decoder, proposal grid and MLP) unchanged up to FP32 summation
    order, and the loss with it."""
    case("proposal training pass, full batch: 4096 rays, ray order")
    m = nerf_model("bf16")
    m.enable_proposal(reso=64, n_fine=48, table_dtype="bf16")
    dev = torch.device("cuda")
    B = 4096
    g = torch.Generator(device=dev).manual_seed(18)
    rays = W.NerfScene(device=dev).random_rays(B, g)
    target = torch.rand((B, 3), generator=g, device=dev)
    prop_u = torch.rand((B, 64), generator=g, device=dev)
    fine_u = torch.rand((B, 48), generator=g, device=dev)
    perm = torch.randperm(B, generator=g, device=dev)
    params = list(m.parameters())
    assert any(p is m.proposal.grid for p in params)
    runs = []
    for order in (None, perm):
        def sel(t):
            return t if order is None else t[order].contiguous()
        m.zero_grad(set_to_none=True)
        loss = m.render_train_step_proposal(sel(rays), sel(target), white_bg=True, prop_u=sel(prop_u),
                                            fine_u=sel(fine_u))
        runs.append((float(loss), m.last_rgb_map().clone(),
                     [None if p.grad is None else p.grad.detach().clone() for p in params]))
    (l0, r0, g0), (l1, r1, g1) = runs
    assert torch.equal(r0[perm], r1)
    assert abs(l0 - l1) <= 1e-5 * abs(l0)
    n_grads = 0
    for k, (a, b) in enumerate(zip(g0, g1)):
        assert (a is None) == (b is None), f"param {k}"
        if a is not None:
            n_grads += 1
            assert_close_gpu(a, b, 1e-4, f"param {k}", atol_frac=1e-5)
    assert n_grads >= 6 + 1 + 3 + 4     # basis levels, coefficient, decoder, proposal grid + MLP


def test_image_full_batch():
    """2-D image regression (512x512 image, grid basis [4,4,4,2,2,2]@[32..128],
    64x64x18 coefficient grid, MLP 18-64-3, MSE) at its 65,536-pixel batch:
    subset invariance and the oracle for the forward; gradients independent
    of the batch order."""
    case("image full batch: 65536 pixels")
    m = FactorFields(W.image_cfg(512), "cuda")
    randomize(m, 19, scale=0.5)
    assert m.fused_supported()
    dev = torch.device("cuda")
    n = 65536
    g = torch.Generator(device=dev).manual_seed(20)
    coords = W.pixel_coords(512, 512, device=dev)
    x = coords[torch.randint(0, 512 * 512, (n,), generator=g, device=dev)].contiguous()
    t = torch.rand((n, 3), generator=g, device=dev)
    with torch.no_grad():
        out = m.field(x)
        idx = torch.randperm(n, generator=torch.Generator().manual_seed(21))[:4096].to(dev)
        sub = m.field(x[idx].contiguous())
    assert torch.equal(out[idx], sub)
    assert_close(sub, oracle_of(m).field(x[idx].cpu()), 1e-5, "image out")
    perm = torch.randperm(n, generator=g, device=dev)
    params = list(m.basises) + list(m.coeffs) + list(m.linear_mat.parameters())
    grads = []
    for xi, ti in [(x, t), (x[perm].contiguous(), t[perm].contiguous())]:
        m.zero_grad(set_to_none=True)
        m.regression_step(xi, ti, loss="mse")
        grads.append([p.grad.detach().clone() for p in params])
    for k, (a, b) in enumerate(zip(*grads)):
        assert float(b.abs().max()) > 0, f"param {k} has no gradient"
        assert_close_gpu(a, b, 1e-4, f"param {k}", atol_frac=1e-5)


def test_dense_alpha_full_grid():
    """getDenseAlpha (models/FactorFields.py:729-755) on the full 128^3 grid
    of the NeRF model with its 128^3 occupancy mask, one unjittered pass:
    4,096 grid points drawn at random match the oracle's alpha there."""
    from oracle import alpha_grid_sample
    case("dense alpha, 128^3 grid")
    m = nerf_model()
    R = [128, 128, 128]
    alpha, dense = m.getDenseAlpha(R, times=1)
    assert alpha.shape == (128, 128, 128)
    idx = torch.randperm(128 ** 3, generator=torch.Generator().manual_seed(22))[:4096]
    x = dense.reshape(-1, 3).cpu()[idx]
    of = oracle_of(m)
    sig = of.basis2density(of.field(x)[:, 0])
    am = m.alphaMask
    keep = alpha_grid_sample(am.alpha_volume.cpu()[0, 0], am.aabb.cpu(), x) > 0
    units = (m.aabb[1] - m.aabb[0]).cpu() / (torch.tensor(R) - 1)
    length = float(units.mean() * 25.0)
    ref = torch.where(keep, 1 - torch.exp(-sig * length), torch.zeros_like(sig))
    assert bool(keep.any()) and bool((~keep).any())
    assert_close(alpha.reshape(-1)[idx.to(alpha.device)], ref, 1e-5, "alpha")
