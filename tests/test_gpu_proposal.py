"""GPU parity of the learned proposal sampler (ff_proposal_sample /
ff_proposal_bwd) against the oracle restatement (oracle/proposal.py).

Bars: per-ray sample counts exact; fine sample t within 1e-5 of the ray's
[t_near, t_far] span, except for discrete flips: the CDF is a parallel prefix
sum on the GPU and a sequential cumsum on the CPU, and sample_pdf's
`denom < 1e-5 -> 1` rule makes the inverse CDF discontinuous at near-empty
bins, so a draw within rounding of a CDF step can land a bin away (at most 1%
of the samples); gradients within 1e-4 relative."""
import sys

import pytest
import torch

from helpers import assert_close
from ffields.proposal import ProposalNet
from ffields.render import alloc_samples, _samples_struct
from oracle.proposal import interlevel_loss, proposal_forward
from test_gpu_render import make_rays

pytestmark = pytest.mark.gpu

AABB = [[-1.2, -1.0, -1.4], [1.3, 1.1, 0.9]]


def make_prop(seed=0, dtype="fp32"):
    p = ProposalNet(AABB, reso=9, n_fine=48, density_shift=-1.0, table_dtype=dtype, device="cuda", seed=seed)
    g = torch.Generator().manual_seed(seed + 1)
    with torch.no_grad():
        p.grid.copy_(torch.randn(p.grid.shape, generator=g).cuda() * 2.0)
        for m in p.mlp:
            m.weight.copy_(torch.randn(m.weight.shape, generator=g).cuda() * 0.5)
            if m.bias is not None:
                m.bias.copy_(torch.randn(m.bias.shape, generator=g).cuda() * 0.5)
    if dtype == "bf16":
        p.refresh_shadow()
    return p


def oracle_net(p, bf16=False):
    grid = p.grid.detach().float().cpu().contiguous()
    if bf16:
        grid = grid.to(torch.bfloat16).float()
    return dict(grid=grid, w0=p.mlp[0].weight.detach().cpu(), b0=p.mlp[0].bias.detach().cpu(),
                w1=p.mlp[1].weight.detach().cpu(), aabb=torch.tensor(AABB), near=0.05,
                shift=p.density_shift, distance_scale=p.distance_scale, n_fine=p.n_fine)


def run_sample(p, rays, prop_u, fine_u):
    B = rays.shape[0]
    buf = alloc_samples(B, p.n_fine, "cuda")
    s = _samples_struct(buf)
    p.sample(rays.cuda(), s, None if prop_u is None else prop_u.cuda(), None if fine_u is None else fine_u.cuda())
    torch.cuda.synchronize()
    return buf, s


@pytest.mark.parametrize("train", [True, False])
@pytest.mark.parametrize("dtype", ["fp32", "bf16"])
def test_proposal_sample_matches_oracle(train, dtype):
    p = make_prop(dtype=dtype)
    B = 257
    rays = make_rays(B, seed=31)
    rays[:5, 3:6] = -rays[:5, 3:6]                     # a few misses
    g = torch.Generator().manual_seed(32)
    prop_u = torch.rand((B, 64), generator=g) if train else None
    fine_u = torch.rand((B, 48), generator=g) if train else None
    buf, _ = run_sample(p, rays, prop_u, fine_u)
    ref = proposal_forward(oracle_net(p, bf16=dtype == "bf16"), rays, prop_u, fine_u)
    off = buf["ray_off"].cpu()
    hit = ref["hit"]
    assert torch.equal(off[1:] - off[:-1], hit.int() * 48)
    M = int(off[-1])
    t_gpu = buf["z"][:M].cpu().view(-1, 48)
    t_ref = ref["t_fine"][hit]
    span = (ref["bins"][hit, -1] - ref["bins"][hit, 0])[:, None]
    bad = (t_gpu - t_ref).abs() > 1e-5 * span + 1e-6
    assert int(bad.sum()) <= max(1, int(0.01 * bad.numel())), int(bad.sum())
    assert torch.equal(buf["ray"][:M].cpu().view(-1, 48)[:, 0].long(), torch.nonzero(hit)[:, 0])
    assert torch.all(t_gpu[:, 1:] >= t_gpu[:, :-1])   # sorted along each ray
    xyz = buf["xyz"][:M].cpu().view(-1, 48, 3)
    r = rays[hit]
    assert_close(xyz, r[:, None, :3] + r[:, None, 3:6] * t_gpu[..., None], 1e-5, "xyz", atol_frac=1e-6)


def test_proposal_offsets_large_batch():
    """Above 65,536 rays the offsets come from the counting pass and the
    separate scan; they equal the one-launch offsets (pinned by the oracle
    test above) of the same rays taken in chunks."""
    print("[case] proposal offsets, 70,000 rays", file=sys.stderr, flush=True)
    p = make_prop(dtype="bf16")
    B = 70_000
    rays = make_rays(B, seed=33)
    rays[::7, 3:6] = -rays[::7, 3:6]                  # misses scattered through the batch
    big, _ = run_sample(p, rays, None, None)
    counts = []
    for lo in range(0, B, 40_000):
        part, _ = run_sample(p, rays[lo:lo + 40_000], None, None)
        off = part["ray_off"].cpu()
        counts.append(off[1:] - off[:-1])
    want = torch.cat([torch.zeros(1, dtype=torch.int32), torch.cumsum(torch.cat(counts), 0).int()])
    assert torch.equal(big["ray_off"].cpu(), want)
    assert int(want[-1]) > 0 and int(want[-1]) < 48 * B


def test_proposal_backward_matches_oracle():
    p = make_prop(seed=3)
    B = 96
    rays = make_rays(B, seed=41)
    g = torch.Generator().manual_seed(42)
    prop_u = torch.rand((B, 64), generator=g)
    fine_u = torch.rand((B, 48), generator=g)
    buf, s = run_sample(p, rays, prop_u, fine_u)
    M = int(buf["ray_off"][-1])
    wf = torch.rand(M, generator=g) * 0.2
    wf_cuda = torch.zeros(buf["xyz"].shape[0], device="cuda")
    wf_cuda[:M] = wf.cuda()
    loss_sum = torch.zeros((), device="cuda")
    p.backward(rays.cuda(), s, wf_cuda, prop_u=prop_u.cuda(), loss_scale=0.5, loss_sum=loss_sum)
    net = oracle_net(p)
    for k in ("grid", "w0", "b0", "w1"):
        net[k] = net[k].clone().requires_grad_(True)
    ref = proposal_forward(net, rays, prop_u, fine_u)
    hit = ref["hit"]
    t_fine = buf["z"][:M].cpu().view(-1, 48)
    loss = interlevel_loss(ref["bins"][hit], ref["w_prop"][hit], t_fine, wf.view(-1, 48)).sum() * 0.5 / B
    loss.backward()
    assert abs(loss_sum.item() - loss.item()) <= 1e-4 * abs(loss.item()) + 1e-8
    assert_close(p.grid.grad.contiguous(), net["grid"].grad, 1e-4, "grid", atol_frac=1e-5)
    assert_close(p.mlp[0].weight.grad, net["w0"].grad, 1e-4, "w0", atol_frac=1e-5)
    assert_close(p.mlp[0].bias.grad, net["b0"].grad, 1e-4, "b0", atol_frac=1e-5)
    assert_close(p.mlp[1].weight.grad, net["w1"].grad, 1e-4, "w1", atol_frac=1e-5)
