"""What the trained proposal network gains over sampling without it.

A field is trained on the synthetic Gaussian scene with the learned proposal
placing 48 fine samples per ray (the proposal trained online on the field's
own rendering weights).  The same field is then rendered on held-out rays with
48 samples per ray placed (a) by the trained proposal and (b) by an untrained
proposal whose density is nearly constant, i.e. sample_pdf over uniform weights
= stratified uniform placement along the ray/box interval.  The trained
proposal must give a clearly better render (PSNR gain >= 1 dB) at the same
sample budget, and it must keep the PSNR within 0.5 dB of the occupancy
sampler's render of the same field with its full sample count."""
import math
import sys

import pytest
import torch

from ffields import FactorFields, FusedAdam
from ffields import workloads as W
from ffields.params import N_to_reso, cal_n_samples
from ffields.proposal import ProposalNet

pytestmark = pytest.mark.gpu


def psnr(a, b):
    return -10 * math.log10(torch.mean((a - b) ** 2).item())


def test_trained_proposal_beats_uniform_placement():
    print("[case] proposal gain", file=sys.stderr, flush=True)
    dev = torch.device("cuda")
    scene = W.NerfScene(device=dev)
    g = torch.Generator(device=dev).manual_seed(5)
    pool = scene.random_rays(1 << 18, g)
    pool_rgb = scene.render_gt(pool)
    test = scene.random_rays(4096, torch.Generator(device=dev).manual_seed(6))
    test_rgb = scene.render_gt(test)
    cfg = W.nerf_cfg()
    torch.manual_seed(0)
    m = FactorFields(cfg, dev, table_dtype="bf16")
    reso = N_to_reso(cfg.training.volume_resoInit ** 3, m.aabb.cpu())
    ns = min(int(cfg.renderer.max_samples), cal_n_samples(reso, cfg.renderer.step_ratio))
    m.enable_proposal(reso=64, n_fine=48, table_dtype="bf16")
    opt = FusedAdam(m.get_optparam_groups(cfg.training.lr_small, cfg.training.lr_large), betas=(0.9, 0.99),
                    shadow_of=m.shadow_lookup)
    for it in range(1000):
        idx = torch.randint(0, pool.shape[0], (4096,), device=dev, generator=g)
        m.render_train_step_proposal(pool[idx], pool_rgb[idx], white_bg=True, reduce=False)
        opt.step(zero_grad=True)
        if it == 500:
            m.updateAlphaMask(tuple(reso), is_update_alphaMask=True)
    trained = m.render_eval(test, use_proposal=True)
    fresh = ProposalNet(m.aabb.cpu(), reso=64, n_fine=48, distance_scale=float(cfg.renderer.distance_scale),
                        table_dtype="bf16", device=dev, seed=1)
    kept, m.proposal = m.proposal, fresh
    uniform = m.render_eval(test, use_proposal=True)
    m.proposal = kept
    occupancy = m.render_eval(test, use_proposal=False, N_samples=ns)
    p_tr, p_un, p_oc = psnr(trained, test_rgb), psnr(uniform, test_rgb), psnr(occupancy, test_rgb)
    print(f"[proposal gain] trained proposal {p_tr:.2f} dB, uniform placement {p_un:.2f} dB, "
          f"occupancy sampler ({ns} max samples) {p_oc:.2f} dB", file=sys.stderr, flush=True)
    assert p_tr >= p_un + 1.0, (p_tr, p_un)
    assert p_tr >= p_oc - 0.5, (p_tr, p_oc)


def test_referee_matches_same_stream_renders():
    """ProposalReferee's side-stream occupancy render gives the same squared
    error as rendering on the current stream, the proposal's error is taken
    from the render it is given (here an evaluation render), and the
    fallback switch trips only above the bar."""
    from ffields.proposal import ProposalReferee
    from test_gpu_render import make_model, make_rays
    print("[case] proposal referee", file=sys.stderr, flush=True)
    m = make_model(seed=3)
    m.enable_proposal(reso=16, n_fine=48, table_dtype="fp32")
    rays = make_rays(300, seed=40).cuda()
    tgt = torch.rand((300, 3), generator=torch.Generator().manual_seed(41)).cuda()
    ref = ProposalReferee(rays.device, n_rays=200, max_gap_db=0.1)
    prop_all = m.render_eval(rays, use_proposal=True)
    ref.add_proposal(prop_all, tgt)            # no compare yet: not counted
    assert ref.sq_err.abs().sum().item() == 0
    ref.compare(m, rays, tgt)
    ref.add_proposal(prop_all, tgt)
    ref.join()
    got = ref.sq_err.clone()
    occ = m.render_eval(rays[:200], use_proposal=False)
    want = torch.stack([((occ.double() - tgt[:200].double()) ** 2).sum(),
                        ((prop_all[:200].double() - tgt[:200].double()) ** 2).sum()]).float()
    # FP32 sums of 600 squares (ff_sq_err_sum's block order) against FP64 sums
    assert torch.allclose(got, want, rtol=1e-5, atol=0), (got, want)
    gap = ref.decide()
    assert abs(gap - 10 * math.log10(want[1].item() / want[0].item())) < 1e-4
    assert ref.active == (gap <= 0.1)
    assert ref.sq_err.abs().sum().item() == 0


def test_last_rgb_map_is_the_training_render():
    """last_rgb_map() after a fused proposal training pass equals the same
    pass's render of the batch (forward only, same draws)."""
    from test_gpu_render import make_model, make_rays
    print("[case] last rgb map", file=sys.stderr, flush=True)
    m = make_model(seed=6)
    m.enable_proposal(reso=16, n_fine=48, table_dtype="fp32")
    B = 256
    rays = make_rays(B, seed=50).cuda()
    tgt = torch.rand((B, 3), generator=torch.Generator().manual_seed(51)).cuda()
    g = torch.Generator().manual_seed(52)
    pu, fu = torch.rand((B, 64), generator=g).cuda(), torch.rand((B, 48), generator=g).cuda()
    m.render_train_step_proposal(rays, tgt, white_bg=True, prop_u=pu, fine_u=fu)
    got = m.last_rgb_map().clone()
    ls = torch.zeros((), device="cuda")
    m.render_train_step_proposal(rays, tgt, white_bg=True, prop_u=pu, fine_u=fu, loss_sum=ls)
    assert torch.equal(got, m.last_rgb_map())
    assert abs(((got - tgt) ** 2).sum().item() - ls.item()) <= 1e-5 * ls.item()
