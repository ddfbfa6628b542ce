"""The reference's NeRF training driver, train_per_scene.py:89-226, run on the
library with the reference's own model configuration (configs/defaults.yaml
merged with configs/nerf.yaml: MLPRender_Fea colour module, 18 -> 64 -> 32
decoder, total_params 5,308,416) on the synthetic Gaussian scene.

The loop is the driver's, step for step: render_ray(is_train=True), MSE,
torch.optim.Adam (betas 0.9, 0.99) with the exponential lr decay,
updateAlphaMask / shrink / filtering_rays on update_AlphaMask_list and
shrinking_list, upsample_volume_grid with a fresh optimizer on upsamp_list,
and save / load at the end.  The schedule's iteration numbers are scaled down
(the reference's start at 500) so the loop runs in seconds."""
import os
import sys

import pytest
import torch

from ffields import FactorFields, N_to_reso, cal_n_samples
from ffields.config import AttrDict
from ffields.rays import SimpleSampler, gather_batch
from ffields.render import render_ray
from ffields.workloads import NerfScene

pytestmark = pytest.mark.gpu


def nerf_yaml_cfg(aabb=((-1.5, -1.5, -1.5), (1.5, 1.5, 1.5))):
    """configs/defaults.yaml (:1-60) merged with configs/nerf.yaml (:1-92), the
    keys this library reads, values as the two files give them (nerf.yaml
    wins where both set a key); dataset.aabb is what train_per_scene.py:114
    takes from the loader."""
    return AttrDict.wrap(dict(
        defaults=dict(expname="basis_lego", logdir="./logs", mode="reconstruction", progress_refresh_rate=10,
                      add_timestamp=0, ckpt=None),
        model=dict(coeff_reso=32, basis_dims=[4, 4, 4, 2, 2, 2], basis_resos=[32, 51, 70, 89, 108, 128],
                   freq_bands=[2.0, 3.2, 4.4, 5.6, 6.8, 8.0], coef_init=1.0, phases=[0.0], total_params=5308416,
                   T_basis=0, T_coeff=0, coef_mode="bilinear", basis_mode="bilinear", coeff_type="grid",
                   basis_type="grid", basis_mapping="sawtooth", with_dropout=False, in_dim=3, out_dim=32,
                   num_layers=2, hidden_dim=64),
        dataset=dict(dataset_name="blender", datadir="./data/nerf_synthetic/lego", ndc_ray=0, with_depth=0,
                     downsample_train=1.0, downsample_test=1.0, is_unbound=False, N_vis=5, vis_every=100000,
                     scene_reso=768, aabb=[list(aabb[0]), list(aabb[1])]),
        training=dict(batch_size=4096, n_iters=30000, lr_small=0.001, lr_large=0.02, lr_decay_iters=-1,
                      lr_decay_target_ratio=0.1, lr_upsample_reset=1, volume_resoInit=128, volume_resoFinal=300,
                      upsamp_list=[2000, 3000, 4000, 5500, 7000], update_AlphaMask_list=[2500, 4000],
                      shrinking_list=[500], L1_weight_inital=0.0, L1_weight_rest=0.0, Ortho_weight=0.0,
                      TV_weight_density=0.0, TV_weight_app=0.0, coeff=True, basis=True, linear_mat=True,
                      renderModule=True),
        exportation=dict(render_only=0, render_test=1, render_train=0, render_path=0, export_mesh=0,
                         export_mesh_only=0),
        renderer=dict(shadingMode="MLP_Fea", num_layers=3, hidden_dim=128, fea2denseAct="softplus",
                      density_shift=-10, distance_scale=25.0, view_pe=6, fea_pe=2, lindisp=0, perturb=1,
                      step_ratio=0.5, max_samples=1200, alphaMask_thres=0.02, rayMarch_weight_thres=1e-3)))


def test_train_per_scene_loop_with_nerf_yaml(tmp_path):
    device = torch.device("cuda")
    torch.manual_seed(20211202)
    scene = NerfScene(n_views=16, H=64, W=64, device="cuda")
    allrays = torch.cat([scene.rays(v, torch.arange(64 * 64, device=device)) for v in range(16)])
    allrgbs = scene.render_gt(allrays)
    cfg = nerf_yaml_cfg(scene.AABB)
    t = cfg.training
    # the driver's schedule, compressed: n_iters 30000 -> 500; density appears
    # after about 80 iterations on this scene
    t.n_iters = 500
    t.upsamp_list, t.update_AlphaMask_list, t.shrinking_list = [150, 200, 260, 330, 420], [200, 300], [100]
    alpha_mask_from = 150         # train_per_scene.py:198 uses iteration >= 1500
    print("[case] train_per_scene loop, nerf.yaml", file=sys.stderr, flush=True)

    model = FactorFields(cfg, device)
    assert model.basis_reso == [25, 39, 54, 69, 84, 99] and list(model.coeff_reso) == [32, 32, 32]
    grad_vars = model.get_optparam_groups(t.lr_small, t.lr_large)
    t.lr_decay_iters = t.n_iters
    lr_factor = t.lr_decay_target_ratio ** (1 / t.n_iters)
    optimizer = torch.optim.Adam(grad_vars, betas=(0.9, 0.99))
    volume_resoList = torch.linspace(t.volume_resoInit, t.volume_resoFinal,
                                     len(t.upsamp_list)).ceil().long().tolist()
    reso_cur = N_to_reso(t.volume_resoInit ** model.in_dim, model.aabb.cpu())
    nSamples = min(cfg.renderer.max_samples, cal_n_samples(reso_cur, cfg.renderer.step_ratio))
    assert nSamples == 443
    sampler = SimpleSampler(allrays.shape[0], t.batch_size, device=device)
    losses, events = [], []
    reso_mask = None
    for iteration in range(t.n_iters):
        ray_idx = sampler.nextids()
        rays_train, rgb_train = gather_batch(allrays, allrgbs, ray_idx)
        rgb_map, depth_map, coeffs = render_ray(rays_train, model, chunk=t.batch_size, N_samples=nSamples,
                                                white_bg=True, ndc_ray=0, device=device, is_train=True)
        assert coeffs.shape[1] == 18
        loss = torch.mean((rgb_map - rgb_train) ** 2)
        optimizer.zero_grad()
        loss.backward()
        optimizer.step()
        losses.append(loss.detach().item())
        for g in optimizer.param_groups:
            g["lr"] = g["lr"] * lr_factor
        if iteration in t.update_AlphaMask_list or iteration in t.shrinking_list:
            if volume_resoList[0] < 256:
                reso_mask = N_to_reso(volume_resoList[0] ** model.in_dim, model.aabb.cpu())
            new_aabb = model.updateAlphaMask(tuple(reso_mask), is_update_alphaMask=iteration >= alpha_mask_from)
            events.append(("alpha", iteration))
            if iteration in t.shrinking_list:
                old = (model.aabb[1] - model.aabb[0]).prod().item()
                model.shrink(new_aabb)
                assert (model.aabb[1] - model.aabb[0]).prod().item() <= old
                grad_vars = model.get_optparam_groups(t.lr_small, t.lr_large)
                optimizer = torch.optim.Adam(grad_vars, betas=(0.9, 0.99))
                events.append(("shrink", iteration))
            if iteration == t.update_AlphaMask_list[0]:
                n0 = allrays.shape[0]
                allrays, allrgbs = model.filtering_rays(allrays, allrgbs)
                assert 0 < allrays.shape[0] <= n0
                sampler = SimpleSampler(allrgbs.shape[0], t.batch_size, device=device)
                events.append(("filter", iteration))
        if iteration in t.upsamp_list:
            n_voxels = volume_resoList.pop(0)
            reso_cur = N_to_reso(n_voxels ** model.in_dim, model.aabb.cpu())
            nSamples = min(cfg.renderer.max_samples, cal_n_samples(reso_cur, cfg.renderer.step_ratio))
            model.upsample_volume_grid(reso_cur)
            assert model.gridSize.tolist() == reso_cur
            grad_vars = model.get_optparam_groups(t.lr_small, t.lr_large)
            optimizer = torch.optim.Adam(grad_vars, betas=(0.9, 0.99))
            events.append(("upsample", iteration))
    torch.cuda.synchronize()
    print(f"[loop] losses {losses[0]:.4f} -> {losses[-1]:.4f}, events {events}", file=sys.stderr, flush=True)
    assert model.alphaMask is not None
    assert nSamples > 443                               # upsampled grids sample more finely
    assert all(map(lambda v: v == v and v < 1.0, losses))
    # after the table reset of the shrink the loss falls again, well below its start
    assert sum(losses[-10:]) / 10 < 0.2 * sum(losses[101:111]) / 10
    # save and load (train_per_scene.py:115-120, :226)
    path = os.path.join(tmp_path, "basis_lego.th")
    model.save(path)
    ckpt = torch.load(path, map_location=device, weights_only=False)
    m2 = FactorFields(AttrDict.wrap(ckpt["cfg"]), device)
    m2.load(ckpt)
    test = allrays[:1024]
    with torch.no_grad():
        a, _ = render_ray(test, model, chunk=512, N_samples=nSamples, device=device)
        b, _ = render_ray(test, m2, chunk=512, N_samples=nSamples, device=device)
    assert torch.equal(a, b)
