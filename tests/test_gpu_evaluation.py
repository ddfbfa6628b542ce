"""evaluation() renders each test image through render_ray in chunks and
scores it (renderer.py:30-100): its images equal a direct render of the same
rays and the oracle renderer's images (models/FactorFields.py:843-898
restated in oracle/render.py) to 1e-5, and its PSNR and SSIM equal the
metrics applied to those images.  evaluation_path's frames equal direct
renders of the rays of its cameras."""
import sys

import pytest
import torch

from ffields.evaluation import evaluation, psnr, ssim
from helpers import assert_close
from test_gpu_render import make_model, make_rays, oracle_renderer

pytestmark = pytest.mark.gpu


def test_evaluation_matches_direct_render():
    print("[case] evaluation", file=sys.stderr, flush=True)
    m = make_model(seed=8)
    W, H, n = 16, 12, 2
    rays = make_rays(n * H * W, seed=60).reshape(n, H * W, 6)
    gt = torch.rand((n, H * W, 3), generator=torch.Generator().manual_seed(61))
    out = evaluation(rays, gt, m, (W, H), chunk=100, white_bg=True)
    assert out["rgb_maps"].shape == (n, H, W, 3) and out["depth_maps"].shape == (n, H, W)
    orr = oracle_renderer(m)
    for i in range(n):
        with torch.no_grad():
            rgb, depth, _ = m(rays[i].cuda(), white_bg=True, is_train=False)
        want = rgb.clamp(0, 1).reshape(H, W, 3).cpu()
        assert torch.equal(out["rgb_maps"][i], want)
        assert torch.equal(out["depth_maps"][i], depth.reshape(H, W).cpu())
        # the oracle renderer on the same rays (no jitter in evaluation)
        ref_rgb, ref_depth, aux = orr.forward(rays[i], white_bg=True)
        assert float(aux["acc"].max()) > 0.5 and float(aux["acc"].min()) < 0.5
        assert_close(out["rgb_maps"][i].reshape(-1, 3), ref_rgb.clamp(0, 1), 1e-5, "evaluation rgb vs oracle")
        assert_close(out["depth_maps"][i].reshape(-1), ref_depth, 1e-5, "evaluation depth vs oracle")
        # float64 means on the GPU and the CPU differ in the last digits
        assert abs(out["psnr"][i] - psnr(want, gt[i].reshape(H, W, 3))) < 1e-12
        assert abs(out["psnr"][i] - psnr(ref_rgb.clamp(0, 1).reshape(H, W, 3), gt[i].reshape(H, W, 3))) < 1e-4
        # float64 convolutions on the GPU and the CPU differ in the last digits
        assert abs(out["ssim"][i] - ssim(want, gt[i].reshape(H, W, 3))) < 1e-12
    assert abs(out["psnr_mean"] - sum(out["psnr"]) / n) < 1e-12


def test_evaluation_path_matches_direct_render():
    print("[case] evaluation_path", file=sys.stderr, flush=True)
    import numpy as np

    from ffields.evaluation import evaluation_path
    from ffields.rays import get_ray_directions, get_rays

    class Path:
        img_wh = (10, 7)
        directions = get_ray_directions(7, 10, (8.0, 8.0))
        near_far = [2.0, 6.0]

    m = make_model(seed=8)
    # cameras at distance 4 on the -y and +x axes, looking at the origin
    c2ws = [torch.tensor([[1.0, 0, 0, 0], [0, 0, 1, -4], [0, -1, 0, 0]]),
            torch.tensor([[0.0, 0, -1, 4], [0, 1, 0, 0], [1, 0, 0, 0]])]
    out = evaluation_path(Path, m, c2ws, chunk=33, white_bg=True, return_maps=True)
    for k, c2w in enumerate(c2ws):
        o, d = get_rays(Path.directions, c2w)
        with torch.no_grad():
            rgb, _, _ = m(torch.cat([o, d], 1).cuda(), white_bg=True, is_train=False)
        want = (rgb.clamp(0, 1).reshape(7, 10, 3).cpu().numpy() * 255).astype(np.uint8)
        assert (out["rgb_maps"][k] == want).all()
