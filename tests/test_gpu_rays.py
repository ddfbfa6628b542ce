"""ff_camera_rays and ff_gather_rays against the oracle (oracle/rays.py:
dataLoader/ray_utils.py:27-90, dataLoader/blender.py:54-55, :86-96,
train_per_scene.py:151-152)."""
import sys

import pytest
import torch

from ffields.rays import camera_rays, gather_batch
from helpers import assert_close
from oracle import all_rays

pytestmark = pytest.mark.gpu


def _cams(n, seed):
    g = torch.Generator().manual_seed(seed)
    out = []
    for _ in range(n):
        q, _ = torch.linalg.qr(torch.randn(3, 3, generator=g, dtype=torch.float64))
        out.append(torch.cat([q, 3 * torch.randn(3, 1, generator=g, dtype=torch.float64)], 1).float())
    return torch.stack(out)


@pytest.mark.parametrize("blender", [False, True])
@pytest.mark.parametrize("normalize", [False, True])
def test_camera_rays_match_oracle(blender, normalize):
    print(f"[case] camera rays blender={blender} normalize={normalize}", file=sys.stderr, flush=True)
    c2w = _cams(3, seed=1)
    H, W, f, cen = 11, 17, (13.5, 12.25), (8.0, 5.5)
    got = camera_rays(c2w.cuda(), H, W, f, center=cen, blender=blender, normalize=normalize).cpu()
    want = all_rays(c2w, H, W, f, center=cen, blender=blender, normalize=normalize)
    assert got.shape == want.shape == (3 * H * W, 6)
    assert torch.equal(got[:, :3], want[:, :3])                  # origins are copies
    assert_close(got[:, 3:], want[:, 3:], 1e-6, "directions", atol_frac=2e-7)
    # default principal point (W / 2, H / 2)
    got = camera_rays(c2w.cuda(), H, W, f, blender=blender, normalize=normalize).cpu()
    want = all_rays(c2w, H, W, f, blender=blender, normalize=normalize)
    assert_close(got[:, 3:], want[:, 3:], 1e-6, "directions (default centre)", atol_frac=2e-7)


def test_camera_rays_subset_and_invalid_pairs():
    print("[case] camera rays subset", file=sys.stderr, flush=True)
    c2w = _cams(4, seed=2).cuda()
    H, W = 9, 10
    full = camera_rays(c2w, H, W, 7.0, blender=True)
    g = torch.Generator(device="cuda").manual_seed(3)
    view = torch.randint(0, 4, (1000,), device="cuda", generator=g)
    pix = torch.randint(0, H * W, (1000,), device="cuda", generator=g)
    sub = camera_rays(c2w, H, W, 7.0, blender=True, view=view, pix=pix)
    assert torch.equal(sub, full[view * H * W + pix])            # same arithmetic per ray
    bad = camera_rays(c2w, H, W, 7.0, view=torch.tensor([4, -1, 0, 0], device="cuda"),
                      pix=torch.tensor([0, 0, H * W, -1], device="cuda"))
    assert torch.isnan(bad).all()
    assert camera_rays(c2w, H, W, 7.0, view=view[:0], pix=pix[:0]).shape == (0, 6)


def test_gather_batch_is_exact():
    print("[case] gather batch", file=sys.stderr, flush=True)
    N = 1 << 20                                                  # the training pool of the NeRF workload
    g = torch.Generator(device="cuda").manual_seed(4)
    allrays = torch.randn((N, 6), device="cuda", generator=g)
    allrgbs = torch.rand((N, 3), device="cuda", generator=g)
    ids = torch.randperm(N, device="cuda", generator=g)[:4096]
    ids[:7] = ids[7]                                             # repeated indices
    r, c = gather_batch(allrays, allrgbs, ids)
    assert torch.equal(r, allrays[ids]) and torch.equal(c, allrgbs[ids])
    r2, c2 = gather_batch(allrays, None, ids)
    assert c2 is None and torch.equal(r2, r)
    r3, c3 = gather_batch(allrays, allrgbs, ids[:0])
    assert r3.shape == (0, 6) and c3.shape == (0, 3)
    r4, c4 = gather_batch(allrays, allrgbs, torch.tensor([N, -1, 5], device="cuda"))
    assert torch.isnan(r4[:2]).all() and torch.isnan(c4[:2]).all()
    assert torch.equal(r4[2], allrays[5]) and torch.equal(c4[2], allrgbs[5])
