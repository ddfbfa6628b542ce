"""Pins of the oracle's camera rays (oracle/rays.py, restating
dataLoader/ray_utils.py:27-90 and dataLoader/blender.py:54-55): closed-form
directions of chosen pixels, exact axis permutations under 90-degree
rotations, unit norms after normalisation, origins equal to the camera
translation, and the batch gather against explicit loops."""
import math

import torch

from oracle import all_rays, gather_batch, get_ray_directions, get_ray_directions_blender, get_rays


def test_directions_closed_form():
    H, W, f = 4, 6, (2.0, 4.0)
    d = get_ray_directions(H, W, f)
    assert d.shape == (H, W, 3)
    for y in range(H):
        for x in range(W):
            want = torch.tensor([(x + 0.5 - W / 2) / f[0], (y + 0.5 - H / 2) / f[1], 1.0])
            assert torch.equal(d[y, x], want), (y, x)
    db = get_ray_directions_blender(H, W, f, center=[1.0, 3.0])
    for y in range(H):
        for x in range(W):
            want = torch.tensor([(x + 0.5 - 1.0) / f[0], -(y + 0.5 - 3.0) / f[1], -1.0])
            assert torch.equal(db[y, x], want), (y, x)


def test_rotation_permutes_axes_and_origin_is_translation():
    H, W = 3, 5
    d = get_ray_directions(H, W, (1.5, 1.5))
    # 90 degrees about z: (x, y, z) -> (-y, x, z), exact in FP32
    c2w = torch.tensor([[0.0, -1.0, 0.0, 0.25], [1.0, 0.0, 0.0, -2.0], [0.0, 0.0, 1.0, 3.5]])
    o, r = get_rays(d, c2w)
    flat = d.reshape(-1, 3)
    assert torch.equal(r, torch.stack([-flat[:, 1], flat[:, 0], flat[:, 2]], -1))
    assert torch.equal(o, torch.tensor([0.25, -2.0, 3.5]).expand(H * W, 3))


def test_normalised_rays_are_unit_and_concatenated_per_camera():
    H, W = 5, 7
    g = torch.Generator().manual_seed(0)
    c2ws = []
    for k in range(3):
        q, _ = torch.linalg.qr(torch.randn(3, 3, generator=g, dtype=torch.float64))
        c2ws.append(torch.cat([q, torch.randn(3, 1, generator=g, dtype=torch.float64)], 1).float())
    rays = all_rays(torch.stack(c2ws), H, W, (3.0, 3.0), blender=True)
    assert rays.shape == (3 * H * W, 6)
    assert (rays[:, 3:].double().norm(dim=-1) - 1).abs().max() < 1e-6
    for k in range(3):
        blk = rays[k * H * W:(k + 1) * H * W]
        assert torch.equal(blk[:, :3], c2ws[k][:3, 3].expand(H * W, 3))
    # the centre pixel of an odd image looks along the camera axis: (0, 0, -1) rotated
    c = rays[(H // 2) * W + W // 2]
    assert (c[3:] + c2ws[0][:3, 2]).abs().max() < 1e-6


def test_gather_batch_matches_loops():
    allrays = torch.arange(60, dtype=torch.float32).reshape(10, 6)
    allrgbs = torch.arange(30, dtype=torch.float32).reshape(10, 3)
    idx = torch.tensor([3, 3, 9, 0])
    r, c = gather_batch(allrays, allrgbs, idx)
    for b, j in enumerate(idx.tolist()):
        assert r[b].tolist() == [6 * j + k for k in range(6)]
        assert c[b].tolist() == [3 * j + k for k in range(3)]
