"""Camera rays and batch sampling (dataLoader/ray_utils.py:27-90,
train_per_scene.py:16-28): pixel-centre directions by their closed form, a
rotation by c2w checked on known axes, and the sampler's epoch coverage."""
import math

import torch

from ffields.rays import SimpleSampler, get_ray_directions, get_ray_directions_blender, get_rays


def test_directions_closed_form():
    H, W, f = 3, 4, (2.0, 4.0)
    d = get_ray_directions(H, W, f)
    assert d.shape == (H, W, 3)
    for y in range(H):
        for x in range(W):
            want = ((x + 0.5 - W / 2) / f[0], (y + 0.5 - H / 2) / f[1], 1.0)
            assert torch.allclose(d[y, x], torch.tensor(want), rtol=0, atol=1e-7)
    b = get_ray_directions_blender(H, W, f, center=(1.0, 2.0))
    assert torch.allclose(b[2, 3], torch.tensor([(3.5 - 1.0) / 2.0, -(2.5 - 2.0) / 4.0, -1.0]), atol=1e-7)


def test_get_rays_rotation_and_origin():
    d = get_ray_directions(2, 2, (1.0, 1.0))
    c = math.cos(math.pi / 2)
    # camera rotated +90 degrees about z and placed at (1, 2, 3)
    c2w = torch.tensor([[c, -1.0, 0.0, 1.0], [1.0, c, 0.0, 2.0], [0.0, 0.0, 1.0, 3.0]])
    o, r = get_rays(d, c2w)
    assert o.shape == (4, 3) and r.shape == (4, 3)
    assert torch.equal(o, torch.tensor([[1.0, 2.0, 3.0]]).expand(4, 3))
    flat = d.reshape(-1, 3)
    want = torch.stack([-flat[:, 1] + c * flat[:, 0], flat[:, 0] + c * flat[:, 1], flat[:, 2]], -1)
    assert torch.allclose(r, want, atol=1e-6)


def test_simple_sampler_epochs():
    g = torch.Generator().manual_seed(0)
    s = SimpleSampler(10, 3, generator=g)
    seen = [s.nextids() for _ in range(3)]       # one epoch: 3 full batches of a permutation
    allid = torch.cat(seen)
    assert allid.numel() == 9 and allid.unique().numel() == 9
    nxt = s.nextids()                              # 12 > 10: a new permutation starts
    assert nxt.numel() == 3 and s.curr == 0
    assert torch.equal(torch.sort(s.ids).values, torch.arange(10))
