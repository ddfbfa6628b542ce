"""Evaluation metrics (renderer.py:30-100, utils.py:110-156): SSIM against
its closed form on constant images and against a brute-force evaluation of
every window with the explicit 2-D Gaussian weights; PSNR by definition.
Depth colouring at levels fixed by its normalisation, and the frames and
files of evaluation_path (renderer.py:101-155) with a stand-in renderer."""
import math

import numpy as np

import torch

from ffields.evaluation import _gauss_taps, psnr, ssim


def test_taps():
    t = _gauss_taps(11, 1.5, torch.float64, "cpu")
    assert abs(float(t.sum()) - 1.0) < 1e-15
    assert torch.allclose(t, t.flip(0), rtol=0, atol=1e-18)
    assert int(t.argmax()) == 5
    assert abs(float(t[6] / t[5]) - math.exp(-0.5 / 1.5 ** 2)) < 1e-14


def test_ssim_constant_images_closed_form():
    """Constant images a, b have zero variances: SSIM = (2ab + c1) / (a^2 + b^2 + c1)."""
    c1 = (0.01 * 1.0) ** 2
    for a, b in [(0.3, 0.7), (0.0, 0.5), (0.9, 0.9)]:
        x = torch.full((16, 20, 3), a, dtype=torch.float64)
        y = torch.full((16, 20, 3), b, dtype=torch.float64)
        want = (2 * a * b + c1) / (a * a + b * b + c1)
        assert abs(ssim(x, y) - want) < 1e-12, (a, b)


def brute_ssim(x, y, size=11, sigma=1.5, k1=0.01, k2=0.03, max_val=1.0):
    t = _gauss_taps(size, sigma, torch.float64, "cpu")
    w2 = t[:, None] * t[None, :]
    H, W, C = x.shape
    c1, c2 = (k1 * max_val) ** 2, (k2 * max_val) ** 2
    vals = []
    for c in range(C):
        for i in range(H - size + 1):
            for j in range(W - size + 1):
                a = x[i:i + size, j:j + size, c]
                b = y[i:i + size, j:j + size, c]
                m0, m1 = (w2 * a).sum(), (w2 * b).sum()
                v0 = max(float((w2 * a * a).sum() - m0 * m0), 0.0)
                v1 = max(float((w2 * b * b).sum() - m1 * m1), 0.0)
                cv = float((w2 * a * b).sum() - m0 * m1)
                cv = math.copysign(min(math.sqrt(v0 * v1), abs(cv)), cv) if cv != 0 else 0.0
                m0, m1 = float(m0), float(m1)
                vals.append(((2 * m0 * m1 + c1) * (2 * cv + c2)) / ((m0 * m0 + m1 * m1 + c1) * (v0 + v1 + c2)))
    return sum(vals) / len(vals)


def test_ssim_matches_windowed_brute_force():
    g = torch.Generator().manual_seed(0)
    x = torch.rand((14, 13, 3), generator=g, dtype=torch.float64)
    y = (x + 0.2 * torch.randn((14, 13, 3), generator=g, dtype=torch.float64)).clamp(0, 1)
    assert abs(ssim(x, y) - brute_ssim(x, y)) < 1e-12
    smap = ssim(x, y, return_map=True)
    assert smap.shape == (4, 3, 3)


def test_ssim_identity_and_symmetry():
    g = torch.Generator().manual_seed(1)
    x = torch.rand((24, 30, 3), generator=g)
    y = torch.rand((24, 30, 3), generator=g)
    assert abs(ssim(x, x) - 1.0) < 1e-12
    assert abs(ssim(x, y) - ssim(y, x)) < 1e-12
    assert ssim(x, y) < 0.2


def test_psnr_definition():
    x = torch.zeros((4, 5, 3), dtype=torch.float64)
    y = torch.full((4, 5, 3), 0.1, dtype=torch.float64)
    assert abs(psnr(x, y) - 20.0) < 1e-9


def test_visualize_depth_known_levels():
    """Smallest positive depth -> level 0, the largest -> level 254 (255 * (1 -
    1e-8 / range) truncated), the midpoint -> 127; colours from the jet table
    in BGR order (level 0 is dark blue: only the first channel lit)."""
    from ffields.evaluation import _jet_bgr, visualize_depth_numpy
    jet = _jet_bgr()
    assert jet[0, 0] > 100 and jet[0, 1] == 0 and jet[0, 2] == 0
    assert jet[255, 2] > 100 and jet[255, 0] == 0 and jet[255, 1] == 0
    d = np.array([[2.0, 4.0], [3.0, 3.0]])
    img, mm = visualize_depth_numpy(d)
    assert mm == [2.0, 4.0] and img.shape == (2, 2, 3) and img.dtype == np.uint8
    assert (img[0, 0] == jet[0]).all() and (img[0, 1] == jet[254]).all() and (img[1, 0] == jet[127]).all()
    img2, mm2 = visualize_depth_numpy(d, [0.0, 8.0])
    assert mm2 == [0.0, 8.0] and (img2[0, 1] == jet[127]).all() and (img2[0, 0] == jet[63]).all()


class _PathData:
    def __init__(self, H=6, W=8):
        from ffields.rays import get_ray_directions
        self.img_wh = (W, H)
        self.directions = get_ray_directions(H, W, (10.0, 10.0))
        self.near_far = [2.0, 6.0]


def _stub_renderer(rays, model, chunk, N_samples, ndc_ray, white_bg, device):
    """Colour from the ray direction, depth from the origin's distance; checks
    that the chunk size arrives."""
    assert chunk == 8192 and not ndc_ray
    d = rays[:, 3:6]
    rgb = (d / d.norm(dim=1, keepdim=True) * 1.5 + 0.25)        # partly outside [0, 1]: clamped
    return rgb, rays[:, :3].norm(dim=1)


def test_evaluation_path_frames(tmp_path):
    from PIL import Image

    from ffields.evaluation import evaluation_path, visualize_depth_numpy
    from ffields.rays import get_rays
    data = _PathData()
    c2ws = [torch.eye(4)[:3], torch.tensor([[0.0, 0, 1, 3], [0, 1, 0, 0], [-1, 0, 0, 1]])]
    out = evaluation_path(data, None, c2ws, renderer=_stub_renderer, savePath=str(tmp_path), prtx="p_",
                          device="cpu", return_maps=True)
    assert out["psnr"] == [] and out["rgb_maps"].shape == (2, 6, 8, 3) and out["depth_maps"].shape == (2, 6, 8, 3)
    for k, c2w in enumerate(c2ws):
        o, d = get_rays(data.directions, c2w)
        rgb, depth = _stub_renderer(torch.cat([o, d], 1), None, 8192, -1, False, False, "cpu")
        want = (rgb.clamp(0, 1).reshape(6, 8, 3).numpy() * 255).astype(np.uint8)
        assert (out["rgb_maps"][k] == want).all()
        dimg, _ = visualize_depth_numpy(depth.reshape(6, 8).numpy(), data.near_far)
        assert (out["depth_maps"][k] == dimg).all()
        assert (np.asarray(Image.open(tmp_path / f"p_{k:03d}.png")) == want).all()
        rgbd = np.asarray(Image.open(tmp_path / "rgbd" / f"p_{k:03d}.png"))
        assert rgbd.shape == (6, 16, 3) and (rgbd[:, 8:] == dimg).all()
    assert evaluation_path(data, None, c2ws[:1], renderer=_stub_renderer, device="cpu") == []
    # poses as numpy arrays give the same frames
    out2 = evaluation_path(data, None, [c.numpy() for c in c2ws], renderer=_stub_renderer, device="cpu",
                           return_maps=True)
    assert (out2["rgb_maps"] == out["rgb_maps"]).all()
