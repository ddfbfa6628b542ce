"""Evaluation of a trained field on held-out rays (renderer.py:30-100):
render every test image through render_ray in chunks, clamp to [0, 1], and
report PSNR and SSIM per image.

SSIM is the Gaussian-window form the reference uses (utils.py:110-156,
rgb_ssim): an 11-tap Gaussian (sigma 1.5) applied separably over the 'valid'
window, the variances clipped at zero and the covariance clipped to
+-sqrt(var0 * var1), with the constants (k1 * max)^2 and (k2 * max)^2.  It is
computed here in float64 with torch, on the images' device.  LPIPS needs
pretrained network weights that cannot be fetched offline; it is not
computed.

evaluation_path renders a camera path (renderer.py:101-155) and writes the
frames as PNG files with PIL; the reference's MP4 videos need imageio, which
is not installed, so no video is written."""
import math
import os

import numpy as np
import torch
import torch.nn.functional as F

from .rays import get_rays
from .render import render_ray


def _gauss_taps(size, sigma, dtype, device):
    half = size // 2
    shift = (2 * half - size + 1) / 2
    x = (torch.arange(size, dtype=dtype, device=device) - half + shift) / sigma
    t = torch.exp(-0.5 * x * x)
    return t / t.sum()


def ssim(img0, img1, max_val=1.0, filter_size=11, filter_sigma=1.5, k1=0.01, k2=0.03, return_map=False):
    """SSIM of two [H, W, C] images (mean over the valid window and channels)."""
    if img0.dim() != 3 or img0.shape != img1.shape:
        raise ValueError("ssim expects two [H, W, C] images of the same shape")
    if img0.shape[0] < filter_size or img0.shape[1] < filter_size:
        raise ValueError(f"images must be at least {filter_size} pixels on each side")
    a = img0.to(torch.float64).permute(2, 0, 1)[:, None]      # [C, 1, H, W]
    b = img1.to(torch.float64).to(a.device).permute(2, 0, 1)[:, None]
    taps = _gauss_taps(filter_size, filter_sigma, torch.float64, a.device)
    ky, kx = taps.view(1, 1, -1, 1), taps.view(1, 1, 1, -1)

    def blur(z):  # separable, 'valid' window (the taps are symmetric)
        return F.conv2d(F.conv2d(z, ky), kx)

    mu0, mu1 = blur(a), blur(b)
    mu00, mu11, mu01 = mu0 * mu0, mu1 * mu1, mu0 * mu1
    s00 = torch.clamp(blur(a * a) - mu00, min=0.0)
    s11 = torch.clamp(blur(b * b) - mu11, min=0.0)
    s01 = blur(a * b) - mu01
    s01 = torch.sign(s01) * torch.minimum(torch.sqrt(s00 * s11), s01.abs())
    c1, c2 = (k1 * max_val) ** 2, (k2 * max_val) ** 2
    smap = ((2 * mu01 + c1) * (2 * s01 + c2)) / ((mu00 + mu11 + c1) * (s00 + s11 + c2))
    smap = smap[:, 0].permute(1, 2, 0)                          # [H', W', C]
    return smap if return_map else float(smap.mean())


def psnr(img, ref):
    """-10 log10(mean squared error) of two images in [0, 1]."""
    mse = torch.mean((img.to(torch.float64) - ref.to(torch.float64).to(img.device)) ** 2).item()
    return -10.0 * math.log10(mse) if mse > 0 else float("inf")


@torch.no_grad()
def evaluation(rays, rgbs, model, img_wh, chunk=1024, N_samples=-1, white_bg=False, ndc_ray=False,
               compute_extra_metrics=True, device="cuda"):
    """Render n test images and score them (renderer.py:30-100).

    rays [n, H*W, 6] or [n*H*W, 6], rgbs [n, H*W, 3] (or empty), img_wh = (W, H).
    Returns a dict with per-image ``psnr`` and ``ssim`` lists, their means, and
    the rendered ``rgb_maps`` [n, H, W, 3] and ``depth_maps`` [n, H, W] on the CPU."""
    W, H = img_wh
    rays = rays.reshape(-1, H * W, rays.shape[-1])
    has_gt = rgbs is not None and len(rgbs) > 0
    if has_gt:
        rgbs = rgbs.reshape(-1, H * W, 3)
    psnrs, ssims, rgb_maps, depth_maps = [], [], [], []
    for i in range(rays.shape[0]):
        rgb_map, depth_map = render_ray(rays[i], model, chunk=chunk, N_samples=N_samples, ndc_ray=ndc_ray,
                                        white_bg=white_bg, is_train=False, device=device)
        rgb_map = rgb_map.clamp(0.0, 1.0).reshape(H, W, 3)
        depth_map = depth_map.reshape(H, W)
        if has_gt:
            gt = rgbs[i].reshape(H, W, 3).to(rgb_map.device)
            psnrs.append(psnr(rgb_map, gt))
            if compute_extra_metrics:
                ssims.append(ssim(rgb_map, gt, 1.0))
        rgb_maps.append(rgb_map.cpu())
        depth_maps.append(depth_map.cpu())
    out = {"psnr": psnrs, "ssim": ssims, "rgb_maps": torch.stack(rgb_maps), "depth_maps": torch.stack(depth_maps)}
    if psnrs:
        out["psnr_mean"] = sum(psnrs) / len(psnrs)
    if ssims:
        out["ssim_mean"] = sum(ssims) / len(ssims)
    return out


def _jet_bgr():
    """256-entry jet colour table in BGR channel order (the order of OpenCV's
    colour maps), from matplotlib's piecewise-linear jet."""
    from matplotlib import colormaps
    rgb = colormaps["jet"](np.arange(256))[:, :3]
    return np.round(rgb * 255).astype(np.uint8)[:, ::-1].copy()


def visualize_depth_numpy(depth, minmax=None):
    """Colour-mapped depth (utils.py:11-26): NaN -> 0, normalised by
    (x - mi) / (ma - mi + 1e-8) with mi the smallest positive depth and ma the
    largest (or `minmax`), cast to 8 bits as numpy casts, then looked up in a
    jet table.  Returns ([H, W, 3] uint8 BGR, [mi, ma]).  The table is
    matplotlib's jet; OpenCV's, which the reference uses, differs from it by a
    few levels per channel."""
    x = np.nan_to_num(np.asarray(depth))
    if minmax is None:
        mi = np.min(x[x > 0])
        ma = np.max(x)
    else:
        mi, ma = minmax
    x = (x - mi) / (ma - mi + 1e-8)
    x = (255 * x).astype(np.uint8)
    return _jet_bgr()[x], [mi, ma]


@torch.no_grad()
def evaluation_path(test_dataset, factor_fields, c2ws, renderer=render_ray, savePath=None, prtx="", N_samples=-1,
                    white_bg=False, ndc_ray=False, compute_extra_metrics=True, device="cuda", chunk=8192,
                    return_maps=False):
    """Render the views of a camera path (renderer.py:101-155).

    test_dataset supplies img_wh = (W, H), directions [H, W, 3] and near_far;
    each c2w [3, 4] (or [4, 4]) gives the rays of one frame through get_rays.
    Frames are clamped to [0, 1] and converted to 8 bits; depth is colour
    mapped over near_far.  With savePath, frame k is written to
    savePath/{prtx}{k:03d}.png and the frame beside its depth to
    savePath/rgbd/{prtx}{k:03d}.png.  Returns the PSNR list, which is empty as
    in the reference (a path has no ground truth), or with return_maps a dict
    of that list and the uint8 ``rgb_maps`` [n, H, W, 3] and ``depth_maps``
    [n, H, W, 3]."""
    if ndc_ray:
        raise NotImplementedError("NDC rays are not supported")
    W, H = test_dataset.img_wh
    if savePath is not None:
        os.makedirs(savePath, exist_ok=True)
        os.makedirs(os.path.join(savePath, "rgbd"), exist_ok=True)
    near_far = test_dataset.near_far
    PSNRs, rgb_maps, depth_maps = [], [], []
    directions = torch.as_tensor(test_dataset.directions, dtype=torch.float32)
    for idx, c2w in enumerate(c2ws):
        c2w = torch.as_tensor(c2w, dtype=torch.float32, device=directions.device)
        rays_o, rays_d = get_rays(directions, c2w)
        rays = torch.cat([rays_o, rays_d], 1)
        rgb_map, depth_map = renderer(rays, factor_fields, chunk=chunk, N_samples=N_samples, ndc_ray=ndc_ray,
                                      white_bg=white_bg, device=device)
        rgb_map = rgb_map.clamp(0.0, 1.0)
        rgb_map, depth_map = rgb_map.reshape(H, W, 3).cpu(), depth_map.reshape(H, W).cpu()
        depth_img, _ = visualize_depth_numpy(depth_map.numpy(), near_far)
        rgb_img = (rgb_map.numpy() * 255).astype("uint8")
        rgb_maps.append(rgb_img)
        depth_maps.append(depth_img)
        if savePath is not None:
            from PIL import Image
            Image.fromarray(rgb_img).save(os.path.join(savePath, f"{prtx}{idx:03d}.png"))
            Image.fromarray(np.concatenate((rgb_img, depth_img), axis=1)).save(
                os.path.join(savePath, "rgbd", f"{prtx}{idx:03d}.png"))
    if return_maps:
        return {"psnr": PSNRs, "rgb_maps": np.stack(rgb_maps), "depth_maps": np.stack(depth_maps)}
    return PSNRs
