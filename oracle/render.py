"""Oracle: ray sampling, occupancy masking, compositing, SH colour, sample_pdf.

TEST INFRASTRUCTURE ONLY -- see the package docstring.  CPU restatement of the
reference's render path (models/FactorFields.py:82-110, :586-602, :843-898;
models/sh.py:34-85; dataLoader/ray_utils.py:132-174) with the random draws
replaced by injected inputs (per-ray jitter ``u``; white background flag).
"""

import torch
import torch.nn.functional as F

# models/sh.py:4-12
SH_C0 = 0.28209479177387814
SH_C1 = 0.4886025119029199
SH_C2 = [1.0925484305920792, -1.0925484305920792, 0.31539156525252005,
         -1.0925484305920792, 0.5462742152960396]


def sample_point(rays_o, rays_d, aabb, step_size, n_samples, jitter=None):
    """models/FactorFields.py:586-602.  ``jitter`` [B] replaces the per-ray
    ``torch.rand_like(rng[:, [0]])`` draw (:595); None means is_train=False."""
    vec = torch.where(rays_d == 0, torch.full_like(rays_d, 1e-6), rays_d)
    rate_a = (aabb[1] - rays_o) / vec
    rate_b = (aabb[0] - rays_o) / vec
    t_min = torch.minimum(rate_a, rate_b).amax(-1).clamp(min=0.05, max=1e3)
    rng = torch.arange(n_samples)[None].float()
    if jitter is not None:
        rng = rng.repeat(rays_d.shape[-2], 1)
        rng += jitter.view(-1, 1)
    step = step_size * rng
    interpx = t_min[..., None] + step
    rays_pts = rays_o[..., None, :] + rays_d[..., None, :] * interpx[..., None]
    mask_outbbox = ((aabb[0] > rays_pts) | (rays_pts > aabb[1])).any(dim=-1)
    return rays_pts, interpx, ~mask_outbbox


def alpha_grid_sample(alpha_volume, aabb, xyz):
    """AlphaGridMask.sample_alpha, models/FactorFields.py:91-110.  alpha_volume is
    [Dz, Hy, Wx] (0/1 float), aabb [2, 3]."""
    aabbSize = aabb[1] - aabb[0]
    invgridSize = 1.0 / aabbSize * 2
    xyz = (xyz - aabb[0]) * invgridSize - 1
    vol = alpha_volume.view(1, 1, *alpha_volume.shape[-3:])
    return F.grid_sample(vol, xyz.view(1, -1, 1, 1, 3), align_corners=True).view(-1)


def raw2alpha(sigma, dist):
    """models/FactorFields.py:82-88."""
    alpha = 1.0 - torch.exp(-sigma * dist)
    T = torch.cumprod(torch.cat([torch.ones_like(alpha[..., :1]), 1.0 - alpha + 1e-10], -1), -1)
    weights = alpha * T[..., :-1]
    return alpha, weights, T[..., -1:]


def eval_sh(deg, sh, dirs):
    """models/sh.py:34-64 (degrees 0..2)."""
    assert 0 <= deg <= 2 and (deg + 1) ** 2 == sh.shape[-1]
    result = SH_C0 * sh[..., 0]
    if deg > 0:
        x, y, z = dirs[..., 0:1], dirs[..., 1:2], dirs[..., 2:3]
        result = (result - SH_C1 * y * sh[..., 1] + SH_C1 * z * sh[..., 2] - SH_C1 * x * sh[..., 3])
        if deg > 1:
            xx, yy, zz = x * x, y * y, z * z
            xy, yz, xz = x * y, y * z, x * z
            result = (result + SH_C2[0] * xy * sh[..., 4] + SH_C2[1] * yz * sh[..., 5]
                      + SH_C2[2] * (2.0 * zz - xx - yy) * sh[..., 6] + SH_C2[3] * xz * sh[..., 7]
                      + SH_C2[4] * (xx - yy) * sh[..., 8])
    return result


def sh_rgb(viewdirs, feat_sh):
    """Degree-2 SH colour head: feat_sh [N, 27] (channel-major 3 x 9) -> sigmoid of
    eval_sh.  (The reference's unused models/sh.py supplies the polynomials; the
    sigmoid matches MLPRender_Fea's output activation, models/FactorFields.py:202.)"""
    sh = feat_sh.reshape(-1, 3, 9)
    return torch.sigmoid(eval_sh(2, sh, viewdirs))


def positional_encoding(positions, freqs):
    """models/FactorFields.py:74-79: for each channel c and frequency 2^f the
    entry c*freqs + f of sin(...), then the same order of cos(...)."""
    freq_bands = (2 ** torch.arange(freqs).float()).to(positions.device)
    pts = (positions[..., None] * freq_bands).reshape(positions.shape[:-1] + (freqs * positions.shape[-1],))
    return torch.cat([torch.sin(pts), torch.cos(pts)], dim=-1)


def mlp_render_fea(viewdirs, features, layers, viewpe=6, feape=2):
    """MLPRender_Fea.forward, models/FactorFields.py:188-203.  ``layers`` is
    [(weight, bias or None), ...] of torch.nn.Linear layout; ReLU between
    layers, sigmoid at the end."""
    indata = [features, viewdirs]
    if feape > 0:
        indata += [positional_encoding(features, feape)]
    if viewpe > 0:
        indata += [positional_encoding(viewdirs, viewpe)]
    h = torch.cat(indata, dim=-1)
    for l, (w, b) in enumerate(layers):
        h = F.linear(h, w, b)
        if l != len(layers) - 1:
            h = F.relu(h)
    return torch.sigmoid(h)


def composite(sigma, rgb, z_vals, dists, distance_scale, white_bg):
    """models/FactorFields.py:878-896 given per-sample sigma/rgb."""
    alpha, weight, bg_weight = raw2alpha(sigma, dists * distance_scale)
    acc_map = torch.sum(weight, -1)
    rgb_map = torch.sum(weight[..., None] * rgb, -2)
    if white_bg:
        rgb_map = rgb_map + (1.0 - acc_map[..., None])
    rgb_map = rgb_map.clamp(0, 1)
    with torch.no_grad():
        depth_map = torch.sum(weight * z_vals, -1)
    return rgb_map, depth_map, weight, acc_map


class OracleRenderer:
    """FactorFields.forward (models/FactorFields.py:843-898) for the bounded,
    non-NDC case, over an OracleField.  The colour head is the reference's
    MLPRender_Fea (:162-203) when ``render_layers`` is given (its
    [(weight, bias), ...]), else the degree-2 SH head."""

    def __init__(self, field, alpha_volume=None, alpha_aabb=None, step_size=None, n_samples=None,
                 render_layers=None, viewpe=6, feape=2):
        self.field = field
        self.alpha_volume = alpha_volume
        self.alpha_aabb = alpha_aabb
        self.step_size = step_size
        self.n_samples = n_samples
        self.render_layers = render_layers
        self.viewpe, self.feape = viewpe, feape

    def forward(self, rays, jitter=None, white_bg=True, n_samples=-1):
        f = self.field
        cfg = f.cfg
        n_samples = n_samples if n_samples > 0 else self.n_samples
        viewdirs = rays[:, 3:6]
        xyz, z_vals, inner_mask = sample_point(rays[:, :3], viewdirs, f.aabb, self.step_size,
                                               n_samples, jitter)
        dists = torch.cat((z_vals[:, 1:] - z_vals[:, :-1], torch.zeros_like(z_vals[:, :1])), dim=-1)
        viewdirs = viewdirs.view(-1, 1, 3).expand(xyz.shape)
        ray_valid = inner_mask.clone()
        if self.alpha_volume is not None:
            alpha_inner_valid = alpha_grid_sample(self.alpha_volume, self.alpha_aabb,
                                                  xyz[inner_mask]) > 0.5
            ray_valid[inner_mask.clone()] = alpha_inner_valid
        sigma = torch.zeros(xyz.shape[:-1], dtype=xyz.dtype)
        rgb = torch.zeros((*xyz.shape[:2], 3), dtype=xyz.dtype)
        feat = None
        if ray_valid.any():
            feats, _ = f.get_coding(xyz[ray_valid])
            feat = f.linear_mat(feats)
            sigma = sigma.masked_scatter(ray_valid, f.basis2density(feat[..., 0]))
        alpha, weight, bg_weight = raw2alpha(sigma, dists * cfg.renderer.distance_scale)
        app_mask = weight > cfg.renderer.rayMarch_weight_thres
        ray_valid_new = torch.logical_and(ray_valid, app_mask)
        app_sel = ray_valid_new[ray_valid]
        if app_sel.any():
            if self.render_layers is not None:
                valid_rgbs = mlp_render_fea(viewdirs[ray_valid_new], feat[app_sel, 1:], self.render_layers,
                                            self.viewpe, self.feape)
            else:
                valid_rgbs = sh_rgb(viewdirs[ray_valid_new], feat[app_sel, 1:28])
            rgb = rgb.masked_scatter(ray_valid_new[..., None].expand_as(rgb), valid_rgbs)
        acc_map = torch.sum(weight, -1)
        rgb_map = torch.sum(weight[..., None] * rgb, -2)
        if white_bg:
            rgb_map = rgb_map + (1.0 - acc_map[..., None])
        rgb_map = rgb_map.clamp(0, 1)
        with torch.no_grad():
            depth_map = torch.sum(weight * z_vals, -1)
        aux = dict(xyz=xyz, z_vals=z_vals, inner_mask=inner_mask, ray_valid=ray_valid,
                   sigma=sigma, weight=weight, app_mask=ray_valid_new, acc=acc_map,
                   bg=bg_weight)
        return rgb_map, depth_map, aux


def sample_pdf(bins, weights, N_samples, det=False, u=None):
    """dataLoader/ray_utils.py:132-174 with ``u`` injectable (replaces torch.rand)."""
    weights = weights + 1e-5
    pdf = weights / torch.sum(weights, -1, keepdim=True)
    cdf = torch.cumsum(pdf, -1)
    cdf = torch.cat([torch.zeros_like(cdf[..., :1]), cdf], -1)
    if u is None:
        if det:
            u = torch.linspace(0.0, 1.0, steps=N_samples, dtype=cdf.dtype)
            u = u.expand(list(cdf.shape[:-1]) + [N_samples])
        else:
            raise ValueError("pass u for the stochastic case")
    u = u.contiguous()
    inds = torch.searchsorted(cdf.detach(), u, right=True)
    below = torch.max(torch.zeros_like(inds - 1), inds - 1)
    above = torch.min((cdf.shape[-1] - 1) * torch.ones_like(inds), inds)
    inds_g = torch.stack([below, above], -1)
    matched_shape = [inds_g.shape[0], inds_g.shape[1], cdf.shape[-1]]
    cdf_g = torch.gather(cdf.unsqueeze(1).expand(matched_shape), 2, inds_g)
    bins_g = torch.gather(bins.unsqueeze(1).expand(matched_shape), 2, inds_g)
    denom = cdf_g[..., 1] - cdf_g[..., 0]
    denom = torch.where(denom < 1e-5, torch.ones_like(denom), denom)
    t = (u - cdf_g[..., 0]) / denom
    return bins_g[..., 0] + t * (bins_g[..., 1] - bins_g[..., 0])


def remove_small_objects(mask, min_size):
    """skimage.morphology.remove_small_objects(mask, min_size, connectivity=1)
    as updateAlphaMask calls it (models/FactorFields.py:774-775): label the
    face-connected components (scipy.ndimage.label with the 6-neighbour
    structure) and unset the voxels of components with fewer than min_size
    voxels.  mask: numpy bool [D, H, W].  Also returns, for comparisons of
    the labelling itself, each voxel's component as the smallest flat index
    in it (-1 for unset voxels) and each voxel's component size (0 for unset
    voxels)."""
    import numpy as np
    from scipy import ndimage
    lab, n = ndimage.label(mask, structure=ndimage.generate_binary_structure(mask.ndim, 1))
    flat = lab.ravel()
    sizes = np.bincount(flat, minlength=n + 1)
    first = np.full(n + 1, np.iinfo(np.int64).max, dtype=np.int64)
    np.minimum.at(first, flat, np.arange(flat.size, dtype=np.int64))
    first[0] = -1
    small = sizes < min_size
    small[0] = False
    out = mask.copy()
    out[small[lab]] = False
    voxel_sizes = sizes[lab]
    voxel_sizes[lab == 0] = 0
    return out, first[lab], voxel_sizes

