"""Oracle: learned proposal sampler (CPU torch restatement of the proposal
kernels' arithmetic).

TEST INFRASTRUCTURE ONLY -- see the package docstring.  The proposal network
has no counterpart in the reference's training loop; its resampling is the
reference's sample_pdf (dataLoader/ray_utils.py:132-174, restated in
oracle/render.py) and its compositing is raw2alpha
(models/FactorFields.py:82-88).  Pinned by tests/test_oracle_proposal.py
(brute-force loops for the interlevel bound loss and the bin bookkeeping).
"""

import torch
import torch.nn.functional as F

from .render import raw2alpha, sample_pdf


def ray_box(rays, aabb, near=0.05):
    o, d = rays[:, :3], rays[:, 3:6]
    vec = torch.where(d == 0, torch.full_like(d, 1e-6), d)
    ta = (aabb[1] - o) / vec
    tb = (aabb[0] - o) / vec
    tn = torch.minimum(ta, tb).amax(-1).clamp(min=near)
    tf = torch.maximum(ta, tb).amin(-1)
    return tn, tf, tf > tn


def proposal_density(grid, w0, b0, w1, aabb, x, shift):
    """grid [1, 8, D, H, W]; trilinear, align_corners=False, border; MLP
    8 -> 32 -> 1 (ReLU); softplus(out + shift)."""
    pts = ((x - aabb[0]) * (2.0 / (aabb[1] - aabb[0])) - 1).view(1, -1, 1, 1, 3)
    f = F.grid_sample(grid, pts, mode="bilinear", align_corners=False, padding_mode="border")
    f = f.view(8, -1).t()
    h = F.relu(F.linear(f, w0, b0))
    return F.softplus(F.linear(h, w1)[:, 0] + shift)


def proposal_forward(net, rays, prop_u=None, fine_u=None, n_prop=64):
    """Returns dict(bins [B,65], t_prop [B,64], w_prop [B,64], t_fine [B,nf], hit [B])."""
    tn, tf, hit = ray_box(rays, net["aabb"], net["near"])
    span = tf - tn
    j = torch.arange(n_prop, dtype=rays.dtype)
    s = tn[:, None] + span[:, None] * (j / n_prop)
    bins = torch.cat([s, tf[:, None]], -1)
    u = prop_u if prop_u is not None else torch.full_like(s, 0.5)
    t = s + u * (bins[:, 1:] - s)
    x = rays[:, None, :3] + rays[:, None, 3:6] * t[..., None]
    sig = proposal_density(net["grid"], net["w0"], net["b0"], net["w1"], net["aabb"], x.reshape(-1, 3),
                           net["shift"]).view(t.shape)
    delta = (bins[:, 1:] - bins[:, :-1]) * net["distance_scale"]
    _, w, _ = raw2alpha(sig, delta)
    nf = net["n_fine"]
    if fine_u is None:
        tf_s = sample_pdf(bins, w, nf, det=True)
    else:
        u = (torch.arange(nf, dtype=rays.dtype) + fine_u) / nf
        tf_s = sample_pdf(bins, w, nf, u=u)
    return dict(bins=bins, t_prop=t, w_prop=w, t_fine=tf_s, hit=hit, sigma=sig)


def interlevel_loss(bins, w_prop, t_fine, w_fine):
    """Bound loss: for each fine interval [t_k, t_k+1] (k < n_fine-1) the sum of
    the proposal weights of the bins it overlaps must cover its weight w_k:
    sum_k max(0, w_k - bound_k)^2 / (w_k + 1e-7), per ray."""
    B, nf = t_fine.shape
    out = []
    for r in range(B):
        tot = 0
        for k in range(nf - 1):
            t0, t1 = t_fine[r, k], t_fine[r, k + 1]
            # bins j with s_j < t1 and s_{j+1} > t0 (clamped to a valid range)
            j_lo = int((bins[r, 1:] <= t0).sum())
            j_hi = int((bins[r, :-1] < t1).sum()) - 1
            j_lo = min(j_lo, w_prop.shape[1] - 1)
            j_hi = max(0, min(j_hi, w_prop.shape[1] - 1))
            j_hi = max(j_hi, j_lo)
            bound = w_prop[r, j_lo:j_hi + 1].sum()
            gap = torch.clamp(w_fine[r, k] - bound, min=0)
            tot = tot + gap * gap / (w_fine[r, k] + 1e-7)
        out.append(tot)
    return torch.stack(out)
