"""Write the golden fixtures under tests/golden/ from the CPU oracle.

Each fixture is an ``.npz`` (no pickled objects) holding seeded inputs, the
configuration as a JSON string, and the expected outputs:

* ``field_*``   -- get_coding / linear_mat outputs of small fields
                   (models/FactorFields.py:425-533, :113-159), computed in
                   FP64 from FP32-valued tables and points, and in FP32 with
                   the reference's operations (suffix 32);
* ``mapping``   -- grid_mapping for every mapping kind
                   (models/FactorFields.py:11-33), FP64;
* ``render``    -- the bounded-scene render (models/FactorFields.py:843-898)
                   with injected jitter and an occupancy volume, FP32 (the
                   sampler's masks and indices are exact-FP32 quantities);
* ``sample_pdf``-- dataLoader/ray_utils.py:132-174 with injected u, FP32;
* ``proposal``  -- the learned proposal's forward (proposal density, weights
                   and resampled fine t) on a 9^3 grid, FP32;
* ``datasets``  -- values of the synthetic workload generators at fixed
                   sample positions (image, SDF scene, NeRF scene), so that a
                   change to a generator is caught.

Run from the repository root:  python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (ROOT, os.path.join(ROOT, "factor-fields_amd"), os.path.dirname(HERE)):
    if p not in sys.path:
        sys.path.insert(0, p)

from oracle import OracleField, OracleRenderer, grid_mapping, sample_pdf  # noqa: E402
from oracle.proposal import proposal_forward  # noqa: E402

FIELD_CASES = {
    "field_2d_grid": dict(mode="image", D=2, ctype="grid"),
    "field_2d_triangle_nearest": dict(mode="image", D=2, ctype="grid", mapping="triangle", cmode="nearest",
                                      bmode="nearest"),
    "field_3d_grid": dict(mode="sdf", D=3, ctype="grid"),
    "field_3d_vm": dict(mode="sdf", D=3, ctype="vm"),
}

PROP_AABB = [[-1.2, -1.0, -1.4], [1.3, 1.1, 0.9]]


def small_cfg(mode="sdf", D=3, ctype="grid", cmode="bilinear", bmode="bilinear", mapping="sawtooth",
              out_dim=3):
    aabb = [[0.0, 0.0], [8.0, 6.0]] if D == 2 else [[-1.0, -0.5, -1.25], [1.0, 1.5, 0.75]]
    return dict(
        defaults=dict(mode=mode),
        model=dict(basis_dims=[4, 4, 4, 2, 2, 2], basis_resos=[9, 11, 13, 7, 6, 5],
                   freq_bands=[1.0, 1.7, 2.9, 3.3, 4.1, 5.0], coeff_reso=6, coeff_type=ctype,
                   basis_type="grid", coef_mode=cmode, basis_mode=bmode, basis_mapping=mapping,
                   param_derivation="literal", out_dim=out_dim, hidden_dim=64, num_layers=2, coef_init=1.0,
                   total_params=0, T_basis=0, T_coeff=0, with_dropout=False),
        dataset=dict(aabb=aabb, scene_reso=1024, is_unbound=False),
        renderer=dict(fea2denseAct="softplus", density_shift=-10, distance_scale=25.0),
        training=dict())


def render_cfg():
    return dict(
        defaults=dict(mode="reconstruction"),
        model=dict(basis_dims=[4, 4, 4, 2, 2, 2], basis_resos=[9, 11, 13, 7, 6, 5],
                   freq_bands=[1.0, 1.7, 2.9, 3.3, 4.1, 5.0], coeff_reso=8, coeff_type="grid",
                   basis_type="grid", coef_mode="bilinear", basis_mode="bilinear", basis_mapping="sawtooth",
                   param_derivation="literal", out_dim=28, hidden_dim=64, num_layers=2, coef_init=1.0,
                   total_params=0, T_basis=0, T_coeff=0, with_dropout=False),
        dataset=dict(aabb=[[-1.2, -1.0, -1.4], [1.3, 1.1, 0.9]], scene_reso=1024, is_unbound=False),
        renderer=dict(fea2denseAct="softplus", density_shift=-10, distance_scale=25.0, step_ratio=0.5,
                      max_samples=512, rayMarch_weight_thres=1e-3, alphaMask_thres=0.02, shadingMode="SH"),
        training=dict(volume_resoInit=40, volume_resoFinal=40))


def wrap(d):
    from oracle.field import AttrDict
    return AttrDict.wrap(json.loads(json.dumps(d)))


def seeded_field(cfg, seed, scale=1.0, dtype=torch.float64):
    """OracleField whose tables and weights are FP32 draws (stored as float32
    in the fixture) promoted to `dtype`."""
    of = OracleField(wrap(cfg), dtype=dtype)
    g = torch.Generator().manual_seed(seed)
    tabs = {}

    def draw(t):
        return (torch.randn(t.shape, generator=g) * scale).float()

    of.coeffs = [draw(t) for t in of.coeffs]
    of.basises = [draw(t) for t in of.basises]
    of.mlp = [(draw(w), None if b is None else draw(b)) for w, b in of.mlp]
    for k, t in enumerate(of.coeffs):
        tabs[f"coeff{k}"] = t.numpy()
    for k, t in enumerate(of.basises):
        tabs[f"basis{k}"] = t.numpy()
    for k, (w, b) in enumerate(of.mlp):
        tabs[f"w{k}"] = w.numpy()
        if b is not None:
            tabs[f"b{k}"] = b.numpy()
    of.coeffs = [t.to(dtype) for t in of.coeffs]
    of.basises = [t.to(dtype) for t in of.basises]
    of.mlp = [(w.to(dtype), None if b is None else b.to(dtype)) for w, b in of.mlp]
    return of, tabs


def field_points(of, n, seed, margin=0.1):
    g = torch.Generator().manual_seed(seed)
    a = of.aabb.float()
    ext = a[1] - a[0]
    return (a[0] - margin * ext + (1 + 2 * margin) * ext * torch.rand((n, len(a[0])), generator=g)).float()


def make_field(name, kw):
    cfg = small_cfg(**kw)
    of, tabs = seeded_field(cfg, seed=11)
    x = field_points(of, 256, seed=12)
    of32, _ = seeded_field(cfg, seed=11, dtype=torch.float32)
    with torch.no_grad():
        feats, coeff = of.get_coding(x.double())
        out = of.linear_mat(feats)
        feats32, coeff32 = of32.get_coding(x)
        out32 = of32.linear_mat(feats32)
    return dict(cfg=np.array(json.dumps(cfg)), x=x.numpy(), feats=feats.numpy(), coeff=coeff.numpy(),
                out=out.numpy(), feats32=feats32.numpy(), coeff32=coeff32.numpy(), out32=out32.numpy(), **tabs)


def make_mapping():
    g = torch.Generator().manual_seed(21)
    aabb = torch.tensor([[-1.0, -0.5, -1.25], [1.0, 1.5, 0.75]], dtype=torch.float64)
    freq = torch.tensor([1.0, 1.7, 2.9, 3.3, 4.1, 5.0], dtype=torch.float64)
    x = (aabb[0] - 0.2 + (aabb[1] - aabb[0] + 0.4) * torch.rand((128, 3), generator=g, dtype=torch.float64))
    x = x.float().double()
    out = dict(pos=x.numpy(), aabb=aabb.numpy(), freq=freq.numpy())
    for kind in ("sawtooth", "triangle", "sinc", "trigonometric", "x"):
        out[kind] = grid_mapping(x, freq, aabb, kind).numpy()
    return out


def make_rays(n, seed=0, radius=4.0):
    g = torch.Generator().manual_seed(seed)
    o = torch.randn((n, 3), generator=g)
    o = o / o.norm(dim=-1, keepdim=True) * radius
    tgt = (torch.rand((n, 3), generator=g) * 2 - 1) * 1.2
    d = tgt - o
    d = d / d.norm(dim=-1, keepdim=True)
    d[:3, 1] = 0.0                      # the d == 0 branch of sample_point
    d[:3] = d[:3] / d[:3].norm(dim=-1, keepdim=True)
    return torch.cat([o, d], -1)


def make_render():
    cfg = render_cfg()
    of, tabs = seeded_field(cfg, seed=41, scale=0.5, dtype=torch.float32)
    w0, b0 = of.mlp[0]
    of.mlp[0] = (w0 * 20.0, b0)
    tabs["w0"] = of.mlp[0][0].numpy()
    g = torch.Generator().manual_seed(42)
    a = of.aabb
    with torch.no_grad():
        o0 = of.field(a[0] + (a[1] - a[0]) * torch.rand((4096, 3), generator=g))[:, 0]
    shift = float(-torch.quantile(o0, 0.6))
    cfg["renderer"]["density_shift"] = shift
    of.cfg = wrap(cfg)
    vol = (torch.rand((29, 41, 37), generator=g) < 0.6).float()
    vol_aabb = torch.tensor([[-1.1, -0.9, -1.3], [1.2, 1.0, 0.8]])
    # render parameters (models/FactorFields.py:693-699) at the 40^3 initial grid
    from oracle import N_to_reso, cal_n_samples
    reso = N_to_reso(40 ** 3, a)
    units = (a[1] - a[0]) / (torch.tensor(reso, dtype=torch.float32) - 1)
    step = torch.mean(units) * 0.5
    n_samples = min(512, cal_n_samples(reso, 0.5))
    rays = make_rays(96, seed=43)
    jit = torch.rand(96, generator=g)
    rd = OracleRenderer(of, vol, vol_aabb, step, n_samples)
    with torch.no_grad():
        rgb, depth, aux = rd.forward(rays, jitter=jit, white_bg=True)
    valid = aux["ray_valid"]
    return dict(cfg=np.array(json.dumps(cfg)), rays=rays.numpy(), jitter=jit.numpy(), occ=vol.numpy(),
                occ_aabb=vol_aabb.numpy(), step=np.float32(step), n_samples=np.int32(n_samples),
                counts=valid.sum(-1).int().numpy(), idx=torch.nonzero(valid)[:, 1].int().numpy(),
                sigma=aux["sigma"][valid].numpy(), weight=aux["weight"][valid].numpy(),
                rgb_map=rgb.numpy(), depth=depth.numpy(), **tabs)


def make_sample_pdf():
    g = torch.Generator().manual_seed(51)
    B = 64
    s = torch.sort(torch.rand((B, 65), generator=g) * 4 + 1, dim=-1).values
    w = torch.rand((B, 64), generator=g) ** 4
    w[:4] = 0.0                                       # all-zero weights: uniform pdf after +1e-5
    u = torch.sort(torch.rand((B, 48), generator=g), dim=-1).values
    return dict(bins=s.numpy(), weights=w.numpy(), u=u.numpy(), samples=sample_pdf(s, w, 48, u=u).numpy(),
                det=sample_pdf(s, w, 48, det=True).numpy())


def make_proposal():
    g = torch.Generator().manual_seed(61)
    net = dict(grid=torch.randn((1, 8, 9, 9, 9), generator=g) * 2.0,
               w0=torch.randn((32, 8), generator=g) * 0.5, b0=torch.randn(32, generator=g) * 0.5,
               w1=torch.randn((1, 32), generator=g) * 0.5, aabb=torch.tensor(PROP_AABB), near=0.05,
               shift=-1.0, distance_scale=25.0, n_fine=48)
    rays = make_rays(64, seed=62)
    rays[:2, 3:6] = -rays[:2, 3:6]                    # two rays that miss the box
    prop_u = torch.rand((64, 64), generator=g)
    fine_u = torch.rand((64, 48), generator=g)
    ref = proposal_forward(net, rays, prop_u, fine_u)
    det = proposal_forward(net, rays)
    return dict(grid=net["grid"].numpy(), w0=net["w0"].numpy(), b0=net["b0"].numpy(), w1=net["w1"].numpy(),
                aabb=np.array(PROP_AABB, dtype=np.float32), rays=rays.numpy(), prop_u=prop_u.numpy(),
                fine_u=fine_u.numpy(), hit=ref["hit"].numpy(), bins=ref["bins"].numpy(),
                w_prop=ref["w_prop"].numpy(), t_fine=ref["t_fine"].numpy(), t_fine_det=det["t_fine"].numpy())


def make_datasets():
    from ffields import workloads as W
    img = W.synth_image(512)
    g = torch.Generator().manual_seed(71)
    px = torch.randint(0, 512, (64, 2), generator=g)
    scene = W.SdfScene()
    q = (torch.rand((64, 3), generator=g) * 2 - 1)
    nerf = W.NerfScene(device="cpu")
    rays = nerf.random_rays(16, torch.Generator().manual_seed(72))
    return dict(image_px=px.numpy(), image_val=img[px[:, 1], px[:, 0]].numpy(), sdf_q=q.numpy(),
                sdf_val=scene.sdf(q).numpy(), nerf_rays=rays.numpy(), nerf_rgb=nerf.render_gt(rays).numpy())


def main():
    torch.set_num_threads(1)
    out = {}
    for name, kw in FIELD_CASES.items():
        out[name] = make_field(name, kw)
    out["mapping"] = make_mapping()
    out["render"] = make_render()
    out["sample_pdf"] = make_sample_pdf()
    out["proposal"] = make_proposal()
    out["datasets"] = make_datasets()
    for name, arrs in out.items():
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, **arrs)
        print(f"{path}: {os.path.getsize(path)} bytes")


if __name__ == "__main__":
    main()
