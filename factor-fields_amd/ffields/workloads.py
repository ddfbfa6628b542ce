"""The benchmark workloads: model configs and seeded synthetic data.

Each ``*_cfg`` returns a cfg tree with the reference's keys
(configs/defaults.yaml layout) describing one workload; the data generators
are seeded from the reference's base seed 20211202 (train_per_scene.py:240),
config k using 20211202 + 1000*k.
"""

import math

import torch

from .config import AttrDict

BASE_SEED = 20211202
FREQ = [2.0, 3.2, 4.4, 5.6, 6.8, 8.0]
RESOS = [32, 51, 70, 89, 108, 128]


def _cfg(mode, model, dataset, renderer=None, training=None):
    m = dict(basis_dims=[4, 4, 4, 2, 2, 2], basis_resos=list(RESOS), freq_bands=list(FREQ),
             coeff_reso=32, total_params=0, T_basis=0, T_coeff=0, coef_init=1.0,
             coef_mode="bilinear", basis_mode="bilinear", basis_mapping="sawtooth",
             coeff_type="grid", basis_type="grid", with_dropout=False, num_layers=2, hidden_dim=64,
             out_dim=3, param_derivation="literal")
    m.update(model)
    r = dict(fea2denseAct="softplus", density_shift=-10, distance_scale=25.0, step_ratio=0.5,
             max_samples=512, alphaMask_thres=0.02, rayMarch_weight_thres=1e-3, shadingMode="SH")
    r.update(renderer or {})
    t = dict(batch_size=4096, n_iters=2000, lr_small=0.001, lr_large=0.02, volume_resoInit=128,
             volume_resoFinal=128, linear_mat=True, coeff=True, basis=True, renderModule=True)
    t.update(training or {})
    d = dict(is_unbound=False, scene_reso=1024)
    d.update(dataset)
    return AttrDict.wrap(dict(defaults=dict(mode=mode), model=m, dataset=d, renderer=r, training=t))


def image_cfg(size=512):
    """2D image regression: grid basis [4,4,4,2,2,2]@[32..128], 64x64x18
    coefficient grid, sawtooth, MLP 18->64->3, Adam lr 0.02."""
    return _cfg("image", dict(coeff_reso=64, out_dim=3, coef_init=1.0),
                dict(aabb=[[0.0, 0.0], [float(size), float(size)]]),
                training=dict(batch_size=65536, n_iters=2000, lr_small=0.02, lr_large=0.02))


def sdf_cfg():
    """3D SDF regression: 3-D grid basis [4,4,4,2,2,2]@[32..128] over [-1,1]^3,
    VM coefficient (three 128^2 planes + three 128 lines, 18 ch), MLP 18->64->1,
    L1 loss."""
    return _cfg("sdf", dict(coeff_type="vm", coeff_reso=128, out_dim=1, coef_init=0.2),
                dict(aabb=[[-1.0, -1.0, -1.0], [1.0, 1.0, 1.0]]),
                training=dict(batch_size=1 << 20, n_iters=2000, lr_small=0.002, lr_large=0.02))


def nerf_cfg():
    """Per-scene NeRF: 3-D grid basis [4,4,4,2,2,2]@[32..128], 32^3x18
    coefficient grid over [-1.5,1.5]^3 (reference keys: total_params 10,747,166,
    coeff_reso 32, scene_reso 1024), degree-2 SH colour head (out_dim 1 + 27)."""
    return _cfg("reconstruction", dict(coeff_reso=32, total_params=10747166, out_dim=28,
                                       param_derivation="reference"),
                dict(aabb=[[-1.5, -1.5, -1.5], [1.5, 1.5, 1.5]], scene_reso=1024),
                training=dict(batch_size=4096, n_iters=30000, lr_small=0.001, lr_large=0.02))


# ---------------------------------------------------------------- 2D image
def synth_image(size=512, seed=BASE_SEED, n_gauss=64, n_sin=32, device="cpu", compute_device="cpu"):
    """[size, size, 3] float32 in [0,1]: 64 random 2D Gaussians (sigma U[4,64]
    px) + 32 random sinusoids (U[1,40] cycles/image), min-max normalised per
    channel (scaled to `size` from the 512 px definition).  The random draws
    come from a CPU generator; the evaluation (float64) runs on compute_device."""
    g = torch.Generator().manual_seed(seed)
    s = size / 512.0
    cd = torch.device(compute_device)
    yy, xx = torch.meshgrid(torch.arange(size, dtype=torch.float64, device=cd) + 0.5,
                            torch.arange(size, dtype=torch.float64, device=cd) + 0.5, indexing="ij")
    img = torch.zeros((size, size, 3), dtype=torch.float64, device=cd)
    for _ in range(n_gauss):
        c = (torch.rand(2, generator=g, dtype=torch.float64) * size).to(cd)
        sig = ((4 + torch.rand(1, generator=g, dtype=torch.float64) * 60) * s).to(cd)
        amp = torch.rand(3, generator=g, dtype=torch.float64).to(cd)
        r2 = (xx - c[0]) ** 2 + (yy - c[1]) ** 2
        img += torch.exp(-r2 / (2 * sig ** 2))[..., None] * amp
    for _ in range(n_sin):
        f = (1 + torch.rand(1, generator=g, dtype=torch.float64) * 39).to(cd)
        th = (torch.rand(1, generator=g, dtype=torch.float64) * 2 * math.pi).to(cd)
        ph = (torch.rand(1, generator=g, dtype=torch.float64) * 2 * math.pi).to(cd)
        amp = (torch.rand(3, generator=g, dtype=torch.float64) / math.sqrt(n_sin)).to(cd)
        u = (xx * torch.cos(th) + yy * torch.sin(th)) / size
        img += torch.sin(2 * math.pi * f * u + ph)[..., None] * amp
    lo = img.amin(dim=(0, 1), keepdim=True)
    hi = img.amax(dim=(0, 1), keepdim=True)
    img = (img - lo) / (hi - lo)
    return img.float().to(device)


def _image_params(seed, n_gauss=64, n_sin=32):
    """synth_image's random draws for one image, in its order (CPU generator)."""
    g = torch.Generator().manual_seed(seed)
    gs, ss = [], []
    for _ in range(n_gauss):
        c = torch.rand(2, generator=g, dtype=torch.float64)
        sig = torch.rand(1, generator=g, dtype=torch.float64)
        amp = torch.rand(3, generator=g, dtype=torch.float64)
        gs.append(torch.cat([c, sig, amp]))
    for _ in range(n_sin):
        f = torch.rand(1, generator=g, dtype=torch.float64)
        th = torch.rand(1, generator=g, dtype=torch.float64)
        ph = torch.rand(1, generator=g, dtype=torch.float64)
        amp = torch.rand(3, generator=g, dtype=torch.float64)
        ss.append(torch.cat([f, th, ph, amp]))
    return torch.stack(gs), torch.stack(ss)


def synth_image_set(n=1024, size=256, device="cuda", chunk=8):
    """The dictionary image set: n images [size, size, 3], image i equal to
    synth_image(size, seed=BASE_SEED + 4000 + i) (the same draws, evaluated in
    float64 on `device`, `chunk` images at a time).  Returns
    [n, size*size, 3] float32 (row-major pixels)."""
    dev = torch.device(device)
    sc = size / 512.0
    params = [_image_params(BASE_SEED + 4000 + i) for i in range(n)]
    G = torch.stack([p[0] for p in params]).to(dev)        # [n, 64, 6]: cx, cy, sig, amp rgb
    Sn = torch.stack([p[1] for p in params]).to(dev)       # [n, 32, 6]: f, th, ph, amp rgb
    yy, xx = torch.meshgrid(torch.arange(size, dtype=torch.float64, device=dev) + 0.5,
                            torch.arange(size, dtype=torch.float64, device=dev) + 0.5, indexing="ij")
    xx, yy = xx.reshape(-1), yy.reshape(-1)
    out = torch.empty((n, size * size, 3), device=dev)
    for i0 in range(0, n, chunk):
        g = G[i0:i0 + chunk]
        sn = Sn[i0:i0 + chunk]
        k = g.shape[0]
        img = torch.zeros((k, size * size, 3), dtype=torch.float64, device=dev)
        for j in range(g.shape[1]):       # same accumulation order as synth_image
            cx, cy = g[:, j, 0:1] * size, g[:, j, 1:2] * size
            sig = (4 + g[:, j, 2:3] * 60) * sc
            r2 = (xx[None] - cx) ** 2 + (yy[None] - cy) ** 2
            img += torch.exp(-r2 / (2 * sig ** 2))[..., None] * g[:, j, None, 3:6]
        for j in range(sn.shape[1]):
            f = 1 + sn[:, j, 0:1] * 39
            th = sn[:, j, 1:2] * 2 * math.pi
            ph = sn[:, j, 2:3] * 2 * math.pi
            amp = sn[:, j, 3:6] / math.sqrt(sn.shape[1])
            u = (xx[None] * torch.cos(th) + yy[None] * torch.sin(th)) / size
            img += torch.sin(2 * math.pi * f * u + ph)[..., None] * amp[:, None, :]
        lo = img.amin(dim=1, keepdim=True)
        hi = img.amax(dim=1, keepdim=True)
        out[i0:i0 + k] = ((img - lo) / (hi - lo)).float()
    return out


def cell_order(size, coeff_reso, device="cpu"):
    """Pixel indices (row-major, [size*size]) ordered by the coefficient cell
    whose four texels a pixel interpolates (cells of c = size // coeff_reso
    pixels, offset by c/2 as the texel centres of align_corners=False are),
    then by row and column inside the cell."""
    c = max(size // coeff_reso, 1)
    y, x = torch.meshgrid(torch.arange(size), torch.arange(size), indexing="ij")
    y, x = y.reshape(-1), x.reshape(-1)
    cy, cx = (y + c // 2) // c, (x + c // 2) // c
    key = ((cy * (size + 1) + cx) * size + y) * size + x
    return torch.argsort(key).to(device)


def image_set_batch(n, n_images, size, generator, device, order=None):
    """n (image, pixel) pairs drawn uniformly with replacement
    (dataLoader/image_set.py:55-56).  Returns (img, pix) int64 tensors.

    order=None: independent draws in draw order (randint, as the reference).
    order=cell_order(...): the same distribution, delivered sorted by image
    and coefficient cell: a sorted draw of n keys on [0, n_images * P)
    (`sorted_batch`, no sort runs) split into image and cell-ordered pixel."""
    P = size * size
    if order is None:
        img = torch.randint(0, n_images, (n,), device=device, generator=generator)
        pix = torch.randint(0, P, (n,), device=device, generator=generator)
        return img, pix
    key = sorted_batch(n_images * P, n, generator, device).long()
    return key // P, order[key % P]


def dict_cfg(n_images=1024, size=256):
    """Dictionary image set: shared 2-D basis [32,32,32,16,16,16]@[32..128]
    (144 channels, sawtooth), per-image 32x32x144 coefficient grids, decoder
    144 -> 64 -> 3, 2^22 pixels per step."""
    return _cfg("images", dict(basis_dims=[32, 32, 32, 16, 16, 16], coeff_reso=32, out_dim=3, coef_init=1.0),
                dict(aabb=[[0.0, 0.0, 0.0], [float(size), float(size), float(n_images)]]),
                training=dict(batch_size=1 << 22, n_iters=2000, lr_small=0.02, lr_large=0.02))


def pixel_coords(h, w, device="cpu"):
    """Pixel centres (x+0.5, y+0.5) in row-major order (dataLoader/image.py:42-43)."""
    y, x = torch.meshgrid(torch.arange(h, device=device), torch.arange(w, device=device), indexing="ij")
    return (torch.stack((x, y), -1).float() + 0.5).reshape(-1, 2)


# ---------------------------------------------------------------- 3D SDF
def morton_order(pts, aabb, bits=10):
    """Permutation that sorts points [N, 3] by the Morton (Z-order) key of
    their cell in a 2^bits-per-axis grid over the box aabb [2, 3].

    A batch drawn from a pool in this order with sorted indices
    (`sorted_batch`) visits the field's texels in spatial runs: the
    backward's run-merging scatter then adds one run per texel instead of
    one value per point.  For the 2^20-point SDF batch (4x pool, VM
    coefficient, six basis levels) this halves the atomic line requests
    (46.6 -> 23.0 per point, counted on the CPU through the scatter's run
    rules).  The losses are means and the gradients sums over the batch, so
    the order changes only the FP32 summation order."""
    import torch
    lo = torch.as_tensor(aabb[0], dtype=torch.float32, device=pts.device)
    hi = torch.as_tensor(aabb[1], dtype=torch.float32, device=pts.device)
    side = (1 << bits)
    q = ((pts.float() - lo) / (hi - lo) * side).floor().clamp(0, side - 1).long()
    key = torch.zeros(len(pts), dtype=torch.long, device=pts.device)
    for b in range(bits):
        for a in range(3):
            key |= ((q[:, a] >> b) & 1) << (3 * b + a)
    return torch.sort(key, stable=True).indices


def sorted_batch(pool_n, n, generator, device):
    """n indices drawn uniformly with replacement from [0, pool_n), in
    ascending order, so a batch of a Morton-ordered pool keeps the pool's
    order.  No sort: the sorted values of n independent U(0,1) draws have the
    distribution of the normalised partial sums S_k / S_{n+1} of n + 1
    independent Exp(1) spacings (the Renyi representation of uniform order
    statistics), and floor(u * pool_n) is monotone, so the indices are
    distributed as a sorted draw of n uniform integers.  The partial sums are
    FP64; the indices int32 (int64 for pools beyond 2^31 - 1)."""
    import torch
    e = torch.empty(n + 1, dtype=torch.float64, device=device).exponential_(generator=generator)
    c = torch.cumsum(e, 0)
    u = c[:n] * (pool_n / c[n])
    return u.floor_().clamp_(max=pool_n - 1).to(torch.int32 if pool_n <= 2 ** 31 - 1 else torch.int64)


class SdfScene:
    """Union of 32 spheres (r U[0.05,0.25]) and 16 axis-aligned boxes
    (half-extents U[0.05,0.3]) with centres U[-1,1]^3; exact SDF."""

    def __init__(self, seed=BASE_SEED + 1000, n_sph=32, n_box=16):
        g = torch.Generator().manual_seed(seed)
        d = torch.float64
        self.sc = torch.rand((n_sph, 3), generator=g, dtype=d) * 2 - 1
        self.sr = 0.05 + torch.rand(n_sph, generator=g, dtype=d) * 0.2
        self.bc = torch.rand((n_box, 3), generator=g, dtype=d) * 2 - 1
        self.bh = 0.05 + torch.rand((n_box, 3), generator=g, dtype=d) * 0.25
        self.g = g

    def sdf(self, p):
        p = p.to(torch.float64)
        s = (torch.cdist(p, self.sc.to(p.device)) - self.sr.to(p.device)).amin(-1)
        q = (p[:, None, :] - self.bc.to(p.device)).abs() - self.bh.to(p.device)
        b = q.clamp(min=0).norm(dim=-1) + q.amax(-1).clamp(max=0)
        return torch.minimum(s, b.amin(-1))

    def surface_points(self, n, chunk=1 << 18):
        """Area-weighted surface samples of the primitives kept where they lie
        on the union's surface (|SDF| < 1e-5), offset by N(0, 0.01)."""
        g = self.g
        area_s = 4 * math.pi * self.sr ** 2
        hx, hy, hz = self.bh[:, 0], self.bh[:, 1], self.bh[:, 2]
        area_b = 8 * (hx * hy + hy * hz + hx * hz)
        areas = torch.cat([area_s, area_b])
        out = []
        got = 0
        while got < n:
            k = torch.multinomial(areas, chunk, replacement=True, generator=g)
            pts = torch.empty((chunk, 3), dtype=torch.float64)
            is_s = k < len(self.sr)
            ks = k[is_s]
            v = torch.randn((len(ks), 3), generator=g, dtype=torch.float64)
            pts[is_s] = self.sc[ks] + self.sr[ks, None] * v / v.norm(dim=-1, keepdim=True)
            kb = k[~is_s] - len(self.sr)
            h = self.bh[kb]
            faces = torch.stack([h[:, 1] * h[:, 2], h[:, 0] * h[:, 2], h[:, 0] * h[:, 1]], -1)
            ax = torch.multinomial(faces, 1, generator=g)[:, 0]
            u = (torch.rand((len(kb), 3), generator=g, dtype=torch.float64) * 2 - 1) * h
            sign = torch.where(torch.rand(len(kb), generator=g) < 0.5, -1.0, 1.0).to(torch.float64)
            u[torch.arange(len(kb)), ax] = sign * h[torch.arange(len(kb)), ax]
            pts[~is_s] = self.bc[kb] + u
            keep = self.sdf(pts).abs() < 1e-5
            pts = pts[keep]
            out.append(pts)
            got += len(pts)
        pts = torch.cat(out)[:n]
        return pts + torch.randn(pts.shape, generator=g, dtype=torch.float64) * 0.01

    def training_pool(self, n):
        """n points (half uniform in [-1,1]^3, half near-surface) and their SDF
        (in generation order; `morton_order` gives the scatter-friendly one)."""
        g = self.g
        nu = n // 2
        uni = torch.rand((nu, 3), generator=g, dtype=torch.float64) * 2 - 1
        surf = self.surface_points(n - nu)
        pts = torch.cat([uni, surf])
        return pts.float(), self.sdf(pts).float()[:, None]


# ---------------------------------------------------------------- NeRF scene
class NerfScene:
    """Density = sum of 48 seeded anisotropic Gaussians (centres U[-1,1]^3,
    per-axis std U[0.05,0.4], random rotation, peak sigma U[10,100]) inside
    [-1.5,1.5]^3; RGB = 0.5 + 0.5*sin(2*pi*f_c.x + phi_c) per channel with
    |f_c| U[0.1,0.5] cycles/unit.  Cameras: `n_views` pinhole views of HxW,
    FoV 40 deg, on a radius-4 Fibonacci sphere looking at the origin.  Ground
    truth: 1,024 stratified (mid-point) samples per ray over the ray/box
    interval, white background."""

    AABB = ((-1.5, -1.5, -1.5), (1.5, 1.5, 1.5))

    def __init__(self, seed=BASE_SEED + 2000, n_gauss=48, n_views=100, H=800, W=800, fov_deg=40.0,
                 radius=4.0, device="cuda"):
        g = torch.Generator().manual_seed(seed)
        d = torch.float64
        self.device = torch.device(device)
        self.mu = torch.rand((n_gauss, 3), generator=g, dtype=d) * 2 - 1
        std = 0.05 + torch.rand((n_gauss, 3), generator=g, dtype=d) * 0.35
        q, r = torch.linalg.qr(torch.randn((n_gauss, 3, 3), generator=g, dtype=d))
        q = q * torch.sign(torch.diagonal(r, dim1=-2, dim2=-1))[:, None, :]
        self.prec = q @ torch.diag_embed(1.0 / std ** 2) @ q.transpose(-1, -2)
        self.peak = 10 + torch.rand(n_gauss, generator=g, dtype=d) * 90
        fdir = torch.randn((3, 3), generator=g, dtype=d)
        fdir = fdir / fdir.norm(dim=-1, keepdim=True)
        self.freq = fdir * (0.1 + torch.rand((3, 1), generator=g, dtype=d) * 0.4)
        self.phase = torch.rand(3, generator=g, dtype=d) * 2 * math.pi
        self.H, self.W = H, W
        self.focal = 0.5 * W / math.tan(0.5 * math.radians(fov_deg))
        self.c2w = self._cameras(n_views, radius)
        self._dev = {k: v.to(self.device, torch.float32) for k, v in
                     dict(mu=self.mu, prec=self.prec, peak=self.peak, freq=self.freq,
                          phase=self.phase).items()}

    @staticmethod
    def _cameras(n, radius):
        ga = math.pi * (3 - math.sqrt(5))
        out = []
        for i in range(n):
            z = 1 - 2 * (i + 0.5) / n
            rr = math.sqrt(max(0.0, 1 - z * z))
            p = torch.tensor([rr * math.cos(ga * i), rr * math.sin(ga * i), z], dtype=torch.float64) * radius
            fwd = -p / p.norm()
            up = torch.tensor([0.0, 0.0, 1.0], dtype=torch.float64)
            if abs(float(fwd @ up)) > 0.99:
                up = torch.tensor([0.0, 1.0, 0.0], dtype=torch.float64)
            right = torch.linalg.cross(fwd, up)
            right = right / right.norm()
            up2 = torch.linalg.cross(right, fwd)
            # camera looks along -z (OpenGL / Blender convention)
            out.append(torch.stack([right, up2, -fwd, p], -1))
        return torch.stack(out)                         # [n, 3, 4]

    def rays(self, view, pix):
        """Rays [len(pix), 6] of pixel indices `pix` (row-major) of one view
        (directions through pixel centres, normalised)."""
        c2w = self.c2w[view].to(self.device, torch.float32)
        j = (pix // self.W).float() + 0.5
        i = (pix % self.W).float() + 0.5
        dirs = torch.stack([(i - self.W / 2) / self.focal, -(j - self.H / 2) / self.focal,
                            -torch.ones_like(i)], -1)
        d = dirs @ c2w[:, :3].T
        d = d / d.norm(dim=-1, keepdim=True)
        o = c2w[:, 3].expand_as(d)
        return torch.cat([o, d], -1)

    def random_rays(self, n, generator):
        view = torch.randint(0, len(self.c2w), (n,), device=self.device, generator=generator)
        pix = torch.randint(0, self.H * self.W, (n,), device=self.device, generator=generator)
        out = torch.empty((n, 6), device=self.device)
        for v in range(len(self.c2w)):
            sel = view == v
            if bool(sel.any()):
                out[sel] = self.rays(v, pix[sel])
        return out

    def _quad_coeffs(self):
        """q_g(x) = (x-mu_g)^T P_g (x-mu_g) as coefficients of the quadratic
        features [x^2, y^2, z^2, xy, xz, yz, x, y, z, 1] (float64 -> float32)."""
        if getattr(self, "_A", None) is None:
            P, mu = self.prec, self.mu
            Pm = (P @ mu[:, :, None])[..., 0]
            A = torch.stack([P[:, 0, 0], P[:, 1, 1], P[:, 2, 2], 2 * P[:, 0, 1], 2 * P[:, 0, 2],
                             2 * P[:, 1, 2], -2 * Pm[:, 0], -2 * Pm[:, 1], -2 * Pm[:, 2],
                             (mu * Pm).sum(-1)], -1)
            self._A = A.to(self.device, torch.float32)
        return self._A

    def density(self, x):
        A = self._quad_coeffs()
        x0, x1, x2 = x[:, 0], x[:, 1], x[:, 2]
        F = torch.stack([x0 * x0, x1 * x1, x2 * x2, x0 * x1, x0 * x2, x1 * x2, x0, x1, x2,
                         torch.ones_like(x0)], -1)
        q = (F @ A.T).clamp(min=0)
        return torch.exp(-0.5 * q) @ self._dev["peak"]

    def color(self, x):
        p = self._dev
        return 0.5 + 0.5 * torch.sin(2 * math.pi * (x @ p["freq"].T) + p["phase"])

    @torch.no_grad()
    def render_gt(self, rays, n_samples=1024, chunk=4096, white_bg=True):
        out = []
        lo = torch.tensor(self.AABB[0], device=self.device)
        hi = torch.tensor(self.AABB[1], device=self.device)
        for k in range(0, rays.shape[0], chunk):
            r = rays[k:k + chunk]
            o, d = r[:, :3], r[:, 3:6]
            vec = torch.where(d == 0, torch.full_like(d, 1e-6), d)
            ta, tb = (hi - o) / vec, (lo - o) / vec
            t0 = torch.minimum(ta, tb).amax(-1).clamp(min=0.0)
            t1 = torch.maximum(ta, tb).amin(-1)
            hit = t1 > t0
            t1 = torch.where(hit, t1, t0)
            u = (torch.arange(n_samples, device=self.device).float() + 0.5) / n_samples
            t = t0[:, None] + (t1 - t0)[:, None] * u
            delta = ((t1 - t0) / n_samples)[:, None].expand_as(t)
            x = (o[:, None, :] + d[:, None, :] * t[..., None]).reshape(-1, 3)
            sig = self.density(x).view(t.shape)
            col = self.color(x).view(*t.shape, 3)
            alpha = 1 - torch.exp(-sig * delta)
            T = torch.cumprod(torch.cat([torch.ones_like(alpha[:, :1]), 1 - alpha], -1), -1)
            w = alpha * T[:, :-1]
            c = (w[..., None] * col).sum(1)
            if white_bg:
                c = c + T[:, -1:]
            out.append(c.clamp(0, 1))
        return torch.cat(out)
