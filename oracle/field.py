"""Oracle: field evaluation (mapping, basis/coefficient gathers, decoder).

TEST INFRASTRUCTURE ONLY -- see the package docstring.  CPU restatement of
``models/FactorFields.py`` of the reference, written with the same ATen ops in
the same order so FP32 results match the reference's semantics.  Parameters are
held in the reference layout (channel-major ``[1, C, (D,) H, W]``).
"""

import math

import numpy as np
import torch
import torch.nn.functional as F


class AttrDict(dict):
    """Nested dict with attribute access (stands in for an OmegaConf node)."""

    def __getattr__(self, k):
        try:
            return self[k]
        except KeyError as e:
            raise AttributeError(k) from e

    def __setattr__(self, k, v):
        self[k] = v

    @staticmethod
    def wrap(d):
        if isinstance(d, dict):
            return AttrDict({k: AttrDict.wrap(v) for k, v in d.items()})
        if isinstance(d, list):
            return [AttrDict.wrap(v) for v in d]
        return d


def make_cfg(mode, model, dataset, renderer=None, training=None):
    """Build a cfg tree with the reference's defaults (configs/defaults.yaml:12-33,
    configs/nerf.yaml:73-92) overridden by the given sections."""
    m = dict(basis_dims=[4, 4, 4, 2, 2, 2], basis_resos=[32, 51, 70, 89, 108, 128],
             coeff_reso=32, total_params=10744166, T_basis=0, T_coeff=0, coef_init=1.0,
             coef_mode="bilinear", basis_mode="bilinear",
             freq_bands=[1.0, 1.33, 1.7689, 2.3526, 3.129, 4.1616], basis_mapping="sawtooth",
             with_dropout=False, in_dim=3, out_dim=32, num_layers=2, hidden_dim=128,
             coeff_type="grid", basis_type="grid", param_derivation="reference")
    m.update(model)
    r = dict(fea2denseAct="softplus", density_shift=-10, distance_scale=25.0, step_ratio=0.5,
             max_samples=1200, alphaMask_thres=0.02, rayMarch_weight_thres=1e-3)
    r.update(renderer or {})
    t = dict(batch_size=4096, n_iters=30000, lr_small=0.001, lr_large=0.02,
             volume_resoInit=128, volume_resoFinal=300)
    t.update(training or {})
    d = dict(is_unbound=False, scene_reso=768)
    d.update(dataset)
    return AttrDict.wrap(dict(defaults=dict(mode=mode), model=m, dataset=d, renderer=r, training=t))


# ---------------------------------------------------------------------------
# utils.py:53-70
# ---------------------------------------------------------------------------

def N_to_reso(n_voxels, bbox):
    """utils.py:53-57."""
    xyz_min, xyz_max = bbox
    dim = len(xyz_min)
    voxel_size = ((xyz_max - xyz_min).prod() / n_voxels).pow(1 / dim)
    return torch.round((xyz_max - xyz_min) / voxel_size).long().tolist()


def N_to_vm_reso(n_voxels, bbox):
    """utils.py:59-67."""
    xyz_min, xyz_max = bbox
    dim = len(xyz_min)
    voxel_size = ((xyz_max - xyz_min).prod() / n_voxels).pow(1 / dim)
    reso = (xyz_max - xyz_min) / voxel_size
    n_mat = reso[0] * reso[1] + reso[0] * reso[2] + reso[1] * reso[2]
    scale = math.sqrt(n_voxels / n_mat)
    return torch.round(reso * scale).long().tolist()


def cal_n_samples(reso, step_ratio=0.5):
    """utils.py:69-70."""
    return int(np.linalg.norm(reso) / step_ratio)


# ---------------------------------------------------------------------------
# models/FactorFields.py:11-33  grid_mapping
# ---------------------------------------------------------------------------

def grid_mapping(positions, freq_bands, aabb, basis_mapping="sawtooth"):
    """models/FactorFields.py:11-33.  positions [N, D] -> [N, D, L] ([N, D, 2L] for
    'trigonometric')."""
    aabbSize = max(aabb[1] - aabb[0])
    scale = aabbSize[..., None] / freq_bands
    if basis_mapping == "triangle":
        pts_local = (positions - aabb[0]).unsqueeze(-1) % scale
        pts_local_int = ((positions - aabb[0]).unsqueeze(-1) // scale) % 2
        pts_local = pts_local / (scale / 2) - 1
        pts_local = torch.where(pts_local_int == 1, -pts_local, pts_local)
    elif basis_mapping == "sawtooth":
        pts_local = (positions - aabb[0])[..., None] % scale
        pts_local = pts_local / (scale / 2) - 1
        pts_local = pts_local.clamp(-1.0, 1.0)
    elif basis_mapping == "sinc":
        pts_local = torch.sin((positions - aabb[0])[..., None] / (scale / np.pi) - np.pi / 2)
    elif basis_mapping == "trigonometric":
        pts_local = (positions - aabb[0])[..., None] / scale * 2 * np.pi
        pts_local = torch.cat((torch.sin(pts_local), torch.cos(pts_local)), dim=-1)
    elif basis_mapping == "x":
        pts_local = (positions - aabb[0]).unsqueeze(-1) / scale
    else:
        raise ValueError(basis_mapping)
    return pts_local


# ---------------------------------------------------------------------------
# models/FactorFields.py:36-71  dct_dict
# ---------------------------------------------------------------------------

def dct_dict(n_atoms_fre, size, n_selete, dim=2):
    """models/FactorFields.py:36-71: DCT dictionary, kron'd to `dim` dims, rows
    sub-selected with array_split, rows L2-normalised."""
    p = n_atoms_fre
    dct = np.zeros((p, size))
    for k in range(p):
        basis = np.cos(np.arange(size) * k * math.pi / p)
        if k > 0:
            basis = basis - np.mean(basis)
        dct[k] = basis
    kron = np.kron(dct, dct)
    if dim == 3:
        kron = np.kron(kron, dct)
    if n_selete < kron.shape[0]:
        idx = [x[0] for x in np.array_split(np.arange(kron.shape[0]), n_selete)]
        kron = kron[idx]
    for col in range(kron.shape[0]):
        norm = np.linalg.norm(kron[col]) or 1
        kron[col] /= norm
    return torch.FloatTensor(kron)


# ---------------------------------------------------------------------------
# models/FactorFields.py:262-307  setup_params
# ---------------------------------------------------------------------------

def setup_params(cfg, aabb):
    """models/FactorFields.py:262-307 plus the literal path ('param_derivation:
    literal' takes basis_resos, freq_bands and a per-axis coeff_reso as given).
    Returns an AttrDict with in_dim, aabb, basis_reso, coeff_reso (DHW order),
    freq_bands (float32 tensor), T_basis, T_coeff, n_scene."""
    mode = cfg.defaults.mode
    p = AttrDict()
    p.in_dim = len(aabb[0]) - 1 if mode in ("images", "reconstructions") else len(aabb[0])
    p.aabb = torch.FloatTensor(aabb)[:, :p.in_dim]
    dims = list(cfg.model.basis_dims)
    ctype = cfg.model.coeff_type
    if cfg.model.get("param_derivation", "reference") == "literal":
        p.basis_reso = list(cfg.model.basis_resos)
        p.freq_bands = torch.FloatTensor(cfg.model.freq_bands)
        p.T_basis = int(sum(np.power(np.array(p.basis_reso), p.in_dim) * np.array(dims)))
        p.coeff_reso = N_to_reso(cfg.model.coeff_reso ** p.in_dim, p.aabb)[::-1]
        p.T_coeff = int(sum(dims) * np.prod(p.coeff_reso))
        p.n_scene = 1
        if mode == "images":
            p.coeff_reso = [int(aabb[1][-1])] + p.coeff_reso
            p.aabb = torch.FloatTensor(aabb)
        return p
    if "reconstruction" not in mode:
        p.basis_reso = (list(cfg.model.basis_resos) if "image" in mode else
                        np.round(np.array(cfg.model.basis_resos) * (min(aabb[1][:p.in_dim]) + 1)
                                 / 1024.0).astype("int").tolist())
        p.T_basis = (cfg.model.T_basis if cfg.model.T_basis > 0 else
                     int(sum(np.power(np.array(p.basis_reso), p.in_dim) * np.array(dims))))
        T_coeff = cfg.model.T_coeff if cfg.model.T_coeff > 0 else cfg.model.total_params - p.T_basis
        p.T_coeff = T_coeff if T_coeff > 0 else 8 ** p.in_dim * sum(dims)
        if mode == "image":
            p.freq_bands = max(aabb[1][:p.in_dim]) / torch.FloatTensor(p.basis_reso)
        else:
            p.freq_bands = torch.FloatTensor(cfg.model.freq_bands)
        p.coeff_reso = N_to_reso(p.T_coeff // sum(dims), p.aabb[:, :p.in_dim])[::-1]
        p.n_scene = 1
        if mode == "sdf":
            p.freq_bands = p.freq_bands * 0.5
        elif mode == "images":
            p.coeff_reso = [aabb[1][-1]] + p.coeff_reso
            p.aabb = torch.FloatTensor(aabb)
        if "vec" in ctype or "cp" in ctype or "vm" in ctype:
            p.coeff_reso = aabb[1]
    else:
        p.coeff_reso = N_to_reso(cfg.model.coeff_reso ** p.in_dim, p.aabb[:, :p.in_dim])[::-1]
        p.T_coeff = sum(dims) * int(np.prod(p.coeff_reso))
        p.T_basis = cfg.model.total_params - p.T_coeff
        scale = p.T_basis / sum(np.power(np.array(cfg.model.basis_resos), p.in_dim) * np.array(dims))
        scale = np.power(scale, 1.0 / p.in_dim)
        btype = cfg.model.basis_type
        p.basis_reso = (list(cfg.model.basis_resos) if ("vec" in btype or "cp" in btype) else
                        np.round(np.array(cfg.model.basis_resos) * scale).astype("int").tolist())
        if mode == "reconstructions" or "x" in btype or "vec" in btype or "cp" in btype:
            p.freq_bands = torch.FloatTensor(cfg.model.freq_bands)
        else:
            p.freq_bands = torch.FloatTensor(cfg.model.freq_bands) * (
                cfg.dataset.scene_reso / float(max(p.basis_reso)) / max(cfg.model.freq_bands))
        p.n_scene = int(aabb[1][-1]) if mode == "reconstructions" else 1
    return p


def vm_plane_reso(cfg, aabb_t):
    """Per-axis (x, y, z) resolution of the VM coefficient planes/lines: the
    cfg's coeff_reso cubed, distributed over the aabb like utils.py:53-57."""
    return N_to_reso(cfg.model.coeff_reso ** 3, aabb_t[:, :3])


MAT_MODE = [[0, 1], [0, 2], [1, 2]]   # models/FactorFields.py:213
VEC_MODE = [2, 1, 0]                  # models/FactorFields.py:214


class OracleField:
    """CPU restatement of FactorFields' field path (models/FactorFields.py:206-533)
    for basis_type in {grid, none} and coeff_type in {grid, vm, none}.

    ``vm`` with a grid basis is the TensoRF-style sum over m of plane_m * line_m
    (18 channels for the 18-channel basis); the reference's own ``vm``
    coefficient (three concatenated lines, :318-323/:451-460) cannot be
    multiplied with a grid basis (:524-527), so this combination has new,
    documented semantics.  Parameters are float32 (or float64 for FD checks).
    """

    def __init__(self, cfg, dtype=torch.float32, seed=0):
        self.cfg = cfg
        self.p = setup_params(cfg, cfg.dataset.aabb)
        self.in_dim = self.p.in_dim
        self.aabb = self.p.aabb.to(dtype)
        self.freq_bands = self.p.freq_bands.to(dtype)
        self.dtype = dtype
        dims = list(cfg.model.basis_dims)
        C = sum(dims)
        self.C = C
        mode = cfg.defaults.mode
        # coefficients (models/FactorFields.py:311-332)
        ctype = cfg.model.coeff_type
        self.coeffs = []
        if ctype == "grid":
            self.coeffs = [cfg.model.coef_init * torch.ones((1, C, *self.p.coeff_reso), dtype=dtype)]
        elif ctype == "vm":
            r = vm_plane_reso(cfg, self.p.aabb)
            self.vm_reso = r
            for m in range(3):
                a0, a1 = MAT_MODE[m]
                self.coeffs.append(cfg.model.coef_init * torch.ones((1, C, r[a1], r[a0]), dtype=dtype))
            for m in range(3):
                self.coeffs.append(cfg.model.coef_init * torch.ones((1, C, r[VEC_MODE[m]], 1), dtype=dtype))
        # basis (models/FactorFields.py:398-423, grid branch :405-408)
        self.basises = []
        if cfg.model.basis_type == "grid":
            D = self.in_dim
            for d, reso in zip(dims, self.p.basis_reso):
                b = dct_dict(int(np.power(d, 1.0 / D) + 1), reso, n_selete=d, dim=D)
                self.basises.append(b.reshape([1, d] + [reso] * D).to(dtype))
        # linear_mat (models/FactorFields.py:227-237, MLPMixer :113-159)
        g = torch.Generator().manual_seed(seed)
        in_dim = C
        self.mlp = []
        L, H, O = cfg.model.num_layers, cfg.model.hidden_dim, cfg.model.out_dim
        for l in range(L):
            li = in_dim if l == 0 else H
            lo = O if l == L - 1 else H
            bound = 1.0 / math.sqrt(li)
            w = (torch.rand((lo, li), generator=g, dtype=torch.float64) * 2 - 1) * bound
            b = None if l == L - 1 else (torch.rand((lo,), generator=g, dtype=torch.float64) * 2 - 1) * bound
            self.mlp.append((w.to(dtype), None if b is None else b.to(dtype)))
        self.mode = mode

    # -- helpers -----------------------------------------------------------
    def parameters(self):
        out = list(self.coeffs) + list(self.basises)
        for w, b in self.mlp:
            out.append(w)
            if b is not None:
                out.append(b)
        return out

    def requires_grad_(self, flag=True):
        for t in self.parameters():
            t.requires_grad_(flag)
        return self

    def normalize_coord(self, xyz):
        """models/FactorFields.py:635-637."""
        invaabbSize = 2.0 / (self.aabb[1] - self.aabb[0])
        return (xyz - self.aabb[0]) * invaabbSize - 1

    # -- models/FactorFields.py:425-465 ------------------------------------
    def get_coeff(self, xyz):
        N, dim = xyz.shape
        ctype = self.cfg.model.coeff_type
        pts = self.normalize_coord(xyz).view([1, -1] + [1] * (dim - 1) + [dim])
        if ctype == "grid":
            return F.grid_sample(self.coeffs[0], pts, mode=self.cfg.model.coef_mode,
                                 align_corners=False, padding_mode="border").view(-1, N).t()
        if ctype == "vm":
            pts = pts.view(-1, dim)
            out = None
            for m in range(3):
                a0, a1 = MAT_MODE[m]
                pc = torch.stack((pts[:, a0], pts[:, a1]), -1).view(1, -1, 1, 2)
                lc = torch.stack((torch.zeros_like(pts[:, 0]), pts[:, VEC_MODE[m]]), -1).view(1, -1, 1, 2)
                plane = F.grid_sample(self.coeffs[m], pc, mode=self.cfg.model.coef_mode,
                                      align_corners=False, padding_mode="border").view(-1, N).t()
                line = F.grid_sample(self.coeffs[3 + m], lc, mode=self.cfg.model.coef_mode,
                                     align_corners=False, padding_mode="border").view(-1, N).t()
                term = plane * line
                out = term if out is None else out + term
            return out
        raise ValueError(ctype)

    # -- models/FactorFields.py:467-516 ------------------------------------
    def get_basis(self, x):
        N = x.shape[0]
        if self.mode == "images":
            x = x[..., :-1]
        L = len(self.freq_bands)
        xyz = grid_mapping(x, self.freq_bands, self.aabb[:, :self.in_dim],
                           self.cfg.model.basis_mapping).view(1, *([1] * (self.in_dim - 1)), -1,
                                                              self.in_dim, L)
        outs = []
        for i in range(L):
            outs.append(F.grid_sample(self.basises[i], xyz[..., i], mode=self.cfg.model.basis_mode,
                                      align_corners=True).view(-1, N).T)
        return torch.cat(outs, dim=-1)

    # -- models/FactorFields.py:523-533 ------------------------------------
    def get_coding(self, x):
        has_c = self.cfg.model.coeff_type != "none"
        has_b = self.cfg.model.basis_type != "none"
        if has_c and has_b:
            coeff = self.get_coeff(x)
            basises = self.get_basis(x)
            return basises * coeff, coeff
        if has_c:
            coeff = self.get_coeff(x)
            return coeff, coeff
        basises = self.get_basis(x)
        return basises, basises

    # -- models/FactorFields.py:144-159 (MLPMixer.forward, pe = 0) ---------
    def linear_mat(self, h, keep=None, p=0.1):
        """``keep`` (bool, the shape of h) replaces the Bernoulli draw of the
        input dropout F.dropout(h, p=0.1) (:150-151): kept entries are scaled
        by 1/(1-p), computed as torch's CPU dropout does (mask / (1-p), then
        the product)."""
        if keep is not None:
            noise = keep.to(h.dtype)
            noise = noise.div_(1 - p)
            h = h * noise
        n = len(self.mlp)
        for l, (w, b) in enumerate(self.mlp):
            h = F.linear(h, w, b)
            if l != n - 1:
                h = F.relu(h)
        return h

    def basis2density(self, f):
        """models/FactorFields.py:639-643."""
        if self.cfg.renderer.fea2denseAct == "softplus":
            return F.softplus(f + self.cfg.renderer.density_shift)
        return F.relu(f + self.cfg.renderer.density_shift)

    def field(self, x):
        feats, _ = self.get_coding(x)
        return self.linear_mat(feats)

    def n_parameters(self):
        return sum(t.numel() for t in self.parameters())
