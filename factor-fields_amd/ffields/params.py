"""Host-side shape derivation and parameter initialisation.

``setup_params`` follows models/FactorFields.py:262-307 (mode-dependent
derivation of in_dim, aabb, basis_reso, coeff_reso, freq_bands, T_basis,
T_coeff, n_scene) with the same FP32 torch arithmetic, plus a literal path
(``model.param_derivation: literal``) that takes basis_resos, freq_bands and
coeff_reso as given.  ``dct_basis`` builds the DCT dictionary of
models/FactorFields.py:36-71 directly from separable outer products of the
selected rows instead of forming the full Kronecker product.
"""

import math

import numpy as np
import torch


def N_to_reso(n_voxels, bbox):
    """utils.py:53-57."""
    xyz_min, xyz_max = bbox
    dim = len(xyz_min)
    voxel_size = ((xyz_max - xyz_min).prod() / n_voxels).pow(1 / dim)
    return torch.round((xyz_max - xyz_min) / voxel_size).long().tolist()


def cal_n_samples(reso, step_ratio=0.5):
    """utils.py:69-70."""
    return int(np.linalg.norm(reso) / step_ratio)


class Shapes:
    """Derived shapes of a FactorFields model (attributes as in the reference)."""

    def __repr__(self):
        return "Shapes(" + ", ".join(f"{k}={v!r}" for k, v in vars(self).items()) + ")"


def setup_params(cfg, aabb):
    mode = cfg.defaults.mode
    s = Shapes()
    s.in_dim = len(aabb[0]) - 1 if mode in ("images", "reconstructions") else len(aabb[0])
    s.aabb = torch.FloatTensor(aabb)[:, :s.in_dim]
    dims = list(cfg.model.basis_dims)
    ctype = cfg.model.coeff_type
    btype = cfg.model.basis_type
    if cfg.model.get("param_derivation", "reference") == "literal":
        s.basis_reso = [int(r) for r in cfg.model.basis_resos]
        s.freq_bands = torch.FloatTensor(cfg.model.freq_bands)
        s.T_basis = int(sum(np.power(np.array(s.basis_reso), s.in_dim) * np.array(dims)))
        s.coeff_reso = N_to_reso(cfg.model.coeff_reso ** s.in_dim, s.aabb)[::-1]
        s.T_coeff = int(sum(dims) * np.prod(s.coeff_reso))
        s.n_scene = 1
        if mode == "images":
            s.coeff_reso = [int(aabb[1][-1])] + s.coeff_reso
            s.aabb = torch.FloatTensor(aabb)
        return s
    if "reconstruction" not in mode:
        s.basis_reso = ([int(r) for r in cfg.model.basis_resos] if "image" in mode else
                        np.round(np.array(cfg.model.basis_resos) * (min(aabb[1][:s.in_dim]) + 1)
                                 / 1024.0).astype("int").tolist())
        s.T_basis = (cfg.model.T_basis if cfg.model.T_basis > 0 else
                     int(sum(np.power(np.array(s.basis_reso), s.in_dim) * np.array(dims))))
        T_coeff = cfg.model.T_coeff if cfg.model.T_coeff > 0 else cfg.model.total_params - s.T_basis
        s.T_coeff = T_coeff if T_coeff > 0 else 8 ** s.in_dim * sum(dims)
        if mode == "image":
            s.freq_bands = max(aabb[1][:s.in_dim]) / torch.FloatTensor(s.basis_reso)
        else:
            s.freq_bands = torch.FloatTensor(cfg.model.freq_bands)
        s.coeff_reso = N_to_reso(s.T_coeff // sum(dims), s.aabb[:, :s.in_dim])[::-1]
        s.n_scene = 1
        if mode == "sdf":
            s.freq_bands = s.freq_bands * 0.5
        elif mode == "images":
            s.coeff_reso = [int(aabb[1][-1])] + s.coeff_reso
            s.aabb = torch.FloatTensor(aabb)
        if "vec" in ctype or "cp" in ctype or "vm" in ctype:
            s.coeff_reso = [int(v) for v in aabb[1]]
    else:
        s.coeff_reso = N_to_reso(cfg.model.coeff_reso ** s.in_dim, s.aabb[:, :s.in_dim])[::-1]
        s.T_coeff = sum(dims) * int(np.prod(s.coeff_reso))
        s.T_basis = cfg.model.total_params - s.T_coeff
        scale = s.T_basis / sum(np.power(np.array(cfg.model.basis_resos), s.in_dim) * np.array(dims))
        scale = np.power(scale, 1.0 / s.in_dim)
        s.basis_reso = ([int(r) for r in cfg.model.basis_resos] if ("vec" in btype or "cp" in btype) else
                        np.round(np.array(cfg.model.basis_resos) * scale).astype("int").tolist())
        if mode == "reconstructions" or "x" in btype or "vec" in btype or "cp" in btype:
            s.freq_bands = torch.FloatTensor(cfg.model.freq_bands)
        else:
            s.freq_bands = torch.FloatTensor(cfg.model.freq_bands) * (
                cfg.dataset.scene_reso / float(max(s.basis_reso)) / max(cfg.model.freq_bands))
        s.n_scene = int(aabb[1][-1]) if mode == "reconstructions" else 1
    return s


def vm_reso(cfg, aabb):
    """Per-axis (x, y, z) resolution of the VM coefficient planes and lines:
    coeff_reso^3 voxels spread over the box like utils.py:53-57."""
    return N_to_reso(int(cfg.model.coeff_reso) ** 3, aabb[:, :3])


def dct_basis(d, reso, dim):
    """[d, reso^dim] float32 rows of the DCT dictionary of models/FactorFields.py:36-71
    with p = int(d^(1/dim) + 1) frequencies per axis, rows picked at the starts
    of np.array_split(arange(p^dim), d), each L2-normalised."""
    p = int(np.power(d, 1.0 / dim) + 1)
    one = np.zeros((p, reso))
    for k in range(p):
        v = np.cos(np.arange(reso) * k * math.pi / p)
        if k > 0:
            v = v - np.mean(v)
        one[k] = v
    total = p ** dim
    picks = ([int(a[0]) for a in np.array_split(np.arange(total), d)] if d < total
             else list(range(total)))
    rows = np.zeros((len(picks), reso ** dim))
    for n, r in enumerate(picks):
        digits = []
        q = r
        for _ in range(dim):
            digits.append(q % p)
            q //= p
        digits.reverse()                  # first Kronecker factor = slowest axis
        v = one[digits[0]]
        for dg in digits[1:]:
            v = np.multiply.outer(v, one[dg])
        v = v.reshape(-1)
        nrm = np.linalg.norm(v) or 1
        rows[n] = v / nrm
    return torch.FloatTensor(rows)
