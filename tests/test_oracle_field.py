"""Pins the oracle's field path against closed forms, an independent numpy
brute-force interpolator and the parameter counts printed by the reference's
notebooks (scripts/2D_regression.ipynb:115, scripts/sdf_regression.ipynb:207)."""
import numpy as np
import pytest
import torch

import npref
from oracle import (AttrDict, N_to_reso, OracleField, dct_dict, grid_mapping, make_cfg,
                    setup_params)


# ---------------------------------------------------------------- mapping
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_sawtooth_closed_form(dtype):
    g = torch.Generator().manual_seed(1)
    aabb = torch.tensor([[-1.5, -1.5, -1.5], [1.5, 1.5, 1.5]], dtype=dtype)
    freq = torch.tensor([2.0, 3.2, 4.4, 5.6, 6.8, 8.0], dtype=dtype)
    p = (torch.rand((4000, 3), generator=g, dtype=dtype) * 3.4 - 1.7)
    out = grid_mapping(p, freq, aabb, "sawtooth")
    assert out.shape == (4000, 3, 6)
    scale = (3.0 / freq.numpy().astype(np.float64))
    ref = npref.sawtooth(p.numpy().astype(np.float64)[..., None], -1.5, scale)
    tol = 1e-12 if dtype == torch.float64 else 2e-5
    # remainder near a period boundary can legitimately wrap to the other end
    d = np.abs(out.numpy() - ref)
    wrap = np.abs(np.abs(out.numpy() - ref) - 2.0) < 1e-4
    assert np.all((d < tol) | wrap)
    assert (out >= -1).all() and (out <= 1).all()


def test_sawtooth_identity_special_case():
    """With freq = 1 the sawtooth of a point inside the box is the plain
    align_corners=True normalisation (p - a0) / (size / 2) - 1."""
    aabb = torch.tensor([[0.0, 0.0], [512.0, 512.0]], dtype=torch.float64)
    p = torch.rand((1000, 2), dtype=torch.float64) * 511.9
    out = grid_mapping(p, torch.tensor([1.0], dtype=torch.float64), aabb, "sawtooth")[..., 0]
    assert torch.allclose(out, (p - 0.0) / 256.0 - 1.0, atol=1e-12)


def test_sawtooth_periodic():
    aabb = torch.tensor([[0.0, 0.0, 0.0], [2.0, 2.0, 2.0]], dtype=torch.float64)
    freq = torch.tensor([2.0, 4.0], dtype=torch.float64)
    p = torch.rand((500, 3), dtype=torch.float64) * 0.9 + 0.05
    a = grid_mapping(p, freq, aabb, "sawtooth")
    b = grid_mapping(p + 1.0, freq, aabb, "sawtooth")   # 1.0 = one period of freq 2, two of freq 4
    assert torch.allclose(a, b, atol=1e-9)


def test_triangle_and_others_closed_form():
    aabb = torch.tensor([[0.0, 0.0], [4.0, 4.0]], dtype=torch.float64)
    freq = torch.tensor([1.0, 2.5], dtype=torch.float64)
    p = torch.rand((800, 2), dtype=torch.float64) * 4.0
    scale = 4.0 / freq.numpy()
    tri = grid_mapping(p, freq, aabb, "triangle").numpy()
    assert np.allclose(tri, npref.triangle(p.numpy()[..., None], 0.0, scale), atol=1e-9)
    sinc = grid_mapping(p, freq, aabb, "sinc").numpy()
    assert np.allclose(sinc, np.sin(p.numpy()[..., None] / (scale / np.pi) - np.pi / 2), atol=1e-12)
    x = grid_mapping(p, freq, aabb, "x").numpy()
    assert np.allclose(x, p.numpy()[..., None] / scale, atol=1e-12)
    tr = grid_mapping(p, freq, aabb, "trigonometric").numpy()
    ang = p.numpy()[..., None] / scale * 2 * np.pi
    assert np.allclose(tr, np.concatenate([np.sin(ang), np.cos(ang)], -1), atol=1e-12)


def test_fp32_remainder_is_fmod_plus_divisor():
    """torch.remainder in float32 equals C fmod, plus the divisor when the signs
    differ -- the formula the HIP mapping kernel implements."""
    g = torch.Generator().manual_seed(7)
    a = (torch.rand(200000, generator=g) * 6 - 3).float()
    for b in (0.375, 0.9375, 1.5, 0.3, 1.7):
        bt = torch.tensor(b, dtype=torch.float32)
        r = torch.remainder(a, bt).numpy()
        f = np.fmod(a.numpy(), np.float32(b))
        f = np.where((f != 0) & ((f < 0) != (np.float32(b) < 0)), f + np.float32(b), f).astype(np.float32)
        assert np.array_equal(r, f)


# ---------------------------------------------------------------- init / shapes
@pytest.mark.parametrize("d,D,rows", [(4, 2, [0, 3, 5, 7]), (2, 2, [0, 2]), (4, 3, [0, 2, 4, 6]),
                                      (2, 3, [0, 4])])
def test_dct_selected_rows(d, D, rows):
    p = int(np.power(d, 1.0 / D) + 1)
    ref, starts = npref.dct_rows(p, 9, D, d)
    assert starts == rows
    got = dct_dict(p, 9, n_selete=d, dim=D).numpy()
    assert np.allclose(got, ref, atol=1e-6)
    assert np.allclose(np.linalg.norm(got, axis=1), 1.0, atol=1e-6)


def test_dct_large_set():
    p = int(np.power(32, 0.5) + 1)
    assert p == 6
    ref, starts = npref.dct_rows(p, 16, 2, 32)
    assert starts[:8] == [0, 2, 4, 6, 8, 9, 10, 11]
    assert np.allclose(dct_dict(p, 16, 32, 2).numpy(), ref, atol=1e-6)


def _param_count(cfg, aabb, out_dim):
    p = setup_params(cfg, aabb)
    C = sum(cfg.model.basis_dims)
    mlp = 0
    L, H = cfg.model.num_layers, cfg.model.hidden_dim
    for l in range(L):
        li = C if l == 0 else H
        lo = out_dim if l == L - 1 else H
        mlp += li * lo + (0 if l == L - 1 else lo)
    return p.T_basis + C * int(np.prod(p.coeff_reso)), mlp, p


def test_param_count_image_notebook():
    """configs/image.yaml on a square image: 1,429,376 parameters
    (scripts/2D_regression.ipynb:115), coefficient 1x144x63x63 (:117)."""
    cfg = make_cfg("image", dict(basis_dims=[32, 32, 32, 16, 16, 16], total_params=1426063,
                                 coef_init=0.001, out_dim=4, num_layers=2, hidden_dim=64,
                                 freq_bands=[2.0, 3.2, 4.4, 5.6, 6.8, 8.0], in_dim=2),
                   dict(aabb=[[0.0, 0.0], [1024.0, 1024.0]]))
    fields, mlp, p = _param_count(cfg, cfg.dataset.aabb, 4)
    assert p.coeff_reso == [63, 63]
    assert p.T_basis == 848304
    assert fields + mlp == 1429376


def test_param_count_sdf_notebook():
    """configs/sdf.yaml on the [0,640]^3 SDF box: 5,342,274 parameters
    (scripts/sdf_regression.ipynb:207)."""
    cfg = make_cfg("sdf", dict(total_params=5313942, coeff_reso=32, coef_init=0.05, out_dim=1,
                               num_layers=1, hidden_dim=64, freq_bands=[2.0, 3.2, 4.4, 5.6, 6.8, 8.0]),
                   dict(aabb=[[0.0, 0.0, 0.0], [640.0, 640.0, 640.0]]))
    fields, mlp, p = _param_count(cfg, cfg.dataset.aabb, 1)
    assert p.basis_reso == [20, 32, 44, 56, 68, 80]
    assert torch.allclose(p.freq_bands, torch.tensor([1.0, 1.6, 2.2, 2.8, 3.4, 4.0]))
    assert p.coeff_reso == [54, 54, 54]
    assert fields + mlp == 5342274


def test_nerf_literal_via_reference_keys():
    """The NeRF workload's literal basis resos are reachable via reference keys:
    total_params 10,747,166, coeff_reso 32, scene_reso 1024."""
    cfg = make_cfg("reconstruction", dict(total_params=10747166, coeff_reso=32,
                                          freq_bands=[2.0, 3.2, 4.4, 5.6, 6.8, 8.0]),
                   dict(aabb=[[-1.5, -1.5, -1.5], [1.5, 1.5, 1.5]], scene_reso=1024))
    p = setup_params(cfg, cfg.dataset.aabb)
    assert p.basis_reso == [32, 51, 70, 89, 108, 128]
    assert p.coeff_reso == [32, 32, 32]
    assert torch.allclose(p.freq_bands, torch.tensor([2.0, 3.2, 4.4, 5.6, 6.8, 8.0]), rtol=1e-6)


def test_n_to_reso():
    assert N_to_reso(128 ** 3, torch.tensor([[-1.5] * 3, [1.5] * 3])) == [128, 128, 128]
    assert N_to_reso(4096, torch.tensor([[0.0, 0.0], [512.0, 256.0]])) == [91, 45]


# ---------------------------------------------------------------- gathers
def _small_cfg(mode, D, ctype, cmode="bilinear", bmode="bilinear", mapping="sawtooth"):
    if D == 2:
        aabb = [[0.0, 0.0], [8.0, 6.0]]
    else:
        aabb = [[-1.0, -1.0, -1.0], [1.0, 1.0, 1.0]]
    return make_cfg(mode, dict(basis_dims=[4, 2, 2], basis_resos=[3, 4, 5], freq_bands=[1.0, 1.7, 2.9],
                               coeff_reso=4, coeff_type=ctype, coef_mode=cmode, basis_mode=bmode,
                               basis_mapping=mapping, param_derivation="literal", out_dim=3,
                               hidden_dim=5, num_layers=2),
                    dict(aabb=aabb))


def _randomise(of, seed=0):
    g = torch.Generator().manual_seed(seed)
    for t in of.coeffs + of.basises:
        t.copy_(torch.randn(t.shape, generator=g, dtype=t.dtype))


@pytest.mark.parametrize("D,ctype,cmode,bmode", [(2, "grid", "bilinear", "bilinear"),
                                                 (2, "grid", "nearest", "nearest"),
                                                 (3, "grid", "bilinear", "bilinear"),
                                                 (3, "vm", "bilinear", "bilinear"),
                                                 (3, "grid", "nearest", "bilinear")])
def test_gathers_vs_numpy_bruteforce(D, ctype, cmode, bmode):
    cfg = _small_cfg("sdf" if D == 3 else "image", D, ctype, cmode, bmode)
    of = OracleField(cfg, dtype=torch.float64)
    _randomise(of)
    g = torch.Generator().manual_seed(3)
    lo, hi = of.aabb[0], of.aabb[1]
    x = lo + (hi - lo) * (torch.rand((60, D), generator=g, dtype=torch.float64) * 1.2 - 0.1)
    # basis: each level is grid_sample(align_corners=True, zeros) at the sawtooth coordinate
    xyz = grid_mapping(x, of.freq_bands, of.aabb, "sawtooth").numpy()
    ref_b = np.concatenate([npref.grid_sample(of.basises[i][0].numpy(), xyz[..., i], True, "zeros", bmode)
                            for i in range(3)], -1)
    assert np.allclose(of.get_basis(x).numpy(), ref_b, atol=1e-12)
    # coefficient: align_corners=False, border
    pts = ((x - of.aabb[0]) * (2.0 / (of.aabb[1] - of.aabb[0])) - 1).numpy()
    if ctype == "grid":
        ref_c = npref.grid_sample(of.coeffs[0][0].numpy(), pts, False, "border", cmode)
    else:
        ref_c = 0
        for m, (a0, a1) in enumerate([[0, 1], [0, 2], [1, 2]]):
            pl = npref.grid_sample(of.coeffs[m][0].numpy(), pts[:, [a0, a1]], False, "border", cmode)
            lc = np.stack([np.zeros(len(pts)), pts[:, [2, 1, 0][m]]], -1)
            li = npref.grid_sample(of.coeffs[3 + m][0].numpy(), lc, False, "border", cmode)
            ref_c = ref_c + pl * li
    assert np.allclose(of.get_coeff(x).numpy(), ref_c, atol=1e-12)
    feats, coeff = of.get_coding(x)
    assert np.allclose(feats.numpy(), ref_b * ref_c, atol=1e-12)
    ref_out = npref.mlp(ref_b * ref_c, [(w.numpy(), None if b is None else b.numpy()) for w, b in of.mlp])
    assert np.allclose(of.linear_mat(feats).numpy(), ref_out, atol=1e-12)


def test_partition_of_unity():
    cfg = _small_cfg("sdf", 3, "grid")
    of = OracleField(cfg, dtype=torch.float64)
    for t in of.coeffs + of.basises:
        t.fill_(2.5)
    x = torch.rand((100, 3), dtype=torch.float64) * 2 - 1
    assert torch.allclose(of.get_coeff(x), torch.full((100, 8), 2.5, dtype=torch.float64))
    assert torch.allclose(of.get_basis(x), torch.full((100, 8), 2.5, dtype=torch.float64))


def test_oracle_gradients_fp64_central_differences():
    cfg = _small_cfg("sdf", 3, "vm")
    of = OracleField(cfg, dtype=torch.float64)
    _randomise(of, 5)
    x = torch.rand((7, 3), dtype=torch.float64) * 2 - 1
    params = of.parameters()
    of.requires_grad_(True)
    loss = (of.field(x) ** 2).sum()
    grads = torch.autograd.grad(loss, params)
    of.requires_grad_(False)
    g = torch.Generator().manual_seed(11)
    eps = 1e-6
    for t, gt in zip(params, grads):
        flat = t.view(-1)
        for _ in range(4):
            j = int(torch.randint(0, flat.numel(), (1,), generator=g))
            old = flat[j].item()
            flat[j] = old + eps
            lp = (of.field(x) ** 2).sum().item()
            flat[j] = old - eps
            lm = (of.field(x) ** 2).sum().item()
            flat[j] = old
            fd = (lp - lm) / (2 * eps)
            assert abs(fd - gt.view(-1)[j].item()) <= 1e-6 * max(1.0, abs(fd)), (t.shape, j)


def test_attrdict():
    a = AttrDict.wrap({"x": {"y": [1, {"z": 2}]}})
    assert a.x.y[1].z == 2
