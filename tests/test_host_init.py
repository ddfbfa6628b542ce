"""The library's own host-side initialisation (CPU): the DCT dictionary it
writes into the basis tables, the shapes and parameter counts its
setup_params derives, and the SH polynomials the colour head uses.

Pins: the oracle's dct_dict (itself pinned by closed-form rows in
tests/test_oracle_field.py), the parameter counts the reference's notebooks
print (scripts/2D_regression.ipynb:115-130, scripts/sdf_regression.ipynb:207),
and scipy's complex spherical harmonics for models/sh.py:4-64.
"""
import numpy as np
import pytest
import torch

from ffields import FactorFields
from ffields.params import dct_basis, setup_params as ff_setup_params
from oracle import dct_dict, eval_sh, make_cfg, setup_params as oracle_setup_params

RESOS = [32, 51, 70, 89, 108, 128]
SDF_RESOS = [20, 32, 44, 56, 68, 80]     # sdf.yaml's resos on the [0, 640]^3 box


@pytest.mark.parametrize("d,D,resos", [(4, 2, RESOS), (2, 2, RESOS), (32, 2, RESOS), (16, 2, RESOS),
                                       (4, 3, RESOS[:4] + SDF_RESOS), (2, 3, RESOS[:4] + SDF_RESOS)])
def test_dct_basis_matches_oracle(d, D, resos):
    """ffields.params.dct_basis (separable outer products) equals the oracle's
    Kronecker-product dct_dict (models/FactorFields.py:36-71) for every level
    shape the configs use."""
    p = int(np.power(d, 1.0 / D) + 1)
    for r in resos:
        got = dct_basis(d, r, D)
        ref = dct_dict(p, r, n_selete=d, dim=D)
        assert got.shape == ref.shape == (d, r ** D)
        assert torch.allclose(got, ref, atol=2e-7, rtol=0), (d, D, r, (got - ref).abs().max().item())


def _image_notebook_cfg():
    # configs/image.yaml:10-36 with the notebook's 1024x1024 image and out_dim 4
    # (scripts/2D_regression.ipynb:117-130 prints Linear(144, 64) and Linear(64, 4))
    return make_cfg("image", dict(basis_dims=[32, 32, 32, 16, 16, 16], basis_resos=RESOS,
                                  freq_bands=[2.0, 3.2, 4.4, 5.6, 6.8, 8.0], total_params=1426063,
                                  coef_init=0.001, coef_mode="nearest", basis_mode="nearest",
                                  in_dim=2, out_dim=4, num_layers=2, hidden_dim=64),
                    dict(aabb=[[0.0, 0.0], [1024.0, 1024.0]]))


def _sdf_notebook_cfg():
    # configs/sdf.yaml:9-31 on the [0, 640]^3 box of dataLoader/sdf.py:22-27
    return make_cfg("sdf", dict(total_params=5313942, coeff_reso=32, coef_init=0.05, out_dim=1,
                                num_layers=1, hidden_dim=64, freq_bands=[2.0, 3.2, 4.4, 5.6, 6.8, 8.0]),
                    dict(aabb=[[0.0, 0.0, 0.0], [640.0, 640.0, 640.0]], scene_reso=384))


def test_product_image_model_matches_notebook():
    """FactorFields(image.yaml) built by the library: 1,429,376 parameters
    (scripts/2D_regression.ipynb:115), coefficient 1x144x63x63 and basis levels
    1x32x{32,51,70}^2, 1x16x{89,108,128}^2 (:117-128)."""
    m = FactorFields(_image_notebook_cfg(), "cpu")
    assert m.n_parameters() == 1429376
    assert tuple(m.coeffs[0].shape) == (1, 144, 63, 63)
    assert [tuple(b.shape) for b in m.basises] == [(1, 32, 32, 32), (1, 32, 51, 51), (1, 32, 70, 70),
                                                   (1, 16, 89, 89), (1, 16, 108, 108), (1, 16, 128, 128)]
    assert [tuple(l.weight.shape) for l in m.linear_mat.backbone] == [(64, 144), (4, 64)]


def test_product_sdf_model_matches_notebook():
    """FactorFields(sdf.yaml) built by the library: 5,342,274 parameters
    (scripts/sdf_regression.ipynb:207), resos scaled to the box, freq halved."""
    m = FactorFields(_sdf_notebook_cfg(), "cpu")
    assert m.n_parameters() == 5342274
    assert m.basis_reso == SDF_RESOS
    assert torch.allclose(m.freq_bands.cpu(), torch.tensor([1.0, 1.6, 2.2, 2.8, 3.4, 4.0]))
    assert tuple(m.coeffs[0].shape) == (1, 18, 54, 54, 54)


@pytest.mark.parametrize("cfg_fn", [_image_notebook_cfg, _sdf_notebook_cfg])
def test_product_tables_are_the_reference_init(cfg_fn):
    """The basis tables hold dct_dict's rows and the coefficient coef_init
    (models/FactorFields.py:311-332, :405-408), in the reference's shapes."""
    cfg = cfg_fn()
    m = FactorFields(cfg, "cpu")
    D = m.in_dim
    for d, r, b in zip(m.basis_dims, m.basis_reso, m.basises):
        p = int(np.power(d, 1.0 / D) + 1)
        ref = dct_dict(p, r, n_selete=d, dim=D).reshape([1, d] + [r] * D)
        assert torch.allclose(b.detach(), ref, atol=2e-7, rtol=0)
    assert torch.all(m.coeffs[0].detach() == float(cfg.model.coef_init))


@pytest.mark.parametrize("mode,model,dataset", [
    ("image", dict(basis_dims=[32, 32, 32, 16, 16, 16], total_params=1426063, in_dim=2, out_dim=3),
     dict(aabb=[[0.0, 0.0], [512.0, 384.0]])),
    ("sdf", dict(total_params=5313942, coeff_reso=32, out_dim=1, num_layers=1),
     dict(aabb=[[0.0, 0.0, 0.0], [640.0, 640.0, 640.0]])),
    ("reconstruction", dict(total_params=5308416, coeff_reso=32,
                            freq_bands=[2.0, 3.2, 4.4, 5.6, 6.8, 8.0], hidden_dim=64),
     dict(aabb=[[-1.5, -1.5, -1.5], [1.5, 1.5, 1.5]], scene_reso=768)),
    ("images", dict(basis_dims=[32, 32, 32, 16, 16, 16], total_params=1024000, coeff_reso=32, out_dim=3),
     dict(aabb=[[0.0, 0.0, 0.0], [256.0, 256.0, 64.0]])),
])
def test_product_setup_params_equals_oracle(mode, model, dataset):
    """The library's setup_params (models/FactorFields.py:262-307) gives the
    oracle's shapes, resolutions and FP32 frequencies in every mode."""
    cfg = make_cfg(mode, model, dataset)
    a = ff_setup_params(cfg, cfg.dataset.aabb)
    b = oracle_setup_params(cfg, cfg.dataset.aabb)
    assert a.in_dim == b.in_dim
    assert a.basis_reso == b.basis_reso
    assert list(a.coeff_reso) == list(b.coeff_reso)
    assert a.T_basis == b.T_basis and a.T_coeff == b.T_coeff
    assert torch.equal(a.freq_bands, b.freq_bands)
    assert torch.equal(a.aabb, b.aabb)


def _real_sh_scipy(dirs):
    """Real spherical harmonics of degree 0..2 in the m = -l..l order of
    models/sh.py, from scipy's complex Y_l^m (Condon-Shortley phase kept):
    sqrt(2) Im Y_l^|m| for m < 0, Y_l^0, sqrt(2) Re Y_l^m for m > 0."""
    from scipy.special import sph_harm_y
    theta = np.arccos(np.clip(dirs[:, 2], -1, 1))
    phi = np.arctan2(dirs[:, 1], dirs[:, 0])
    cols = []
    for l in range(3):
        for m in range(-l, l + 1):
            y = sph_harm_y(l, abs(m), theta, phi)
            cols.append(np.sqrt(2) * y.imag if m < 0 else (np.sqrt(2) * y.real if m > 0 else y.real))
    return np.stack(cols, 1)


def test_eval_sh_matches_scipy_real_harmonics():
    """Every coefficient, sign and index of the oracle's degree-2 eval_sh
    (models/sh.py:34-64) against scipy's spherical harmonics on random unit
    directions and on the six axis directions (where each linear term and
    each quadratic term is isolated)."""
    rng = np.random.default_rng(7)
    d = rng.normal(size=(256, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    axes = np.concatenate([np.eye(3), -np.eye(3), np.array([[1, 1, 0], [0, 1, 1], [1, 0, 1]]) / np.sqrt(2)])
    d = np.concatenate([d, axes])
    ref = _real_sh_scipy(d)
    dt = torch.tensor(d)
    got = torch.stack([eval_sh(2, torch.eye(9, dtype=torch.float64)[k].expand(len(d), 1, 9), dt)[:, 0]
                       for k in range(9)], 1).numpy()
    assert np.abs(got - ref).max() < 1e-12
