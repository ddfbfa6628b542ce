"""Checkpoint format (models/FactorFields.py:669-691): state_dict keys and
shapes in the reference's layout, a save/load round trip (values, the
texel-major storage of the tables, the packed occupancy mask), on the CPU."""
import torch

from ffields import FactorFields
from ffields.render import AlphaGridMask
from helpers import small_cfg


def _model(seed=0, **kw):
    m = FactorFields(small_cfg(**kw), "cpu")
    g = torch.Generator().manual_seed(seed)
    with torch.no_grad():
        for p in m.parameters():
            p.copy_(torch.randn(p.shape, generator=g))
    return m


def test_state_dict_uses_reference_names_and_shapes():
    m = _model(mode="sdf", D=3, ctype="grid")
    sd = m.state_dict()
    assert [k for k in sd if k.startswith("basises.")] == [f"basises.{i}" for i in range(6)]
    assert "coeffs.0" in sd
    assert {"linear_mat.backbone.0.weight", "linear_mat.backbone.0.bias", "linear_mat.backbone.1.weight"} <= set(sd)
    # reference shapes [1, C, D, H, W] (models/FactorFields.py:311-332, :405-408)
    assert tuple(sd["coeffs.0"].shape)[:2] == (1, 18)
    assert [tuple(sd[f"basises.{i}"].shape) for i in range(6)] == [
        (1, d, r, r, r) for d, r in zip([4, 4, 4, 2, 2, 2], [9, 11, 13, 7, 6, 5])]


def test_save_load_round_trip(tmp_path):
    cfg_kw = dict(mode="sdf", D=3, ctype="vm")
    a = _model(seed=1, **cfg_kw)
    vol = (torch.rand((7, 9, 8), generator=torch.Generator().manual_seed(2)) < 0.5).float()
    a.alphaMask = AlphaGridMask(torch.device("cpu"), torch.tensor([[-1.0, -0.5, -1.2], [0.9, 1.4, 0.7]]), vol)
    path = tmp_path / "ckpt.th"
    a.save(str(path))
    ckpt = torch.load(str(path), weights_only=False)   # written by this test
    b = _model(seed=3, **cfg_kw)
    b.load(ckpt)
    for (k, va), (k2, vb) in zip(a.state_dict().items(), b.state_dict().items()):
        assert k == k2 and torch.equal(va, vb), k
    for t in list(b.basises) + list(b.coeffs):     # tables stay texel-major after loading
        assert t.is_contiguous(memory_format=torch.channels_last_3d if t.dim() == 5 else torch.channels_last)
    assert torch.equal(b.alphaMask.alpha_volume.cpu(), a.alphaMask.alpha_volume.cpu())
    assert torch.equal(b.alphaMask.aabb.cpu(), a.alphaMask.aabb.cpu())


def test_proposal_state_is_kept_beside_the_reference_state_dict(tmp_path):
    """A model with a learned proposal saves the reference's state_dict keys
    only (its strict load_state_dict, models/FactorFields.py:686, would reject
    proposal.*), and the proposal under its own entry; loading restores it."""
    cfg_kw = dict(mode="sdf", D=3, ctype="grid")
    a = _model(seed=4, **cfg_kw)
    ref_keys = list(a.state_dict().keys())
    pr = a.enable_proposal(reso=8, n_fine=16, table_dtype="fp32", density_shift=-3.0, seed=5)
    with torch.no_grad():
        pr.grid.copy_(torch.randn(pr.grid.shape, generator=torch.Generator().manual_seed(6)))
    path = tmp_path / "ckpt.th"
    a.save(str(path))
    ckpt = torch.load(str(path), weights_only=False)   # written by this test
    assert list(ckpt["state_dict"].keys()) == ref_keys
    assert ckpt["proposal"]["args"] == {"reso": 8, "n_fine": 16, "density_shift": -3.0, "table_dtype": "fp32"}
    assert isinstance(ckpt["cfg"], dict) and not hasattr(ckpt["cfg"], "to_dict")
    b = _model(seed=7, **cfg_kw)
    b.load(ckpt)
    assert b.proposal is not None and b.proposal.n_fine == 16
    for (k, va), (k2, vb) in zip(a.proposal.state_dict().items(), b.proposal.state_dict().items()):
        assert k == k2 and torch.equal(va, vb), k
    for (k, va), (k2, vb) in zip(a.state_dict().items(), b.state_dict().items()):
        assert k == k2 and torch.equal(va, vb), k


def test_load_is_strict(tmp_path):
    import pytest
    a = _model(seed=8, mode="sdf", D=3, ctype="grid")
    path = tmp_path / "ckpt.th"
    a.save(str(path))
    good = torch.load(str(path), weights_only=False)   # written by this test
    b = _model(seed=9, mode="sdf", D=3, ctype="grid")
    before = {k: v.clone() for k, v in b.state_dict().items()}
    for mutate, what in ((lambda sd: sd.pop("basises.3"), "missing"),
                         (lambda sd: sd.__setitem__("renderModule.mlp.0.weight", torch.zeros(2, 2)), "unexpected"),
                         (lambda sd: sd.__setitem__("coeffs.0", sd["coeffs.0"][..., :-1]), "shapes differ")):
        ck = dict(good)
        ck["state_dict"] = dict(good["state_dict"])
        mutate(ck["state_dict"])
        with pytest.raises(RuntimeError, match=what):
            b.load(ck)
        # nothing was copied before the check failed
        for k, v in b.state_dict().items():
            assert torch.equal(v, before[k]), k


def test_config_from_any_mapping():
    """A mapping-like config (an OmegaConf DictConfig in a reference
    checkpoint; a minimal Mapping here, omegaconf is not installed) wraps into
    attribute access and saves as plain dicts."""
    from collections.abc import Mapping

    from ffields.config import AttrDict, plain_cfg

    class Cfg(Mapping):
        def __init__(self, d):
            self._d = d

        def __getitem__(self, k):
            v = self._d[k]
            return Cfg(v) if isinstance(v, dict) else v

        def __iter__(self):
            return iter(self._d)

        def __len__(self):
            return len(self._d)

    c = Cfg({"model": {"basis_dims": [4, 2], "coeff_type": "grid"}, "renderer": {"step_ratio": 0.5}})
    w = AttrDict.wrap(c)
    assert w.model.basis_dims == [4, 2] and w.model.coeff_type == "grid" and w.renderer.step_ratio == 0.5
    p = plain_cfg(c)
    assert type(p) is dict and type(p["model"]) is dict and p == {"model": {"basis_dims": [4, 2], "coeff_type": "grid"},
                                                                  "renderer": {"step_ratio": 0.5}}


def test_upsample_google_objs_coefficient_grid():
    """models/FactorFields.py:704-707: a single-scene google_objs model with a
    grid coefficient resizes it by 1.3 per axis (trilinear, align_corners);
    other configurations keep their coefficients.  (CPU: only the render
    parameters and the resize run, no kernel.)"""
    import torch.nn.functional as F
    from ffields.ops import is_texel_major
    from test_gpu_render import nerf_small_cfg
    cfg = nerf_small_cfg()
    cfg.dataset.dataset_name = "google_objs"
    m = FactorFields(cfg, "cpu")
    g = torch.Generator().manual_seed(5)
    with torch.no_grad():
        m.coeffs[0].copy_(torch.randn(m.coeffs[0].shape, generator=g))
    old = m.coeffs[0].detach().clone()
    m.upsample_volume_grid([20, 22, 24])
    want = F.interpolate(old.contiguous(), scale_factor=1.3, align_corners=True, mode="trilinear")
    c = m.coeffs[0]
    assert tuple(c.shape) == tuple(want.shape) and list(m.coeff_reso) == list(want.shape[2:])
    assert torch.allclose(c.detach(), want, rtol=0, atol=1e-6)
    assert is_texel_major(c)
    assert m._spec.coeff_reso_xyz == list(want.shape[2:])[::-1]
    assert m.gridSize.tolist() == [20, 22, 24]
    m2 = FactorFields(nerf_small_cfg(), "cpu")
    shape = tuple(m2.coeffs[0].shape)
    m2.upsample_volume_grid([20, 22, 24])
    assert tuple(m2.coeffs[0].shape) == shape
