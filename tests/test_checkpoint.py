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
