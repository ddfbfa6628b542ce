"""GPU tests of occupancy-grid maintenance (models/FactorFields.py:710-793)."""
import pytest
import torch
import torch.nn.functional as F

from helpers import assert_close, oracle_of
from test_gpu_render import make_model
from oracle import alpha_grid_sample
import ffields._lib as L

pytestmark = pytest.mark.gpu


def test_pool_threshold_matches_max_pool3d():
    g = torch.Generator().manual_seed(0)
    a = torch.rand((13, 17, 11), generator=g) * 3 - 1     # values outside [0,1] exercise the clamp
    out = torch.empty((13, 17, 11), dtype=torch.uint8, device="cuda")
    ac = a.cuda()
    L.check(L.lib().ff_occupancy_pool(L.ptr(ac), 13, 17, 11, 2.0, 0.3, L.ptr(out), L.stream_of(ac)))
    ref = F.max_pool3d((a / 2.0).clamp(0, 1)[None, None], kernel_size=3, padding=1, stride=1)[0, 0] >= 0.3
    assert torch.equal(out.cpu().bool(), ref)


def test_dense_alpha_matches_oracle():
    m = make_model(seed=3)
    torch.manual_seed(0)
    alpha, dense = m.getDenseAlpha([9, 10, 11], times=1)
    of = oracle_of(m)
    x = dense.reshape(-1, 3).cpu()
    sig = of.basis2density(of.field(x)[:, 0])
    am = m.alphaMask
    keep = alpha_grid_sample(am.alpha_volume.cpu()[0, 0], am.aabb.cpu(), x) > 0
    units = (m.aabb[1] - m.aabb[0]).cpu() / (torch.tensor([9, 10, 11]) - 1)
    length = float(units.mean() * 25.0)                # getDenseAlpha's own step (FactorFields.py:735-737)
    ref = torch.where(keep, 1 - torch.exp(-sig * length), torch.zeros_like(sig))
    assert_close(alpha.reshape(-1), ref, 1e-5, "alpha")


def test_update_alpha_mask_runs_and_shrinks_box():
    m = make_model(seed=5, occupancy=False)
    box = m.updateAlphaMask((16, 16, 16), is_update_alphaMask=True, times=2)
    assert m.alphaMask is not None
    assert box.shape == (2, 3)
    occ = m.alphaMask.occ
    assert occ.shape == (16, 16, 16) and int(occ.sum()) > 0
