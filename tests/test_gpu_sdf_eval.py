"""SDF evaluation helpers on a VM-coefficient SDF model: the grid volume of
eval_sdf against the oracle's field at the same points, the L1 error and
outside-IoU against the same metrics of the oracle's predictions, and the
level-set export read back from its PLY file."""
import sys

import pytest
import torch

from ffields.mesh import read_ply
from ffields.sdf import cal_l1_iou, eval_sdf, export_sdf_mesh
from helpers import assert_close, new_model, oracle_of, small_cfg

pytestmark = pytest.mark.gpu

BOX = [[0.0, 0.0, 0.0], [2.0, 2.0, 2.0]]


def make():
    print("[case] sdf vm model", file=sys.stderr, flush=True)
    return new_model(small_cfg(mode="sdf", D=3, ctype="vm", out_dim=1, aabb=BOX), seed=4)


def test_eval_sdf_matches_oracle():
    m = make()
    reso = (20, 22, 24)
    vol = eval_sdf(m, reso, 2.0, chunk=3000)
    g = [torch.linspace(0, 2, r) for r in reso]
    p = torch.stack(torch.meshgrid(*g, indexing="ij"), -1).reshape(-1, 3)
    ref = oracle_of(m).field(p).reshape(reso)
    assert_close(vol, ref, 1e-5, "sdf volume")


def test_l1_iou_matches_oracle():
    m = make()
    g = torch.Generator().manual_seed(11)
    p = torch.rand((50_000, 3), generator=g) * 2
    pred_ref = oracle_of(m).field(p).reshape(-1).double()
    sdf = pred_ref.float() + torch.randn(p.shape[0], generator=g) * pred_ref.abs().mean().float()
    l1, iou = cal_l1_iou(m, sdf[:, None], p, chunk=7000)
    s = sdf.double()
    l1_ref = (pred_ref - s).abs().mean().item()
    inter = ((s > 0) & (pred_ref > 0)).sum().item()
    union = ((s > 0) | (pred_ref > 0)).sum().item()
    assert abs(l1.item() - l1_ref) <= 1e-5 * l1_ref
    # a prediction within FP32 rounding of zero may land on the other side
    near = (pred_ref.abs() < 1e-5 * pred_ref.abs().max()).sum().item()
    assert abs(iou.item() - inter / union) <= (near + 1e-9) / union + 1e-6


def test_export_sdf_mesh(tmp_path):
    m = make()
    reso = (24, 24, 24)
    level = float(eval_sdf(m, reso, 2.0).median())
    path = tmp_path / "sdf.ply"
    verts, faces = export_sdf_mesh(m, str(path), reso, 2.0, level=level)
    assert faces.shape[0] > 0
    v, f = read_ply(path)
    assert v.shape[0] == verts.shape[0] and f.shape[0] == faces.shape[0]
    assert float(abs(v).max()) <= 1.0 + 1e-6
    assert int(f.min()) >= 0 and int(f.max()) < v.shape[0]
    # the same level set as the oracle's marching tetrahedra of the same volume
    from oracle.mesh import marching_tetrahedra as oracle_mt
    from ffields.sdf import marchcude_to_world
    vol = eval_sdf(m, reso, 2.0)
    v_ref, f_ref = oracle_mt(vol.cpu().double(), level, False)
    assert torch.equal(verts.cpu(), marchcude_to_world(v_ref, reso))
    assert sorted(map(tuple, faces.cpu().tolist())) == sorted(map(tuple, f_ref.tolist()))
