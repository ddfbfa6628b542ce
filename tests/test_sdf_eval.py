"""SDF evaluation helpers (scripts/sdf_regression.ipynb, cell 1) on analytic
fields: the grid layout of eval_sdf against a linear field's closed form, the
L1 error and outside-IoU of two concentric spheres against their closed
forms, and the zero level set export against the sphere it comes from."""
import math

import torch

from ffields.mesh import read_ply
from ffields.sdf import cal_l1_iou, eval_sdf, export_sdf_mesh, marchcude_to_world


class Analytic:
    """Stands in for a model: field(x) [N, 3] -> [N, 1] on the CPU."""

    device = "cpu"

    def __init__(self, fn):
        self.fn = fn
        self.calls = 0

    def field(self, x):
        self.calls += 1
        return self.fn(x.double()).float()[:, None]


def test_eval_sdf_layout():
    m = Analytic(lambda p: p[:, 0] + 10 * p[:, 1] + 100 * p[:, 2])
    reso, ext = (5, 6, 7), (2.0, 3.0, 4.0)
    want = (torch.linspace(0, 2, 5)[:, None, None] + 10 * torch.linspace(0, 3, 6)[None, :, None]
            + 100 * torch.linspace(0, 4, 7)[None, None, :])
    for chunk in (1, 84, 100, 1 << 20):      # one slab, two slabs with a ragged end, all at once
        m.calls = 0
        vol = eval_sdf(m, reso, ext, chunk=chunk)
        assert vol.shape == (5, 6, 7)
        assert torch.allclose(vol, want, rtol=0, atol=1e-4)
        assert m.calls == {1: 5, 84: 3, 100: 3, 1 << 20: 1}[chunk]
    cube = eval_sdf(m, (4, 4, 4), 3.0)
    g = torch.linspace(0, 3, 4)
    assert torch.allclose(cube, g[:, None, None] + 10 * g[None, :, None] + 100 * g[None, None, :], atol=1e-4)


def test_l1_iou_spheres():
    r_true, r_pred = 0.6, 0.5
    n = 96
    c = (torch.arange(n, dtype=torch.float64) + 0.5) / n * 2 - 1      # cell centres of [-1, 1]
    p = torch.stack(torch.meshgrid(c, c, c, indexing="ij"), -1).reshape(-1, 3).float()
    sdf = (p.double().norm(dim=1) - r_true).float()[:, None]
    m = Analytic(lambda x: x.norm(dim=1) - r_pred)
    l1, iou = cal_l1_iou(m, sdf, p, chunk=200_000)
    assert abs(l1.item() - (r_true - r_pred)) < 1e-6
    # outside both / outside either = outside the larger / outside the smaller sphere
    want = (8 - 4 / 3 * math.pi * r_true ** 3) / (8 - 4 / 3 * math.pi * r_pred ** 3)
    assert abs(iou.item() - want) < 2e-3, (iou.item(), want)
    l1b, ioub = cal_l1_iou(m, sdf, p, chunk=1 << 30)
    assert torch.equal(iou, ioub) and abs(l1b.item() - l1.item()) < 1e-7


def test_marchcude_to_world():
    v = torch.tensor([[0.0, 0.0, 0.0], [9.0, 4.0, 19.0], [4.5, 2.0, 9.5]])
    w = marchcude_to_world(v, (10, 5, 20))
    assert torch.allclose(w, torch.tensor([[-1.0, -1, -1], [1, 1, 1], [0, 0, 0]], dtype=torch.float64))


def test_export_sdf_mesh(tmp_path, monkeypatch):
    """The export's wiring (grid, world coordinates, PLY) on the CPU, with the
    oracle's marching tetrahedra in place of the device kernels."""
    import ffields.sdf
    import oracle.mesh
    monkeypatch.setattr(ffields.sdf, "marching_tetrahedra", oracle.mesh.marching_tetrahedra)
    R = 0.6
    m = Analytic(lambda x: (x - 1.0).norm(dim=1) - R)                # sphere at the centre of [0, 2]^3
    reso = (33, 33, 33)
    path = tmp_path / "sdf.ply"
    verts, faces = export_sdf_mesh(m, str(path), reso, 2.0)
    h = 2.0 / 32
    assert faces.shape[0] > 500
    assert (verts.norm(dim=1) - R).abs().max().item() < 0.1 * h
    a, b, c = verts[faces[:, 0]], verts[faces[:, 1]], verts[faces[:, 2]]
    vol6 = (a * torch.cross(b, c, dim=1)).sum().item()
    assert abs(vol6 / 6 / (4 / 3 * math.pi * R ** 3) - 1) < 0.03      # outward faces: positive volume
    v, f = read_ply(path)
    assert v.shape == tuple(verts.shape) and f.shape == tuple(faces.shape)
    assert abs(float(v[0, 0]) - float(verts[0, 0])) < 1e-6
