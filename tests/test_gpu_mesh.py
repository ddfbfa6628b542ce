"""export_mesh on a render-path model: the dense alpha volume from the HIP
kernels, its isosurface, and a PLY file that reads back with every vertex in
the box (in the (z, y, x) axis order of the reference's export)."""
import sys

import pytest
import torch

from ffields.mesh import export_mesh, read_ply
from test_gpu_render import make_model

pytestmark = pytest.mark.gpu


def test_export_mesh(tmp_path):
    print("[case] export mesh", file=sys.stderr, flush=True)
    m = make_model(seed=9)
    alpha, _ = m.getDenseAlpha((24, 20, 16), times=1)
    level = float(torch.quantile(alpha.flatten().float(), 0.7))
    path = tmp_path / "field.ply"
    pts, faces = export_mesh(m, str(path), gridSize=(24, 20, 16), level=level, times=1)
    assert faces.shape[0] > 0
    v, f = read_ply(path)
    assert v.shape[0] == pts.shape[0] and f.shape[0] == faces.shape[0]
    box = m.aabb.cpu().double()
    lo = torch.stack([box[0, 2], box[0, 1], box[0, 0]])
    hi = torch.stack([box[1, 2], box[1, 1], box[1, 0]])
    p = torch.from_numpy(v.copy()).double()
    assert bool(((p >= lo - 1e-5) & (p <= hi + 1e-5)).all())
    assert int(f.min()) >= 0 and int(f.max()) < v.shape[0]
