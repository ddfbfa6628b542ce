"""Marching tetrahedra on the device (ff_mt_count / ff_mt_emit / ff_mt_verts)
against the oracle's on the same FP32 volumes: the same vertices bit for bit
(both sorted by grid edge) and the same set of oriented faces; the sphere's
closed forms on a volume the oracle would take long on; export_mesh on a
render-path model: the dense alpha volume from the HIP kernels, its
isosurface, and a PLY file that reads back with every vertex in the box (in
the (z, y, x) axis order of the reference's export)."""
import math
import sys

import pytest
import torch

from ffields.mesh import export_mesh, marching_tetrahedra, read_ply
from oracle.mesh import marching_tetrahedra as oracle_mt
from test_gpu_render import make_model

pytestmark = pytest.mark.gpu


def _sphere(shape, r, centre=(0.0, 0.0, 0.0)):
    g = [torch.linspace(-1, 1, n, dtype=torch.float64) for n in shape]
    z, y, x = torch.meshgrid(*g, indexing="ij")
    return ((x - centre[0]) ** 2 + (y - centre[1]) ** 2 + (z - centre[2]) ** 2).sqrt() - r


def _rows(faces):
    """Faces as a sorted list of rows (the face order is not part of the result)."""
    return sorted(map(tuple, faces.cpu().tolist()))


def _compare(vol32, level, inside_above):
    v_ref, f_ref = oracle_mt(vol32.double(), level, inside_above)
    v_dev, f_dev = marching_tetrahedra(vol32.cuda(), level, inside_above)
    assert v_dev.dtype == torch.float64 and f_dev.dtype == torch.int64
    assert torch.equal(v_dev.cpu(), v_ref), (v_dev.shape, v_ref.shape)
    assert _rows(f_dev) == _rows(f_ref)
    return v_dev, f_dev


@pytest.mark.parametrize("case", ["sdf", "density", "noise", "ties", "flat", "ragged"])
def test_matches_oracle(case):
    print(f"[case] marching tetrahedra vs oracle: {case}", file=sys.stderr, flush=True)
    g = torch.Generator().manual_seed(11)
    if case == "sdf":
        vol, level, above = _sphere((20, 23, 17), 0.55, (0.1, -0.05, 0.0)).float(), 0.0, False
    elif case == "density":
        vol, level, above = (-_sphere((18, 16, 21), 0.5)).float(), 0.02, True
    elif case == "noise":            # every tetrahedron case, many two-triangle ones
        vol, level, above = torch.rand((13, 11, 9), generator=g), 0.5, True
    elif case == "ties":             # corners exactly at the level (outside; vertices on grid points)
        vol, level, above = torch.randint(0, 3, (9, 10, 8), generator=g).float() * 0.25, 0.25, True
    elif case == "flat":             # one plane crossing, planar mesh
        vol, level, above = torch.arange(6.0)[:, None, None].expand(6, 5, 4).contiguous(), 2.5, False
    else:                            # two points on one axis; a non-integer level
        vol, level, above = torch.rand((2, 7, 3), generator=g), 0.37, False
    v, f = _compare(vol, level, above)
    assert f.shape[0] > 0


def test_empty_and_degenerate():
    print("[case] marching tetrahedra: empty volumes", file=sys.stderr, flush=True)
    for vol in (torch.zeros((5, 6, 7)), torch.ones((3, 3, 3)), torch.zeros((1, 4, 4)), torch.zeros((4, 1, 9))):
        v, f = marching_tetrahedra(vol.cuda(), 0.5)
        assert v.shape == (0, 3) and f.shape == (0, 3) and v.is_cuda
    with pytest.raises(ValueError):
        marching_tetrahedra(torch.zeros((4, 4, 4)), 0.5)          # a host volume is refused


def test_sphere_closed_forms_large():
    """A 160^3 signed distance: the oracle's mesh, vertices on the sphere,
    closed and consistently oriented, area and outward volume within 0.1 % of
    the closed forms (the oracle's own errors there: -0.008 %, -0.016 %)."""
    print("[case] marching tetrahedra: 160^3 sphere", file=sys.stderr, flush=True)
    n, r = 160, 0.7
    h = 2.0 / (n - 1)
    verts, faces = _compare(_sphere((n, n, n), r).float(), 0.0, False)
    p = verts * h - 1.0
    assert (p.norm(dim=1) - r).abs().max().item() < 0.02 * h
    e = torch.cat([faces[:, [0, 1]], faces[:, [1, 2]], faces[:, [2, 0]]])
    key = e[:, 0] * verts.shape[0] + e[:, 1]
    rkey = e[:, 1] * verts.shape[0] + e[:, 0]
    assert torch.unique(key).numel() == key.numel()
    assert torch.equal(torch.sort(key).values, torch.sort(rkey).values)
    a, b, c = p[faces[:, 0]], p[faces[:, 1]], p[faces[:, 2]]
    area = 0.5 * torch.cross(b - a, c - a, dim=1).norm(dim=1).sum().item()
    vol6 = (a * torch.cross(b, c, dim=1)).sum().item()
    assert abs(area / (4 * math.pi * r * r) - 1) < 1e-3, area
    assert abs(vol6 / 6 / (4 / 3 * math.pi * r ** 3) - 1) < 1e-3, vol6


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


def test_emit_without_cube_list():
    """ff_mt_emit with no cube list (one thread per cube of the grid) writes
    the same keys at the same offsets as with the list of cubes that have
    triangles."""
    import ctypes
    import ffields._lib as L
    print("[case] marching tetrahedra: emit over every cube", file=sys.stderr, flush=True)
    vol = torch.rand((11, 9, 13), generator=torch.Generator().manual_seed(4)).cuda()
    A, B, C = vol.shape
    lev = ctypes.c_double(0.5)
    counts = torch.empty((A - 1) * (B - 1) * (C - 1), dtype=torch.int32, device="cuda")
    L.check(L.lib().ff_mt_count(L.ptr(vol), A, B, C, lev, 1, L.ptr(counts), L.stream_of(vol)))
    ends = torch.cumsum(counts, 0, dtype=torch.int64)
    offsets = ends - counts
    n_tri = int(ends[-1])
    cubes = torch.nonzero(counts).squeeze(1)
    k_all = torch.empty(3 * n_tri, dtype=torch.int64, device="cuda")
    k_list = torch.empty(3 * n_tri, dtype=torch.int64, device="cuda")
    L.check(L.lib().ff_mt_emit(L.ptr(vol), A, B, C, lev, 1, None, 0, L.ptr(offsets), L.ptr(k_all),
                               L.stream_of(vol)))
    L.check(L.lib().ff_mt_emit(L.ptr(vol), A, B, C, lev, 1, L.ptr(cubes), cubes.numel(), L.ptr(offsets),
                               L.ptr(k_list), L.stream_of(vol)))
    assert n_tri > 0 and torch.equal(k_all, k_list)


def test_export_mesh_matches_oracle(tmp_path):
    """export_mesh's points and faces equal the oracle's marching tetrahedra
    of the same dense alpha volume, mapped into the box the same way
    (utils.py:215-218 axis order, 1 / (shape - 1) scaling)."""
    print("[case] export mesh vs oracle", file=sys.stderr, flush=True)
    m = make_model(seed=9)
    torch.manual_seed(3)
    alpha, _ = m.getDenseAlpha((24, 20, 16), times=1)
    level = float(torch.quantile(alpha.flatten().float(), 0.7))
    pts, faces = export_mesh(m, str(tmp_path / "f.ply"), gridSize=(24, 20, 16), level=level, times=1)
    v_ref, f_ref = oracle_mt(alpha.cpu().double(), level, True)
    shape = torch.tensor(alpha.shape, dtype=torch.float64)
    box = m.aabb.cpu().double()
    lo = torch.stack([box[0, 2], box[0, 1], box[0, 0]])
    size = torch.stack([box[1, 2] - box[0, 2], box[1, 1] - box[0, 1], box[1, 0] - box[0, 0]])
    want = lo + v_ref / (shape - 1) * size
    assert torch.equal(pts.cpu(), want)
    assert _rows(faces) == _rows(f_ref)

