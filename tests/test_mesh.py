"""Isosurface extraction and PLY export (train_per_scene.py:32-41,
utils.py:178-249), on the CPU: the oracle's marching tetrahedra on a sphere's
signed distance give a closed, consistently oriented mesh whose vertices lie
on the sphere and whose area and enclosed volume match the closed forms; the
PLY file round-trips.  The device path is compared with this oracle in
tests/test_gpu_mesh.py."""
import math

import torch

from ffields.mesh import read_ply, write_ply
from oracle.mesh import marching_tetrahedra


def sphere(n=40, r=0.6):
    g = torch.linspace(-1, 1, n, dtype=torch.float64)
    z, y, x = torch.meshgrid(g, g, g, indexing="ij")
    return (x * x + y * y + z * z).sqrt() - r, 2.0 / (n - 1)


def test_sphere_sdf_mesh():
    vol, h = sphere()
    r = 0.6
    verts, faces = marching_tetrahedra(vol, 0.0, inside_above=False)
    assert faces.shape[0] > 1000
    p = verts * h - 1.0                                        # world coordinates
    # vertices on the sphere (linear interpolation of the distance on grid edges)
    assert (p.norm(dim=1) - r).abs().max().item() < 0.1 * h
    # closed and consistently oriented: every directed edge once, with its reverse once
    e = torch.cat([faces[:, [0, 1]], faces[:, [1, 2]], faces[:, [2, 0]]])
    key = e[:, 0] * verts.shape[0] + e[:, 1]
    rkey = e[:, 1] * verts.shape[0] + e[:, 0]
    assert torch.unique(key).numel() == key.numel()
    assert torch.equal(torch.sort(key).values, torch.sort(rkey).values)
    # area and (outward) enclosed volume
    a, b, c = p[faces[:, 0]], p[faces[:, 1]], p[faces[:, 2]]
    cr = torch.cross(b - a, c - a, dim=1)
    area = 0.5 * cr.norm(dim=1).sum().item()
    vol6 = (a * torch.cross(b, c, dim=1)).sum().item()
    assert abs(area / (4 * math.pi * r * r) - 1) < 0.02, area
    assert abs(vol6 / 6 / (4 / 3 * math.pi * r ** 3) - 1) < 0.02, vol6


def test_density_orientation_and_empty():
    vol, h = sphere(24, 0.5)
    dens = -vol                                                 # inside where density > level
    v1, f1 = marching_tetrahedra(dens, 0.0, inside_above=True)
    p = v1 * h - 1.0
    a, b, c = p[f1[:, 0]], p[f1[:, 1]], p[f1[:, 2]]
    assert (a * torch.cross(b, c, dim=1)).sum().item() > 0     # normals point out of the dense region
    v0, f0 = marching_tetrahedra(torch.zeros((5, 6, 7)), 0.5)
    assert v0.shape == (0, 3) and f0.shape == (0, 3)


def test_ply_round_trip(tmp_path):
    vol, _ = sphere(16, 0.5)
    verts, faces = marching_tetrahedra(vol, 0.0, inside_above=False)
    path = tmp_path / "m.ply"
    write_ply(path, verts.numpy(), faces.numpy())
    v, f = read_ply(path)
    assert torch.allclose(torch.from_numpy(v.copy()).double(), verts, atol=1e-5)
    assert torch.equal(torch.from_numpy(f.copy()).long(), faces)
    head = open(path, "rb").read(300).split(b"end_header")[0].decode()
    assert "property list uchar int vertex_indices" in head and "binary_little_endian" in head
