"""Isosurface extraction and PLY export of a field (train_per_scene.py:32-41,
utils.py:178-249).

The reference runs skimage's marching cubes on the dense alpha volume and
writes the mesh with plyfile; neither library is available here.  This module
extracts the same isosurface with marching tetrahedra on the GPU
(ff_mt_count / ff_mt_emit / ff_mt_verts): every grid cube is split into six
tetrahedra along its main diagonal (the split is the same in neighbouring
cubes, so the mesh is closed), each tetrahedron whose corners straddle the
level contributes one or two triangles, and every vertex is the linear
interpolation of the level on a grid edge, shared by all the triangles that
use that edge.  The triangulation differs from marching cubes'; the surface,
the vertex positions on the grid edges and the file layout are the
reference's."""
import ctypes

import torch

from . import _lib as L


def marching_tetrahedra(vol, level, inside_above=True):
    """Isosurface of a [A, B, C] volume (on the GPU; taken as FP32) at
    `level`.  Corners with value above the level are inside when
    inside_above (density), below it otherwise (signed distance).  Returns
    (verts [V, 3] float64 in index coordinates of the volume's axes, sorted
    by the grid edge they lie on, faces [F, 3] int64 in cube order), every
    face oriented with its normal pointing from inside to outside.

    Three launches and torch steps between them: the triangle count of every
    cube, its exclusive prefix sum (torch.cumsum) and the list of cubes with
    triangles (torch.nonzero), the edge keys of every triangle written at its
    cube's offset (one thread per listed cube), the distinct keys and every
    corner's rank among them (torch.unique), and one vertex per distinct
    key."""
    if not vol.is_cuda:
        raise ValueError("marching_tetrahedra runs on the GPU: pass the volume as a CUDA tensor")
    if vol.dim() != 3:
        raise ValueError("expected a [A, B, C] volume")
    v = vol.to(torch.float32).contiguous()
    dev = v.device
    A, B, C = v.shape
    empty = (torch.zeros((0, 3), dtype=torch.float64, device=dev), torch.zeros((0, 3), dtype=torch.int64, device=dev))
    if min(A, B, C) < 2:
        return empty
    lib, st, lev = L.lib(), L.stream_of(v), ctypes.c_double(float(level))
    counts = torch.empty((A - 1) * (B - 1) * (C - 1), dtype=torch.int32, device=dev)
    L.check(lib.ff_mt_count(L.ptr(v), A, B, C, lev, int(bool(inside_above)), L.ptr(counts), st), "ff_mt_count")
    ends = torch.cumsum(counts, 0, dtype=torch.int64)
    n_tri = int(ends[-1])
    if n_tri == 0:
        return empty
    offsets = ends - counts
    cubes = torch.nonzero(counts).squeeze(1)                # cubes with triangles
    keys = torch.empty(3 * n_tri, dtype=torch.int64, device=dev)
    L.check(lib.ff_mt_emit(L.ptr(v), A, B, C, lev, int(bool(inside_above)), L.ptr(cubes), cubes.numel(),
                           L.ptr(offsets), L.ptr(keys), st), "ff_mt_emit")
    uniq, inv = torch.unique(keys, return_inverse=True)
    verts = torch.empty((uniq.numel(), 3), dtype=torch.float64, device=dev)
    L.check(lib.ff_mt_verts(L.ptr(v), A, B, C, lev, L.ptr(uniq), uniq.numel(), L.ptr(verts), st), "ff_mt_verts")
    return verts, inv.reshape(-1, 3)


def write_ply(path, verts, faces):
    """Binary little-endian PLY with a float vertex (x, y, z) element and a face
    element of int vertex_indices lists, the layout plyfile writes for the
    reference (utils.py:233-249)."""
    import numpy as np
    verts = np.ascontiguousarray(verts, dtype="<f4")
    faces = np.ascontiguousarray(faces, dtype="<i4")
    header = ("ply\nformat binary_little_endian 1.0\n"
              f"element vertex {verts.shape[0]}\nproperty float x\nproperty float y\nproperty float z\n"
              f"element face {faces.shape[0]}\nproperty list uchar int vertex_indices\nend_header\n")
    rec = np.empty(faces.shape[0], dtype=[("n", "u1"), ("i", "<i4", (3,))])
    rec["n"] = 3
    rec["i"] = faces
    with open(path, "wb") as fh:
        fh.write(header.encode("ascii"))
        fh.write(verts.tobytes())
        fh.write(rec.tobytes())


def read_ply(path):
    """Inverse of write_ply (for checks): returns (verts [V, 3], faces [F, 3])."""
    import numpy as np
    with open(path, "rb") as fh:
        data = fh.read()
    end = data.index(b"end_header\n") + len(b"end_header\n")
    head = data[:end].decode("ascii").split("\n")
    nv = int(next(h for h in head if h.startswith("element vertex")).split()[-1])
    nf = int(next(h for h in head if h.startswith("element face")).split()[-1])
    verts = np.frombuffer(data, dtype="<f4", count=3 * nv, offset=end).reshape(nv, 3)
    rec = np.frombuffer(data, dtype=[("n", "u1"), ("i", "<i4", (3,))], count=nf, offset=end + 12 * nv)
    return verts, rec["i"]


@torch.no_grad()
def export_mesh(model, path, gridSize=None, level=0.005, times=16):
    """The reference's mesh export (train_per_scene.py:32-41, :83-86): the
    dense alpha volume of getDenseAlpha, its isosurface at `level`, vertices
    scaled by 1 / (shape - 1) (utils.py:178-179) and placed in the box in the
    volume's (z, y, x) axis order as utils.py:215-218 does.  Faces point out of
    the region where alpha exceeds the level.  Returns (verts, faces)."""
    alpha, _ = model.getDenseAlpha(gridSize, times=times)                    # [Z, Y, X]
    verts, faces = marching_tetrahedra(alpha, level, inside_above=True)
    shape = torch.tensor(alpha.shape, dtype=torch.float64, device=verts.device)
    unit = verts / (shape - 1)
    box = model.aabb.to(torch.float64).to(verts.device)
    lo = torch.stack([box[0, 2], box[0, 1], box[0, 0]])
    size = torch.stack([box[1, 2] - box[0, 2], box[1, 1] - box[0, 1], box[1, 0] - box[0, 0]])
    points = lo + unit * size
    write_ply(path, points.cpu().numpy(), faces.cpu().numpy())
    return points, faces
