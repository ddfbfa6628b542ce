"""Evaluation of a signed-distance field (scripts/sdf_regression.ipynb, cell 1:
eval_sdf, cal_l1_iou, the marching-cubes export and marchcude_to_world).

The reference evaluates get_coding + linear_mat in chunks and copies every
chunk to the host; here each chunk is one call of the fused field kernel and
the results stay on the model's device.  Chunks only bound the memory of one
call: every point is evaluated independently, so the chunk size does not
change any value."""
import torch

from .mesh import marching_tetrahedra, write_ply


def _device(model):
    return torch.device(model.device)


@torch.no_grad()
def eval_sdf(model, reso, bbox, chunk=1 << 20):
    """The field on a regular grid: returns [reso[0], reso[1], reso[2]] with
    entry (i, j, k) the field at (x_i, y_j, z_k), x = linspace(0, bbox[0],
    reso[0]) and likewise for y and z (bbox a number gives the same extent on
    every axis).  The reference fills the same (x, y, z) order for cubic
    grids, the only ones it calls this with."""
    dev = _device(model)
    ext = [float(bbox)] * 3 if isinstance(bbox, (int, float)) else [float(b) for b in bbox]
    axes = [torch.linspace(0, ext[a], int(reso[a]), device=dev) for a in range(3)]
    out = torch.empty(int(reso[0]) * int(reso[1]) * int(reso[2]), device=dev)
    n_yz = int(reso[1]) * int(reso[2])
    # x-slabs of whole (y, z) planes, as many as fit one chunk
    slab = max(1, chunk // n_yz)
    yy, zz = torch.meshgrid(axes[1], axes[2], indexing="ij")
    yz = torch.stack([yy.reshape(-1), zz.reshape(-1)], -1)
    for i0 in range(0, int(reso[0]), slab):
        xs = axes[0][i0:i0 + slab]
        pts = torch.cat([xs.repeat_interleave(n_yz)[:, None], yz.repeat(len(xs), 1)], 1)
        out[i0 * n_yz:(i0 + len(xs)) * n_yz] = model.field(pts.contiguous()).reshape(-1)
    return out.reshape(int(reso[0]), int(reso[1]), int(reso[2]))


@torch.no_grad()
def cal_l1_iou(model, sdf, coordinates, chunk=1 << 20):
    """Mean absolute error of the field against `sdf` at `coordinates`, and the
    IoU of the two outsides: |{sdf > 0} and {pred > 0}| / |{sdf > 0} or
    {pred > 0}|.  Returns (l1, iou) as 0-d tensors on the host."""
    dev = _device(model)
    pred = torch.cat([model.field(c.to(dev).contiguous()) for c in torch.split(coordinates, chunk)])
    s = sdf.to(dev).reshape(pred.shape)
    l1 = (pred - s).abs().mean()
    iou = ((s > 0) & (pred > 0)).sum() / ((s > 0) | (pred > 0)).sum()
    return l1.cpu(), iou.cpu()


def marchcude_to_world(vertices, reso_WHD):
    """Grid index coordinates to [-1, 1]: v / (reso - 1) * 2 - 1."""
    r = torch.as_tensor(reso_WHD, dtype=torch.float64, device=vertices.device)
    return vertices.to(torch.float64) / (r - 1) * 2 - 1


@torch.no_grad()
def export_sdf_mesh(model, path, reso, bbox, level=0.0, chunk=1 << 20):
    """The zero level set of the field on the grid of eval_sdf, vertices in
    [-1, 1] (marchcude_to_world), written as a PLY file.  Faces point from
    negative to positive distance (out of the shape).  Returns (verts, faces)."""
    vol = eval_sdf(model, reso, bbox, chunk=chunk)
    verts, faces = marching_tetrahedra(vol, level, inside_above=False)
    verts = marchcude_to_world(verts, reso)
    write_ply(path, verts.cpu().numpy(), faces.cpu().numpy())
    return verts, faces
