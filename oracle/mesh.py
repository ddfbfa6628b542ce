"""Oracle marching tetrahedra for the mesh export (utils.py:183-249,
train_per_scene.py:32-41).  TEST INFRASTRUCTURE ONLY (see the package
docstring).

The reference extracts the isosurface with skimage's marching cubes, which is
not installed; the product replaces it by marching tetrahedra and this is the
CPU restatement of that replacement in torch operations, tetrahedron by
tetrahedron rather than cube by cube: every grid cube splits into the six
tetrahedra that share its main diagonal, each tetrahedron whose corners
straddle the level contributes one or two triangles, every vertex is the
linear interpolation of the level on a grid edge (FP64, from the edge's lower
point id), shared by all triangles using that edge, and every triangle is
oriented along its tetrahedron's inside -> outside direction.  Vertices come
sorted by edge key (lo * N + hi); the face order is the order of the case
groups, so comparisons with the device path compare face sets.

Pinned by a sphere's signed distance (vertices on the sphere, closed and
consistently oriented mesh, area and enclosed volume against the closed
forms) and by the density orientation and empty-volume cases
(tests/test_mesh.py)."""
import itertools

import torch

# corners of a unit cube, corner c at offset (c & 1, (c >> 1) & 1, c >> 2) on
# the volume's axes (0, 1, 2)
_OFFS = [((c >> 0) & 1, (c >> 1) & 1, (c >> 2) & 1) for c in range(8)]


def _kuhn_tets():
    """The six tetrahedra of a cube that share its main diagonal: for every
    order (p, q, r) of the axes, corners 0, e_p, e_p + e_q, 7."""
    tets = []
    for p, q, r in itertools.permutations(range(3)):
        a = [0, 0, 0]
        path = [tuple(a)]
        for ax in (p, q):
            a[ax] = 1
            path.append(tuple(a))
        path.append((1, 1, 1))
        tets.append([_OFFS.index(v) for v in path])
    return tets


def _case_table():
    """For each inside mask of a tetrahedron's 4 corners: up to two triangles,
    each as 3 edges (u, v) between an inside corner u and an outside corner v."""
    table = []
    for m in range(16):
        ins = [i for i in range(4) if (m >> i) & 1]
        out = [i for i in range(4) if not (m >> i) & 1]
        tris = []
        if len(ins) == 1:
            tris.append([(ins[0], v) for v in out])
        elif len(ins) == 3:
            tris.append([(u, out[0]) for u in ins])
        elif len(ins) == 2:
            a, b = ins
            c, d = out
            tris.append([(a, c), (a, d), (b, d)])
            tris.append([(a, c), (b, d), (b, c)])
        table.append(tris)
    return table


_TETS = _kuhn_tets()
_CASES = _case_table()


def marching_tetrahedra(vol, level, inside_above=True):
    """Isosurface of a [A, B, C] volume at `level`.  Corners with value above
    the level are inside when inside_above (density), below it otherwise
    (signed distance).  Returns (verts [V, 3] float64 in index coordinates of
    the volume's axes, faces [F, 3] int64), every face oriented with its
    normal pointing from inside to outside."""
    v = vol.to(torch.float64)
    dev = v.device
    A, B, C = v.shape
    if min(A, B, C) < 2:
        return torch.zeros((0, 3), dtype=torch.float64, device=dev), torch.zeros((0, 3), dtype=torch.int64, device=dev)
    corner = torch.stack([v[o[0]:A - 1 + o[0], o[1]:B - 1 + o[1], o[2]:C - 1 + o[2]] for o in _OFFS])  # [8, ...]
    ins = corner > level if inside_above else corner < level
    active = ins.any(0) & (~ins).any(0)
    cube = active.nonzero()                                                   # [M, 3] cube origins
    if cube.shape[0] == 0:
        return torch.zeros((0, 3), dtype=torch.float64, device=dev), torch.zeros((0, 3), dtype=torch.int64, device=dev)
    offs = torch.tensor(_OFFS, device=dev)                                    # [8, 3]
    tets = torch.tensor(_TETS, device=dev)                                    # [6, 4]
    pos = cube[:, None, None, :] + offs[tets][None]                           # [M, 6, 4, 3] grid points
    val = v[pos[..., 0], pos[..., 1], pos[..., 2]]                            # [M, 6, 4]
    tin = val > level if inside_above else val < level
    code = (tin.long() << torch.arange(4, device=dev)).sum(-1)                # [M, 6]
    pos, val, tin, code = pos.reshape(-1, 4, 3), val.reshape(-1, 4), tin.reshape(-1, 4), code.reshape(-1)
    pid = (pos[..., 0] * B + pos[..., 1]) * C + pos[..., 2]                   # [T, 4] grid point ids
    faces, keys, ends = [], [], []
    for m in range(1, 15):
        sel = (code == m).nonzero().squeeze(1)
        if sel.numel() == 0:
            continue
        for tri in _CASES[m]:
            e = torch.tensor(tri, device=dev)                                 # [3, 2] (inside, outside) corners
            pu, pv = pid[sel][:, e[:, 0]], pid[sel][:, e[:, 1]]               # [S, 3]
            keys.append(torch.minimum(pu, pv) * (A * B * C) + torch.maximum(pu, pv))
            ends.append(torch.stack([torch.minimum(pu, pv), torch.maximum(pu, pv)], -1))
            # orientation: from the tetrahedron's inside corners to its outside ones
            w = tin[sel].double()
            cin = (pos[sel].double() * w[..., None]).sum(1) / w.sum(1, keepdim=True)
            cout = (pos[sel].double() * (1 - w)[..., None]).sum(1) / (1 - w).sum(1, keepdim=True)
            faces.append((sel, e, cout - cin))
    key = torch.cat(keys).reshape(-1)
    uniq, inv = torch.unique(key, return_inverse=True)
    end = torch.cat(ends).reshape(-1, 2)
    first = torch.full((uniq.numel(),), -1, dtype=torch.int64, device=dev)
    first.scatter_(0, inv, torch.arange(inv.numel(), device=dev))
    e0, e1 = end[first, 0], end[first, 1]                                     # endpoints, lower id first
    flat = v.reshape(-1)
    f0, f1 = flat[e0], flat[e1]

    def coords(p):
        return torch.stack([p // (B * C), (p // C) % B, p % C], -1).double()

    t = ((level - f0) / (f1 - f0))[:, None]
    verts = coords(e0) + t * (coords(e1) - coords(e0))
    tri = inv.reshape(-1, 3)
    # orient every face along its tetrahedron's inside -> outside direction
    dirs = torch.cat([d for _, _, d in faces])
    p0, p1, p2 = verts[tri[:, 0]], verts[tri[:, 1]], verts[tri[:, 2]]
    flip = (torch.cross(p1 - p0, p2 - p0, dim=1) * dirs).sum(1) < 0
    tri = torch.where(flip[:, None], tri[:, [0, 2, 1]], tri)
    return verts, tri
