"""Camera rays and batch gathers on the CPU (test infrastructure only).

Restates dataLoader/ray_utils.py:27-90 (get_ray_directions,
get_ray_directions_blender, get_rays), the blender loader's ray build
(dataLoader/blender.py:54-55, :86-96: directions divided by their norm, then
rotated per camera, all cameras concatenated) and the training batch gather
(train_per_scene.py:151-152), with the same ATen operations.  kornia's
create_meshgrid(H, W, normalized_coordinates=False) is the (x, y) grid of
integer pixel coordinates; it is written here with torch.meshgrid."""
import torch


def _grid(H, W):
    """create_meshgrid(H, W, normalized_coordinates=False)[0] + 0.5, unbound:
    the column and row coordinates of every pixel centre, [H, W] each."""
    ys, xs = torch.meshgrid(torch.linspace(0, H - 1, H), torch.linspace(0, W - 1, W), indexing="ij")
    return xs + 0.5, ys + 0.5


def get_ray_directions(H, W, focal, center=None):
    """ray_utils.py:27-45."""
    i, j = _grid(H, W)
    cent = center if center is not None else [W / 2, H / 2]
    return torch.stack([(i - cent[0]) / focal[0], (j - cent[1]) / focal[1], torch.ones_like(i)], -1)


def get_ray_directions_blender(H, W, focal, center=None):
    """ray_utils.py:48-66."""
    i, j = _grid(H, W)
    cent = center if center is not None else [W / 2, H / 2]
    return torch.stack([(i - cent[0]) / focal[0], -(j - cent[1]) / focal[1], -torch.ones_like(i)], -1)


def get_rays(directions, c2w):
    """ray_utils.py:69-90."""
    rays_d = directions @ c2w[:3, :3].T
    rays_o = c2w[:3, 3].expand(rays_d.shape)
    return rays_o.reshape(-1, 3), rays_d.reshape(-1, 3)


def all_rays(c2ws, H, W, focal, center=None, blender=False, normalize=True):
    """The blender loader's all_rays: per camera, get_rays of the (normalised)
    directions, concatenated [n_cams * H * W, 6] (blender.py:54-55, :86-96)."""
    dirs = (get_ray_directions_blender if blender else get_ray_directions)(H, W, focal, center)
    if normalize:
        dirs = dirs / torch.norm(dirs, dim=-1, keepdim=True)
    out = []
    for c2w in c2ws:
        o, d = get_rays(dirs, c2w)
        out.append(torch.cat([o, d], 1))
    return torch.cat(out, 0)


def gather_batch(allrays, allrgbs, ray_idx):
    """train_per_scene.py:151-152."""
    return allrays[ray_idx], allrgbs[ray_idx]
