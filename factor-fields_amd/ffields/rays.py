"""Camera rays and training-batch sampling on the device
(dataLoader/ray_utils.py:27-90, train_per_scene.py:16-28).

get_ray_directions / get_ray_directions_blender give the camera-space
direction through every pixel centre; get_rays rotates them into world space
with a camera-to-world matrix.  SimpleSampler walks a fresh random permutation
of the training rays in fixed-size batches, as the reference does, but keeps
the permutation on the device so a batch needs no host-to-device copy."""
import torch


def _pixel_centres(H, W, device):
    """Column and row coordinates (x + 0.5, y + 0.5) of every pixel, [H, W] each."""
    y, x = torch.meshgrid(torch.arange(H, dtype=torch.float32, device=device),
                          torch.arange(W, dtype=torch.float32, device=device), indexing="ij")
    return x + 0.5, y + 0.5


def get_ray_directions(H, W, focal, center=None, device="cpu"):
    """[H, W, 3] directions ((x - cx) / fx, (y - cy) / fy, 1) through the pixel
    centres (ray_utils.py:27-45); focal = (fx, fy), center defaults to (W/2, H/2)."""
    i, j = _pixel_centres(H, W, device)
    cx, cy = center if center is not None else (W / 2, H / 2)
    return torch.stack([(i - cx) / focal[0], (j - cy) / focal[1], torch.ones_like(i)], -1)


def get_ray_directions_blender(H, W, focal, center=None, device="cpu"):
    """As get_ray_directions with the y and z axes flipped (ray_utils.py:48-66):
    ((x - cx) / fx, -(y - cy) / fy, -1)."""
    i, j = _pixel_centres(H, W, device)
    cx, cy = center if center is not None else (W / 2, H / 2)
    return torch.stack([(i - cx) / focal[0], -(j - cy) / focal[1], -torch.ones_like(i)], -1)


def get_rays(directions, c2w):
    """World-space rays of one image (ray_utils.py:69-90): rays_d = directions
    rotated by c2w[:3, :3] (not normalised, as in the reference), rays_o the
    camera centre c2w[:3, 3] for every pixel.  Returns ([H*W, 3], [H*W, 3])."""
    rays_d = directions @ c2w[:3, :3].T
    rays_o = c2w[:3, 3].expand(rays_d.shape)
    return rays_o.reshape(-1, 3), rays_d.reshape(-1, 3)


class SimpleSampler:
    """Batches of `batch` ray indices out of `total` (train_per_scene.py:16-28):
    consecutive slices of a random permutation, redrawn when the next slice
    would run past the end.  The permutation lives on `device`."""

    def __init__(self, total, batch, device="cpu", generator=None):
        self.total = total
        self.batch = batch
        self.curr = total
        self.ids = None
        self.device = torch.device(device)
        self.generator = generator

    def nextids(self):
        self.curr += self.batch
        if self.curr + self.batch > self.total:
            self.ids = torch.randperm(self.total, device=self.device, generator=self.generator)
            self.curr = 0
        return self.ids[self.curr:self.curr + self.batch]
