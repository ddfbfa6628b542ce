"""Camera rays and training-batch sampling on the device
(dataLoader/ray_utils.py:27-90, train_per_scene.py:16-28).

get_ray_directions / get_ray_directions_blender give the camera-space
direction through every pixel centre; get_rays rotates them into world space
with a camera-to-world matrix (the reference's own functions, kept for code
that calls them).  camera_rays does the whole chain -- directions, the
blender loader's normalisation, rotation -- in one HIP kernel for any set of
cameras and pixels.  SimpleSampler walks a fresh random permutation of the
training rays in fixed-size batches, as the reference does, but keeps the
permutation on the device; gather_batch takes a batch's rays and colours with
one HIP kernel."""
import ctypes

import torch

from . import _lib as L


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


def camera_rays(c2w, H, W, focal, center=None, blender=False, normalize=True, view=None, pix=None):
    """World-space rays [n, 6] = (origin, direction) of pinhole cameras,
    computed by ff_camera_rays: the direction through pixel centre
    (x + 0.5, y + 0.5) -- get_ray_directions (ray_utils.py:27-45), or
    get_ray_directions_blender (:48-66) with blender=True -- divided by its
    norm when normalize (blender.py:54-55), rotated by c2w[:3, :3] with
    origin c2w[:3, 3] (get_rays, :69-90).

    c2w: [n_cams, 3, 4] (or [3, 4]) on the GPU; focal = (fx, fy) or a scalar;
    center defaults to (W / 2, H / 2).  view, pix: int64 device tensors of
    camera and row-major pixel indices; both None gives every pixel of every
    camera in order (the blender loader's all_rays)."""
    c2w = c2w.reshape(-1, 3, 4).to(torch.float32).contiguous()
    if not c2w.is_cuda:
        raise ValueError("camera_rays runs on the GPU: pass c2w as a CUDA tensor")
    fx, fy = (focal, focal) if isinstance(focal, (int, float)) else (float(focal[0]), float(focal[1]))
    cx, cy = center if center is not None else (W / 2, H / 2)
    cams = L.Cameras(L.ptr(c2w), c2w.shape[0], int(H), int(W), fx, fy, float(cx), float(cy), int(blender),
                     int(normalize))
    if (view is None) != (pix is None):
        raise ValueError("give both view and pix, or neither")
    if view is None:
        n = c2w.shape[0] * H * W
    else:
        view = view.to(device=c2w.device, dtype=torch.int64).contiguous()
        pix = pix.to(device=c2w.device, dtype=torch.int64).contiguous()
        if view.shape != pix.shape:
            raise ValueError("view and pix differ in shape")
        n = view.numel()
    out = torch.empty((n, 6), device=c2w.device)
    L.check(L.lib().ff_camera_rays(ctypes.byref(cams), L.ptr(view), L.ptr(pix), n, L.ptr(out), L.stream_of(out)),
            "ff_camera_rays")
    return out


def gather_batch(allrays, allrgbs, ids):
    """(allrays[ids], allrgbs[ids]) (train_per_scene.py:151-152) by
    ff_gather_rays; allrgbs may be None.  Device tensors [N, 6], [N, 3] and
    int64 ids; an index outside [0, N) gives a NaN row."""
    if not allrays.is_cuda:
        raise ValueError("gather_batch runs on the GPU: pass CUDA tensors")
    rays = allrays.to(torch.float32).contiguous()
    rgbs = None if allrgbs is None else allrgbs.to(device=rays.device, dtype=torch.float32).contiguous()
    if rays.dim() != 2 or rays.shape[1] != 6 or (rgbs is not None and tuple(rgbs.shape) != (rays.shape[0], 3)):
        raise ValueError("expected allrays [N, 6] and allrgbs [N, 3]")
    ids = ids.to(device=rays.device, dtype=torch.int64).contiguous()
    B = ids.numel()
    out_rays = torch.empty((B, 6), device=rays.device)
    out_rgbs = None if rgbs is None else torch.empty((B, 3), device=rays.device)
    L.check(L.lib().ff_gather_rays(L.ptr(rays), L.ptr(rgbs), rays.shape[0], L.ptr(ids), B, L.ptr(out_rays),
                                   L.ptr(out_rgbs), L.stream_of(out_rays)), "ff_gather_rays")
    return out_rays, out_rgbs
