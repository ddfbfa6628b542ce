"""Volume rendering on the HIP kernels: FactorFields.forward
(models/FactorFields.py:843-898) for bounded scenes, with the sampler
(:586-602), AlphaGridMask occupancy (:91-110), field + density/SH head,
compositing (:82-88, :878-896) and their backward passes.

The colour head is degree-2 spherical harmonics (models/sh.py:34-64) on
decoder outputs 1..27 followed by a sigmoid; it is evaluated inside the field
kernel, so ``renderModule`` holds no parameters.
"""

import ctypes

import numpy as np
import torch
import torch.nn as nn

from . import _lib as L
from .ops import mlp_desc
from .params import N_to_reso, cal_n_samples


class AlphaGridMask:
    """models/FactorFields.py:91-110: a 0/1 occupancy volume over an aabb.
    ``alpha_volume`` is [Dz, Hy, Wx] (or [1,1,Dz,Hy,Wx]) with values in {0,1}."""

    def __init__(self, device, aabb, alpha_volume):
        self.device = device
        self.aabb = aabb.to(device).float()
        self.aabbSize = self.aabb[1] - self.aabb[0]
        self.invgridSize = 1.0 / self.aabbSize * 2
        vol = alpha_volume.to(device).float()
        self.alpha_volume = vol.view(1, 1, *vol.shape[-3:])
        if not bool(((self.alpha_volume == 0) | (self.alpha_volume == 1)).all()):
            raise ValueError("AlphaGridMask volume must hold only 0 and 1")
        self.gridSize = torch.LongTensor([vol.shape[-1], vol.shape[-2], vol.shape[-3]]).to(device)
        self.occ = self.alpha_volume.reshape(vol.shape[-3:]).to(torch.uint8).contiguous()
        self._inv = self.invgridSize.cpu().tolist()
        self._origin = self.aabb[0].cpu().tolist()


class SHRender(nn.Module):
    """Degree-2 SH colour head (no parameters).  It runs inside the render
    kernels; it is not callable on its own."""

    def forward(self, viewdirs, features):
        raise NotImplementedError("the SH colour head is evaluated inside ff_render_field_fwd")


def _samples_struct(buf):
    s = L.Samples()
    s.xyz = buf["xyz"].data_ptr()
    s.z = buf["z"].data_ptr()
    s.dist = buf["dist"].data_ptr()
    s.ray = buf["ray"].data_ptr()
    s.idx = buf["idx"].data_ptr()
    s.ray_off = buf["ray_off"].data_ptr()
    s.cap = buf["xyz"].shape[0]
    return s


def alloc_samples(B, Ns, device):
    cap = max(1, B * Ns)
    return dict(xyz=torch.empty((cap, 3), device=device), z=torch.empty(cap, device=device),
                dist=torch.empty(cap, device=device),
                ray=torch.empty(cap, device=device, dtype=torch.int32),
                idx=torch.empty(cap, device=device, dtype=torch.int32),
                ray_off=torch.empty(B + 1, device=device, dtype=torch.int32))


class RenderMixin:
    """Render-side methods of FactorFields (reconstruction mode)."""

    def _init_render(self):
        cfg = self.cfg
        if cfg.dataset.get("is_unbound", False):
            raise NotImplementedError("unbounded scenes have no HIP sampler")
        if cfg.renderer.get("shadingMode", "SH") not in ("SH",):
            raise NotImplementedError("only the degree-2 SH colour head is built")
        if cfg.model.out_dim != 28:
            raise ValueError("the SH colour head needs out_dim = 1 + 27")
        self.renderModule = SHRender()
        self.is_unbound = False
        self.inward_aabb = self.aabb
        self.cur_volumeSize = N_to_reso(cfg.training.volume_resoInit ** self.in_dim, self.aabb.cpu())
        self.update_renderParams(self.cur_volumeSize)

    # models/FactorFields.py:693-699
    def update_renderParams(self, gridSize):
        self.aabbSize = self.aabb[1] - self.aabb[0]
        self.gridSize = torch.LongTensor(gridSize).to(self.device)
        units = self.aabbSize / (self.gridSize - 1)
        self.stepSize = torch.mean(units) * self.cfg.renderer.step_ratio
        aabbDiag = torch.sqrt(torch.sum(torch.square(self.aabbSize)))
        self.nSamples = int((aabbDiag / self.stepSize).item()) + 1
        self._step = float(self.stepSize.item())

    def _sampler(self, n_samples):
        s = L.RaySampler()
        a = self.aabb.cpu().tolist()
        for k in range(3):
            s.aabb_min[k], s.aabb_max[k] = a[0][k], a[1][k]
        s.step_size = self._step
        s.n_samples = n_samples
        s.distance_scale = float(self.cfg.renderer.distance_scale)
        am = self.alphaMask
        if am is not None:
            s.occ = am.occ.data_ptr()
            s.occ_reso[0], s.occ_reso[1], s.occ_reso[2] = (am.occ.shape[2], am.occ.shape[1],
                                                           am.occ.shape[0])
            for k in range(3):
                s.occ_origin[k], s.occ_inv[k] = am._origin[k], am._inv[k]
        return s

    def _head(self):
        h = L.RenderHead()
        h.density_act = 0 if self.cfg.renderer.fea2denseAct == "softplus" else 1
        h.density_shift = float(self.cfg.renderer.density_shift)
        return h

    def _field_structs(self):
        tabs = self._read_tables()
        d = self._spec.desc(*self._split(tabs))
        m = mlp_desc(self._layers())
        return d, m

    def render_buffers(self, B, Ns):
        key = (B, Ns)
        cache = getattr(self, "_rbuf", None)
        if cache is None or cache[0] != key:
            dev = self.device
            buf = alloc_samples(B, Ns, dev)
            cap = buf["xyz"].shape[0]
            buf.update(sigma=torch.empty(cap, device=dev), rgb=torch.empty((cap, 3), device=dev),
                       d_sample=torch.empty((cap, 4), device=dev))
            self._rbuf = (key, buf)
        return self._rbuf[1]

    def render_samples(self, rays, n_samples, jitter, buf):
        """Sampler + field + head over `rays`; fills buf (packed samples,
        sigma, rgb).  Returns the ff_samples struct."""
        s = _samples_struct(buf)
        st = L.stream_of(rays)
        smp = self._sampler(n_samples)
        L.check(L.lib().ff_sample_rays(ctypes.byref(smp), L.ptr(rays), L.ptr(jitter), rays.shape[0],
                                       ctypes.byref(s), st), "ff_sample_rays")
        d, m = self._field_structs()
        h = self._head()
        L.check(L.lib().ff_render_field_fwd(ctypes.byref(d), ctypes.byref(m), ctypes.byref(h), L.ptr(rays),
                                            rays.shape[0], ctypes.byref(s), L.ptr(buf["sigma"]),
                                            L.ptr(buf["rgb"]), st), "ff_render_field_fwd")
        return s

    def _grads_struct(self):
        self._ensure_grads()
        b, c = self._tables()
        p = self.linear_mat.params_list()
        g = L.Grads()
        for l, t in enumerate(b):
            g.basis[l] = t.grad.data_ptr()
        for k, t in enumerate(c):
            g.coeff[k] = t.grad.data_ptr()
        g.w0, g.b0, g.w1 = p[0].grad.data_ptr(), p[1].grad.data_ptr(), p[2].grad.data_ptr()
        return g

    def _jitter(self, rays, is_train, jitter):
        if jitter is not None:
            return jitter.contiguous()
        if is_train:
            return torch.rand(rays.shape[0], device=rays.device)
        return None

    # models/FactorFields.py:843-898
    def forward(self, rays_chunk, white_bg=True, is_train=False, ndc_ray=False, N_samples=-1,
                jitter=None):
        if ndc_ray:
            raise NotImplementedError("NDC rays have no HIP sampler")
        n_samples = N_samples if N_samples > 0 else self.nSamples
        rays = rays_chunk.contiguous().float()
        jit = self._jitter(rays, is_train, jitter)
        bg = bool(white_bg or (is_train and torch.rand((1,)) < 0.5))
        params = self.linear_mat.params_list() + list(self._tables()[0]) + list(self._tables()[1])
        rgb_map, depth = RenderRays.apply(self, rays, jit, n_samples, bg, *params)
        coeffs = torch.zeros((1, sum(self.basis_dims)), device=rays.device)
        return rgb_map, depth, coeffs

    def render_train_step(self, rays, target, white_bg=True, N_samples=-1, jitter=None, loss_scale=1.0,
                          loss_sum=None):
        """One NeRF training pass fused across kernels: sample, field, composite,
        MSE loss (train_per_scene.py:158), composite backward, field backward
        into each parameter's .grad.  Returns the device scalar sum((rgb-gt)^2)."""
        n_samples = N_samples if N_samples > 0 else self.nSamples
        rays = rays.contiguous()
        target = target.contiguous()
        B = rays.shape[0]
        jit = self._jitter(rays, True, jitter)
        bg = bool(white_bg or torch.rand((1,)) < 0.5)
        buf = self.render_buffers(B, n_samples)
        s = self.render_samples(rays, n_samples, jit, buf)
        st = L.stream_of(rays)
        rgb_map = buf.setdefault("rgb_map", torch.empty((B, 3), device=rays.device))
        if rgb_map.shape[0] != B:
            rgb_map = buf["rgb_map"] = torch.empty((B, 3), device=rays.device)
        rgb_raw = buf.setdefault("rgb_raw", torch.empty((B, 3), device=rays.device))
        if rgb_raw.shape[0] != B:
            rgb_raw = buf["rgb_raw"] = torch.empty((B, 3), device=rays.device)
        thres = float(self.cfg.renderer.rayMarch_weight_thres)
        L.check(L.lib().ff_composite_fwd(ctypes.byref(s), B, L.ptr(buf["sigma"]), L.ptr(buf["rgb"]),
                                         int(bg), thres, L.ptr(rgb_map), L.ptr(rgb_raw), None, None, None,
                                         st), "ff_composite_fwd")
        if loss_sum is None:
            loss_sum = torch.zeros((), device=rays.device)
        L.check(L.lib().ff_composite_bwd(ctypes.byref(s), B, L.ptr(buf["sigma"]), L.ptr(buf["rgb"]),
                                         int(bg), thres, L.ptr(rgb_raw), None, L.ptr(target),
                                         float(loss_scale), L.ptr(loss_sum), L.ptr(buf["d_sample"]), st),
                "ff_composite_bwd")
        d, m = self._field_structs()
        h = self._head()
        g = self._grads_struct()
        ws_bytes = L.lib().ff_render_field_bwd_workspace_bytes(ctypes.byref(d), ctypes.byref(m), s.cap)
        ws = self._workspace(ws_bytes)
        L.check(L.lib().ff_render_field_bwd(ctypes.byref(d), ctypes.byref(m), ctypes.byref(h), L.ptr(rays),
                                            B, ctypes.byref(s), L.ptr(buf["d_sample"]), ctypes.byref(g),
                                            L.ptr(ws), ws_bytes, st), "ff_render_field_bwd")
        return loss_sum


class RenderRays(torch.autograd.Function):
    """rays -> (rgb_map, depth) with the hand-written backward."""

    @staticmethod
    def forward(ctx, model, rays, jitter, n_samples, white_bg, *params):
        B = rays.shape[0]
        dev = rays.device
        buf = alloc_samples(B, n_samples, dev)
        cap = buf["xyz"].shape[0]
        buf.update(sigma=torch.empty(cap, device=dev), rgb=torch.empty((cap, 3), device=dev))
        s = model.render_samples(rays, n_samples, jitter, buf)
        rgb_map = torch.empty((B, 3), device=dev)
        rgb_raw = torch.empty((B, 3), device=dev)
        depth = torch.empty(B, device=dev)
        thres = float(model.cfg.renderer.rayMarch_weight_thres)
        L.check(L.lib().ff_composite_fwd(ctypes.byref(s), B, L.ptr(buf["sigma"]), L.ptr(buf["rgb"]),
                                         int(white_bg), thres, L.ptr(rgb_map), L.ptr(rgb_raw),
                                         L.ptr(depth), None, None, L.stream_of(rays)), "ff_composite_fwd")
        ctx.model, ctx.white_bg, ctx.thres, ctx.buf = model, white_bg, thres, buf
        ctx.save_for_backward(rays, rgb_raw)
        ctx.mark_non_differentiable(depth)
        return rgb_map, depth

    @staticmethod
    def backward(ctx, d_rgb, d_depth):
        rays, rgb_raw = ctx.saved_tensors
        model, buf = ctx.model, ctx.buf
        B = rays.shape[0]
        s = _samples_struct(buf)
        st = L.stream_of(rays)
        cap = buf["xyz"].shape[0]
        d_sample = torch.empty((cap, 4), device=rays.device)
        L.check(L.lib().ff_composite_bwd(ctypes.byref(s), B, L.ptr(buf["sigma"]), L.ptr(buf["rgb"]),
                                         int(ctx.white_bg), ctx.thres, L.ptr(rgb_raw),
                                         L.ptr(d_rgb.contiguous()), None, 1.0, None, L.ptr(d_sample), st),
                "ff_composite_bwd")
        d, m = model._field_structs()
        h = model._head()
        p = model.linear_mat.params_list()
        b, c = model._tables()
        gp = [torch.zeros_like(t) for t in p]
        gt = [torch.zeros_like(t, memory_format=torch.preserve_format) for t in list(b) + list(c)]
        g = L.Grads()
        for l in range(len(b)):
            g.basis[l] = gt[l].data_ptr()
        for k in range(len(c)):
            g.coeff[k] = gt[len(b) + k].data_ptr()
        g.w0, g.b0, g.w1 = gp[0].data_ptr(), gp[1].data_ptr(), gp[2].data_ptr()
        ws_bytes = L.lib().ff_render_field_bwd_workspace_bytes(ctypes.byref(d), ctypes.byref(m), cap)
        ws = torch.empty((max(ws_bytes, 4) + 3) // 4, device=rays.device)
        L.check(L.lib().ff_render_field_bwd(ctypes.byref(d), ctypes.byref(m), ctypes.byref(h), L.ptr(rays),
                                            B, ctypes.byref(s), L.ptr(d_sample), ctypes.byref(g), L.ptr(ws),
                                            ws_bytes, st), "ff_render_field_bwd")
        needs = ctx.needs_input_grad[5:]
        grads = [gr if nd else None for gr, nd in zip(gp + gt, needs)]
        return (None, None, None, None, None) + tuple(grads)


def render_ray(rays, factor_fields, chunk=4096, N_samples=-1, ndc_ray=False, white_bg=True,
               is_train=False, device="cuda"):
    """renderer.py:8-27: chunked rendering of [N, 6] rays."""
    rgbs, depth_maps, coeffs = [], [], []
    n = rays.shape[0]
    for k in range(n // chunk + int(n % chunk > 0)):
        rc = rays[k * chunk:(k + 1) * chunk].to(device)
        rgb_map, depth_map, coeff = factor_fields(rc, is_train=is_train, white_bg=white_bg,
                                                  ndc_ray=ndc_ray, N_samples=N_samples)
        rgbs.append(rgb_map)
        depth_maps.append(depth_map)
        if is_train:
            coeffs.append(coeff)
    if is_train:
        return torch.cat(rgbs), torch.cat(depth_maps), torch.cat(coeffs)
    return torch.cat(rgbs), torch.cat(depth_maps)
