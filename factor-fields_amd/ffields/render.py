"""Volume rendering on the HIP kernels: FactorFields.forward
(models/FactorFields.py:843-898) for bounded scenes, with the sampler
(:586-602), AlphaGridMask occupancy (:91-110), field + density head,
colour module, compositing (:82-88, :878-896) and their backward passes.

Two colour modules (cfg.renderer.shadingMode):
- ``MLP_Fea``: the reference's own MLPRender_Fea (:162-203) on decoder
  features 1..31, evaluated only on the app samples (weight >
  rayMarch_weight_thres, :878-885) by ff_color_mlp_fwd / ff_color_mlp_bwd;
- ``SH``: degree-2 spherical harmonics (models/sh.py:34-64) on decoder
  outputs 1..27 followed by a sigmoid, evaluated inside the field kernel
  (``renderModule`` then holds no parameters).
"""

import ctypes
import warnings

import numpy as np
import torch
import torch.nn as nn
import torch.nn.functional as F

from . import _lib as L
from . import ops
from .ops import mlp_desc, texel_major
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
    """Degree-2 SH colour head (no parameters): rgb = sigmoid(eval_sh(2,
    features viewed as [..., 3, 9], viewdirs)) (models/sh.py:34-85 with
    MLPRender_Fea's output sigmoid, models/FactorFields.py:202).  The render
    kernels evaluate it inline; called on its own it runs the ffields::sh_render
    custom op (ff_sh_rgb), differentiable in the features."""

    def forward(self, viewdirs, features):
        if features.shape[-1] != 27:
            raise ValueError(f"the SH head takes 27 features (3 x 9), got {features.shape[-1]}")
        return ops.sh_render_op(viewdirs, features)


class MLPRender_Fea(nn.Module):
    """models/FactorFields.py:162-203 on the HIP colour kernels (FP32 matrix
    cores), with the reference's state_dict keys (mlp.{l}.weight / .bias).
    Built for inChanel 31, fea_pe 2, view_pe 6 or 2, hidden 128, 3 layers (the
    reference's configs); other shapes raise NotImplementedError."""

    def __init__(self, inChanel, num_layers=3, hidden_dim=64, viewpe=6, feape=2):
        super().__init__()
        if (inChanel, num_layers, hidden_dim, feape) != (31, 3, 128, 2) or viewpe not in (6, 2):
            raise NotImplementedError(
                f"MLPRender_Fea(inChanel={inChanel}, num_layers={num_layers}, hidden_dim={hidden_dim}, "
                f"viewpe={viewpe}, feape={feape}) has no HIP kernel (built: 31, 3, 128, 6|2, 2)")
        self.in_mlpC = 3 + inChanel + 2 * viewpe * 3 + 2 * feape * inChanel
        self.num_layers = num_layers
        self.viewpe = viewpe
        self.feape = feape
        self.inChanel = inChanel
        mlp = []
        for l in range(num_layers):
            in_dim = self.in_mlpC if l == 0 else hidden_dim
            out_dim, bias = (3, False) if l == num_layers - 1 else (hidden_dim, True)
            mlp.append(nn.Linear(in_dim, out_dim, bias=bias))
        self.mlp = nn.ModuleList(mlp)

    def params_list(self):
        return [self.mlp[0].weight, self.mlp[0].bias, self.mlp[1].weight, self.mlp[1].bias, self.mlp[2].weight]

    def desc(self):
        d = L.ColorMlpDesc()
        d.in_feat, d.view_pe, d.fea_pe, d.hidden, d.n_layers = self.inChanel, self.viewpe, self.feape, 128, 3
        p = self.params_list()
        d.w0, d.b0, d.w1, d.b1, d.w2 = (t.data_ptr() for t in p)
        return d

    def grads_struct(self, grads):
        g = L.ColorMlpGrads()
        g.w0, g.b0, g.w1, g.b1, g.w2 = (t.data_ptr() for t in grads)
        return g

    def forward(self, viewdirs, features):
        return ColorMLP.apply(self, viewdirs, features, *self.params_list())


def _color_ws(module, n_max, backward, device, cache=None):
    nb = L.lib().ff_color_mlp_workspace_bytes(ctypes.byref(module.desc()), n_max, int(backward))
    if cache is not None:
        return cache(nb), nb
    return torch.empty((max(nb, 4) + 3) // 4, device=device), nb


class ColorMLP(torch.autograd.Function):
    """renderModule(viewdirs, features) -> rgb [N, 3] with the hand-written
    backward (feature and weight gradients; view directions get none)."""

    @staticmethod
    def forward(ctx, module, viewdirs, features, *params):
        v = viewdirs.contiguous().float()
        f = features.contiguous().float()
        n = f.shape[0]
        rgb = torch.empty((n, 3), device=f.device)
        r = L.ColorRows(f.data_ptr(), f.shape[1], v.data_ptr(), 3, None, None, None, n)
        ws, nb = _color_ws(module, n, False, f.device)
        L.check(L.lib().ff_color_mlp_fwd(ctypes.byref(module.desc()), ctypes.byref(r), L.ptr(rgb), 3, L.ptr(ws), nb,
                                         L.stream_of(f)), "ff_color_mlp_fwd")
        ctx.module = module
        ctx.save_for_backward(v, f, *params)
        return rgb

    @staticmethod
    def backward(ctx, d_rgb):
        v, f, *params = ctx.saved_tensors
        module = ctx.module
        n = f.shape[0]
        d_feat = torch.empty_like(f)
        gp = [torch.zeros_like(p) for p in params]
        r = L.ColorRows(f.data_ptr(), f.shape[1], v.data_ptr(), 3, None, None, None, n)
        ws, nb = _color_ws(module, n, True, f.device)
        L.check(L.lib().ff_color_mlp_bwd(ctypes.byref(module.desc()), ctypes.byref(r), L.ptr(d_rgb.contiguous()), 3,
                                         L.ptr(d_feat), f.shape[1], ctypes.byref(module.grads_struct(gp)), L.ptr(ws),
                                         nb, L.stream_of(f)), "ff_color_mlp_bwd")
        needs = ctx.needs_input_grad
        return (None, None, d_feat if needs[2] else None) + tuple(g if nd else None for g, nd in zip(gp, needs[3:]))


def _samples_struct(buf):
    s = L.Samples()
    s.xyz = buf["xyz"].data_ptr()
    s.z = buf["z"].data_ptr()
    s.dist = buf["dist"].data_ptr()
    s.ray = buf["ray"].data_ptr()
    s.idx = buf["idx"].data_ptr()
    s.ray_off = buf["ray_off"].data_ptr()
    s.cap = buf["xyz"].shape[0]
    coding = buf.get("coding")
    s.coding = coding.data_ptr() if coding is not None else None
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
        mode = cfg.renderer.get("shadingMode", "MLP_Fea")   # the reference always builds MLPRender_Fea
        if mode == "MLP_Fea":
            # models/FactorFields.py:243-245
            self.renderModule = MLPRender_Fea(inChanel=cfg.model.out_dim - 1, num_layers=cfg.renderer.num_layers,
                                              hidden_dim=cfg.renderer.hidden_dim, viewpe=cfg.renderer.view_pe,
                                              feape=cfg.renderer.fea_pe).to(self.device)
            if cfg.model.out_dim != 32:
                raise NotImplementedError("the feature decoder of the render kernels has out_dim 32")
        elif mode == "SH":
            if cfg.model.out_dim != 28:
                raise ValueError("the SH colour head needs out_dim = 1 + 27")
            self.renderModule = SHRender()
        else:
            raise NotImplementedError(f"shadingMode {mode!r} has no HIP kernel (built: MLP_Fea, SH)")
        self._feat_head = mode == "MLP_Fea"
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

    @torch.no_grad()
    def upsample_volume_grid(self, res_target):
        """models/FactorFields.py:701-708: new render parameters for the grid
        resolution res_target (step size and sample count); for a
        single-scene google_objs configuration with a grid coefficient, the
        coefficient grid is also resized by 1.3 on every axis (trilinear,
        align_corners=True, as the reference does).  The resized grid is a
        new parameter, so an optimizer must be rebuilt afterwards (the
        reference's training loop does so, train_per_scene.py:213-218)."""
        self.update_renderParams(res_target)
        if (self.cfg.dataset.get("dataset_name") == "google_objs" and self.n_scene == 1
                and self.coeff_type == "grid"):
            with torch.no_grad():
                c = self.coeffs[0].data
                up = F.interpolate(c.contiguous(), size=None, scale_factor=1.3, align_corners=True,
                                   mode="trilinear" if c.dim() == 5 else "bilinear")
            self.coeffs = nn.ParameterList([nn.Parameter(texel_major(up))])
            self.coeff_reso = list(up.shape[2:])
            self._build_spec()

    def _aabb_host(self):
        """Host copy of the box (the kernels take it by value); refreshed only
        when the device tensor object changes, so a training step does not
        synchronise with the device."""
        ref = getattr(self, "_aabb_ref", None)
        if ref is None or ref[0] is not self.aabb or ref[1] != self.aabb._version:
            self._aabb_list = self.aabb.cpu().tolist()
            self._aabb_ref = (self.aabb, self.aabb._version)   # holds the tensor: no id reuse
        return self._aabb_list

    def _sampler(self, n_samples):
        s = L.RaySampler()
        a = self._aabb_host()
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

    def _head(self, buf=None):
        """Density head, and with the MLP_Fea colour module the feature rows of
        `buf` (written by the field forward) and their gradients (read by the
        field backward through app_pos)."""
        h = L.RenderHead()
        h.density_act = 0 if self.cfg.renderer.fea2denseAct == "softplus" else 1
        h.density_shift = float(self.cfg.renderer.density_shift)
        if getattr(self, "_feat_head", False):
            h.color = L.COLOR_FEATURES
            h.feat_stride = 32
            if buf is not None:
                h.feat = buf["feat"].data_ptr()
                h.d_feat = buf["d_feat"].data_ptr()
                h.app_pos = buf["app_pos"].data_ptr()
        return h

    def _field_structs(self):
        """(field, decoder) descriptors, rebuilt only when the spec object or
        a table or weight pointer changed since the last call (building the
        ctypes structs costs tens of microseconds per call)."""
        tabs = self._read_tables()
        layers = self._layers()
        key = (tuple(t.data_ptr() for t in tabs),
               tuple(x.data_ptr() for lw in layers for x in lw if x is not None))
        c = self.__dict__.get("_struct_cache")
        if c is None or c[0] is not self._spec or c[1] != key:
            c = (self._spec, key, self._spec.desc(*self._split(tabs)), mlp_desc(layers))
            self.__dict__["_struct_cache"] = c
        return c[2], c[3]

    def render_buffers(self, B, Ns):
        """Packed-sample and per-sample buffers for B rays x Ns candidates,
        kept between steps (a few shapes are cached).  They include the
        per-sample coding rows, so a field backward on these buffers reads
        what the field forward on them wrote: call the backward right after
        the forward of the same samples, with the same tables."""
        key = (B, Ns)
        cache = self.__dict__.setdefault("_rbufs", {})
        if key not in cache:
            if len(cache) >= 4:
                cache.pop(next(iter(cache)))
            dev = self.device
            buf = alloc_samples(B, Ns, dev)
            cap = buf["xyz"].shape[0]
            # the field forward keeps each sample's basis and coefficient
            # values here for the backward of the same pass (BF16 tables, or
            # FP32 with a grid coefficient, whose backward then runs as a
            # decoder kernel and a scatter kernel and leaves the rows' gradients
            # in place of the rows)
            keep_rows = self.table_dtype == "bf16" or self.coeff_type == "grid"
            buf.update(sigma=torch.empty(cap, device=dev), rgb=torch.empty((cap, 3), device=dev),
                       d_sample=torch.empty((cap, 4), device=dev),
                       coding=(torch.empty((2 * self._spec.C, cap), device=dev) if keep_rows else None))
            self._feature_buffers(buf, B)
            cache[key] = buf
        self._rbuf = (key, cache[key])
        return cache[key]

    def _feature_buffers(self, buf, B):
        """Per-sample buffers of the MLP_Fea colour module: decoder features
        [cap, 32], their gradients (compact, one row per app sample), the app
        list, each sample's position in it, and the per-ray app offsets."""
        if not getattr(self, "_feat_head", False):
            return buf
        cap = buf["xyz"].shape[0]
        dev = buf["xyz"].device
        buf.update(feat=torch.empty((cap, 32), device=dev), d_feat=torch.empty((cap, 32), device=dev),
                   app_rows=torch.empty(cap, device=dev, dtype=torch.int32),
                   app_pos=torch.empty(cap, device=dev, dtype=torch.int32),
                   app_off=torch.empty(B + 1, device=dev, dtype=torch.int32))
        return buf

    def _color_rows(self, rays, buf):
        cap = buf["xyz"].shape[0]
        B = rays.shape[0]
        return L.ColorRows(buf["feat"].data_ptr(), 32, rays.data_ptr() + 12, 6, buf["ray"].data_ptr(),
                           buf["app_rows"].data_ptr(), buf["app_off"].data_ptr() + 4 * B, cap)

    def _color_workspace(self, nbytes):
        ws = getattr(self, "_cws", None)
        if ws is None or ws.numel() * 4 < nbytes:
            ws = torch.empty((max(nbytes, 4) + 3) // 4, device=self.device, dtype=torch.float32)
            self._cws = ws
        return ws

    def render_samples(self, rays, n_samples, jitter, buf):
        """Sampler + field + colour over `rays`; fills buf (packed samples,
        sigma, rgb).  Returns the ff_samples struct."""
        s = _samples_struct(buf)
        st = L.stream_of(rays)
        smp = self._sampler(n_samples)
        L.check(L.lib().ff_sample_rays(ctypes.byref(smp), L.ptr(rays), L.ptr(jitter), rays.shape[0],
                                       ctypes.byref(s), st), "ff_sample_rays")
        self._field_fwd(rays, s, buf, st)
        return s


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
        if is_train and self.cfg.model.get("with_dropout", False):
            raise NotImplementedError("linear_mat input dropout inside the render kernels is not built")
        params = self.linear_mat.params_list() + list(self._tables()[0]) + list(self._tables()[1])
        if getattr(self, "_feat_head", False):
            params = params + self.renderModule.params_list()
        rgb_map, depth = RenderRays.apply(self, rays, jit, n_samples, bg, *params)
        return rgb_map, depth, self._render_coeffs(is_train)

    def _render_coeffs(self, is_train):
        """The coeffs FactorFields.forward returns (models/FactorFields.py:
        871-875): get_coding's coefficient values at the valid samples of the
        last render, zeros (1, C) when there were none.  Only computed for
        training renders, the only ones render_ray keeps (renderer.py:18-21);
        it reads the sample count on the host."""
        C = sum(self.basis_dims)
        buf = getattr(self, "_last_render_buf", None)
        if not is_train or buf is None:
            return torch.zeros((1, C), device=self.device)
        n = int(buf["ray_off"][-1])
        if n == 0 or self.coeff_type == "none":
            return torch.zeros((1, C), device=self.device)
        return self.get_coeff(buf["xyz"][:n])

    # ------------------------------------------------ learned proposal sampler
    def enable_proposal(self, reso=64, n_fine=48, table_dtype="bf16", density_shift=-5.0, seed=0):
        """Attach a ProposalNet (64^3 x 8 grid + MLP 8-32-1) that places n_fine
        samples per ray; the occupancy sampler stays available as fallback
        and reference sampler (use_proposal toggles which one trains)."""
        from .proposal import ProposalNet
        self.proposal = ProposalNet(self.aabb.cpu(), reso=reso, n_fine=n_fine, density_shift=density_shift,
                                    distance_scale=float(self.cfg.renderer.distance_scale),
                                    table_dtype=table_dtype, device=self.device, seed=seed)
        self.use_proposal = True
        return self.proposal

    def shadow_lookup(self, p):
        s = self.shadow_of(p)
        if s is None and getattr(self, "proposal", None) is not None:
            s = self.proposal.shadow_of(p)
        return s

    def _composite(self, s, B, buf, bg, rgb_map, rgb_raw, depth=None, weights=None, stream=None):
        thres = float(self.cfg.renderer.rayMarch_weight_thres)
        L.check(L.lib().ff_composite_fwd(ctypes.byref(s), B, L.ptr(buf["sigma"]), L.ptr(buf["rgb"]), int(bg),
                                         thres, L.ptr(rgb_map), L.ptr(rgb_raw), L.ptr(depth), None,
                                         L.ptr(weights), stream), "ff_composite_fwd")

    def _field_fwd(self, rays, s, buf, stream):
        """Field + density head over the packed samples, then the colour: the
        SH head inside the field kernel, or (MLP_Fea) the app-sample selection
        (models/FactorFields.py:878-882) and the colour module on those
        samples (:883-885); rgb of the other samples is never read."""
        d, m = self._field_structs()
        h = self._head(buf)
        L.check(L.lib().ff_render_field_fwd(ctypes.byref(d), ctypes.byref(m), ctypes.byref(h), L.ptr(rays),
                                            rays.shape[0], ctypes.byref(s), L.ptr(buf["sigma"]),
                                            L.ptr(buf["rgb"]), stream), "ff_render_field_fwd")
        if getattr(self, "_feat_head", False):
            B = rays.shape[0]
            thres = float(self.cfg.renderer.rayMarch_weight_thres)
            L.check(L.lib().ff_select_app(ctypes.byref(s), B, L.ptr(buf["sigma"]), thres, L.ptr(buf["app_off"]),
                                          L.ptr(buf["app_rows"]), L.ptr(buf["app_pos"]), stream), "ff_select_app")
            rm = self.renderModule
            r = self._color_rows(rays, buf)
            nb = L.lib().ff_color_mlp_workspace_bytes(ctypes.byref(rm.desc()), r.n_max, 0)
            ws = self._color_workspace(nb)
            L.check(L.lib().ff_color_mlp_fwd(ctypes.byref(rm.desc()), ctypes.byref(r), L.ptr(buf["rgb"]), 3,
                                             L.ptr(ws), nb, stream), "ff_color_mlp_fwd")

    def _color_bwd(self, rays, buf, grads, stream):
        """MLP_Fea backward on the app samples: dL/drgb from d_sample ->
        compact feature gradients (d_feat) and the colour weights' grads."""
        rm = self.renderModule
        r = self._color_rows(rays, buf)
        nb = L.lib().ff_color_mlp_workspace_bytes(ctypes.byref(rm.desc()), r.n_max, 1)
        ws = self._color_workspace(nb)
        L.check(L.lib().ff_color_mlp_bwd(ctypes.byref(rm.desc()), ctypes.byref(r), buf["d_sample"].data_ptr() + 4, 4,
                                         L.ptr(buf["d_feat"]), 32, ctypes.byref(rm.grads_struct(grads)), L.ptr(ws), nb,
                                         stream), "ff_color_mlp_bwd")

    def _train_from_samples(self, rays, target, s, buf, bg, loss_scale, loss_sum, weights=None):
        """composite forward (+weights) + MSE backward + field backward."""
        B = rays.shape[0]
        st = L.stream_of(rays)
        rgb_map = _buf(buf, "rgb_map", (B, 3), rays.device)
        rgb_raw = _buf(buf, "rgb_raw", (B, 3), rays.device)
        thres = float(self.cfg.renderer.rayMarch_weight_thres)
        self._composite(s, B, buf, bg, rgb_map, rgb_raw, weights=weights, stream=st)
        L.check(L.lib().ff_composite_bwd(ctypes.byref(s), B, L.ptr(buf["sigma"]), L.ptr(buf["rgb"]),
                                         int(bg), thres, L.ptr(rgb_raw), None, L.ptr(target),
                                         float(loss_scale), L.ptr(loss_sum), L.ptr(buf["d_sample"]), st),
                "ff_composite_bwd")
        g = self._train_grads()
        if getattr(self, "_feat_head", False):
            self._color_bwd(rays, buf, [p.grad for p in self.renderModule.params_list()], st)
        d, m = self._field_structs()
        h = self._head(buf)
        ws_bytes = L.lib().ff_render_field_bwd_workspace_bytes(ctypes.byref(d), ctypes.byref(m), s.cap)
        ws = self._workspace(ws_bytes)
        ev = getattr(self, "kernel_events", None)
        if ev is not None:
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
        L.check(L.lib().ff_render_field_bwd(ctypes.byref(d), ctypes.byref(m), ctypes.byref(h), L.ptr(rays),
                                            B, ctypes.byref(s), L.ptr(buf["d_sample"]), ctypes.byref(g),
                                            L.ptr(ws), ws_bytes, st), "ff_render_field_bwd")
        if ev is not None:
            e1.record()
            ev.append((e0, e1))
        return rgb_map

    def render_train_step_proposal(self, rays, target, white_bg=True, loss_scale=1.0, prop_loss_scale=1.0,
                                   loss_sum=None, prop_loss_sum=None, reduce=True, prop_u=None, fine_u=None):
        """NeRF training pass with the learned proposal placing the samples:
        proposal sample -> field -> composite (weights kept) -> MSE backward ->
        field backward -> proposal backward on the field's weights."""
        pr = self.proposal
        rays = rays.contiguous()
        target = target.contiguous()
        B = rays.shape[0]
        nf = pr.n_fine
        if prop_u is None or fine_u is None:
            # both draws from one launch: two contiguous slices of one buffer
            u = torch.rand(B * (64 + nf), device=rays.device)
            prop_u = u[:B * 64].view(B, 64) if prop_u is None else prop_u
            fine_u = u[B * 64:].view(B, nf) if fine_u is None else fine_u
        bg = bool(white_bg or torch.rand((1,)) < 0.5)
        buf = self.render_buffers(B, nf)
        s = _samples_struct(buf)
        st = L.stream_of(rays)
        pr.sample(rays, s, prop_u, fine_u)
        self._field_fwd(rays, s, buf, st)
        if loss_sum is None:
            loss_sum = torch.zeros((), device=rays.device)
        weights = _buf(buf, "weights", (s.cap,), rays.device)
        self._train_from_samples(rays, target, s, buf, bg, loss_scale, loss_sum, weights=weights)
        pr.backward(rays, s, weights, prop_u=prop_u, loss_scale=prop_loss_scale, loss_sum=prop_loss_sum)
        if reduce:
            self.reduce_grad_replicas()
        return loss_sum

    def last_rgb_map(self):
        """rgb_map [B, 3] of the most recent fused training pass
        (render_train_step or render_train_step_proposal); the buffer is
        reused by the next pass."""
        return self._rbuf[1]["rgb_map"]

    @torch.no_grad()
    def render_eval(self, rays, white_bg=True, use_proposal=None, N_samples=-1):
        """Forward-only render of `rays` with the occupancy sampler or the
        proposal (deterministic draws).  Returns rgb_map [B, 3]."""
        use_proposal = getattr(self, "use_proposal", False) if use_proposal is None else use_proposal
        rays = rays.contiguous()
        B = rays.shape[0]
        st = L.stream_of(rays)
        if use_proposal:
            pr = self.proposal
            buf = alloc_samples(B, pr.n_fine, rays.device)
            s = _samples_struct(buf)
            pr.sample(rays, s)
        else:
            n_samples = N_samples if N_samples > 0 else self.nSamples
            buf = alloc_samples(B, n_samples, rays.device)
            s = _samples_struct(buf)
            smp = self._sampler(n_samples)
            L.check(L.lib().ff_sample_rays(ctypes.byref(smp), L.ptr(rays), None, B, ctypes.byref(s), st),
                    "ff_sample_rays")
        cap = buf["xyz"].shape[0]
        buf.update(sigma=torch.empty(cap, device=rays.device), rgb=torch.empty((cap, 3), device=rays.device))
        self._feature_buffers(buf, B)
        self._field_fwd(rays, s, buf, st)
        rgb_map = torch.empty((B, 3), device=rays.device)
        rgb_raw = torch.empty((B, 3), device=rays.device)
        self._composite(s, B, buf, white_bg, rgb_map, rgb_raw, stream=st)
        return rgb_map

    def render_train_step(self, rays, target, white_bg=True, N_samples=-1, jitter=None, loss_scale=1.0,
                          loss_sum=None, reduce=True):
        """One NeRF training pass fused across kernels: sample, field, composite,
        MSE loss (train_per_scene.py:158), composite backward, field backward
        into each parameter's .grad (reduce=False leaves the table gradients
        split over the per-XCD copies for FusedAdam).  Returns the device
        scalar sum((rgb-gt)^2)."""
        n_samples = N_samples if N_samples > 0 else self.nSamples
        rays = rays.contiguous()
        target = target.contiguous()
        B = rays.shape[0]
        jit = self._jitter(rays, True, jitter)
        bg = bool(white_bg or torch.rand((1,)) < 0.5)
        buf = self.render_buffers(B, n_samples)
        s = self.render_samples(rays, n_samples, jit, buf)
        if loss_sum is None:
            loss_sum = torch.zeros((), device=rays.device)
        self._train_from_samples(rays, target, s, buf, bg, loss_scale, loss_sum)
        if reduce:
            self.reduce_grad_replicas()
        return loss_sum


    # ------------------------------------------------ occupancy maintenance
    @torch.no_grad()
    def getDenseAlpha(self, gridSize=None, times=16, generator=None):
        """models/FactorFields.py:729-755: mean over `times` jittered passes of
        1 - exp(-sigma * stepSize * distance_scale) on a dense grid.  Returns
        (alpha [Z, Y, X], dense_xyz [Z, Y, X, 3])."""
        gridSize = self.gridSize.tolist() if gridSize is None else list(gridSize)
        dev = self.device
        aabb = self.inward_aabb
        aabbSize = aabb[1] - aabb[0]
        units = aabbSize / (torch.LongTensor(gridSize).to(dev) - 1)
        units_half = 1.0 / (torch.LongTensor(gridSize) - 1) * 0.5
        stepSize = torch.mean(units)
        samples = torch.stack(torch.meshgrid(
            [torch.linspace(units_half[0], 1 - units_half[0], gridSize[0]),
             torch.linspace(units_half[1], 1 - units_half[1], gridSize[1]),
             torch.linspace(units_half[2], 1 - units_half[2], gridSize[2])], indexing="ij"), -1).to(dev)
        dense_xyz = aabb[0] * (1 - samples) + aabb[1] * samples
        dense_xyz = dense_xyz.transpose(0, 2).contiguous()
        alpha = torch.zeros(dense_xyz.shape[:-1], device=dev)
        length = float(stepSize * self.cfg.renderer.distance_scale)
        d, m = self._field_structs()
        h = self._head()
        occ = self._sampler(1)
        flat = dense_xyz.view(-1, 3)
        for _ in range(times):
            if times > 1:
                shift = (torch.rand(flat.shape, device=dev, generator=generator) * 2 - 1) * (units / 2 * 1.2)
                pts = (flat + shift).contiguous()
            else:
                pts = flat
            L.check(L.lib().ff_render_alpha(ctypes.byref(d), ctypes.byref(m), ctypes.byref(h),
                                            ctypes.byref(occ), L.ptr(pts), pts.shape[0], length,
                                            L.ptr(alpha), L.stream_of(pts)), "ff_render_alpha")
        return alpha / times, dense_xyz

    @torch.no_grad()
    def updateAlphaMask(self, gridSize=(200, 200, 200), is_update_alphaMask=False, times=16, generator=None):
        """models/FactorFields.py:757-793.  Returns the new aabb."""
        gridSize = list(gridSize)
        alpha, dense_xyz = self.getDenseAlpha(gridSize, times=times, generator=generator)
        thres = float(self.cfg.renderer.alphaMask_thres) if is_update_alphaMask else 0.08
        Z, Y, X = alpha.shape
        occ = torch.empty((Z, Y, X), dtype=torch.uint8, device=self.device)
        L.check(L.lib().ff_occupancy_pool(L.ptr(alpha), Z, Y, X, 1.0, thres, L.ptr(occ),
                                          L.stream_of(alpha)), "ff_occupancy_pool")
        keep = remove_small_components(occ, min_size=float(np.mean([Z, Y, X])))
        if not bool(keep.any()):
            # Nothing above the threshold yet (an untrained field): the reference
            # would fail on the empty reduction below; keep the current mask and box.
            warnings.warn("updateAlphaMask: no voxel above the threshold; mask and box unchanged")
            return self.aabb.clone()
        alpha = keep.float()
        if is_update_alphaMask:
            self.alphaMask = AlphaGridMask(self.device, self.inward_aabb, alpha)
        valid_xyz = dense_xyz[alpha > 0.5]
        xyz_min = valid_xyz.amin(0)
        xyz_max = valid_xyz.amax(0)
        pad = (xyz_max - xyz_min) / 20
        return torch.stack((xyz_min - pad, xyz_max + pad))

    @torch.no_grad()
    def shrink(self, new_aabb):
        """models/FactorFields.py:796-809 (re-initialises every table, as the
        reference does)."""
        self.setup_params(new_aabb.tolist())
        if self.coeff_type != "none":
            del self.coeffs
            self.coeffs = self.init_coef()
        if self.basis_type != "none":
            del self.basises
            self.basises = self.init_basis()
        self.aabb = self.inward_aabb = new_aabb.to(self.device)
        self.cfg.dataset.aabb = self.aabb.tolist()
        if self.table_dtype == "bf16":
            self.refresh_shadows()
        self._build_spec()
        self.update_renderParams(self.gridSize.tolist())

    @torch.no_grad()
    def filtering_rays(self, all_rays, all_rgbs, N_samples=256, chunk=10240 * 5, bbox_only=False):
        """models/FactorFields.py:811-841: keep rays that hit the box
        (bbox_only) or reach an occupied cell (> 0 in the occupancy volume)."""
        keep = []
        for k in range(0, all_rays.shape[0], chunk):
            rc = all_rays[k:k + chunk].to(self.device).contiguous()
            if bbox_only or self.alphaMask is None:
                o, d = rc[:, :3], rc[:, 3:6]
                vec = torch.where(d == 0, torch.full_like(d, 1e-6), d)
                ra = (self.aabb[1] - o) / vec
                rb = (self.aabb[0] - o) / vec
                t_min = torch.minimum(ra, rb).amax(-1)
                t_max = torch.maximum(ra, rb).amin(-1)
                keep.append((t_max > t_min).cpu())
            else:
                buf = alloc_samples(rc.shape[0], N_samples, self.device)
                s = _samples_struct(buf)
                smp = self._sampler(N_samples)
                L.check(L.lib().ff_sample_rays(ctypes.byref(smp), L.ptr(rc), None, rc.shape[0],
                                               ctypes.byref(s), L.stream_of(rc)), "ff_sample_rays")
                cnt = buf["ray_off"][1:] - buf["ray_off"][:-1]
                keep.append((cnt > 0).cpu())
        mask = torch.cat(keep)
        return all_rays[mask], all_rgbs[mask]


def _buf(buf, key, shape, device):
    t = buf.get(key)
    if t is None or tuple(t.shape) != tuple(shape):
        t = buf[key] = torch.empty(shape, device=device)
    return t


def remove_small_components(mask, min_size):
    """skimage.morphology.remove_small_objects(mask, min_size, connectivity=1)
    (models/FactorFields.py:774-775) on the GPU by ff_remove_small_components:
    drop face-connected components with fewer than min_size voxels.  mask is
    a [D, H, W] CUDA tensor (non-zero = set); returns a uint8 tensor of 0/1."""
    if not mask.is_cuda or mask.dim() != 3:
        raise ValueError("remove_small_components takes a [D, H, W] CUDA tensor")
    m = (mask != 0).to(torch.uint8).contiguous()
    D, H, W = m.shape
    labels = torch.empty(m.shape, dtype=torch.int32, device=m.device)
    sizes = torch.empty(m.shape, dtype=torch.int32, device=m.device)
    out = torch.empty_like(m)
    L.check(L.lib().ff_remove_small_components(L.ptr(m), D, H, W, float(min_size), L.ptr(labels), L.ptr(sizes),
                                               L.ptr(out), L.stream_of(m)), "ff_remove_small_components")
    return out


class RenderRays(torch.autograd.Function):
    """rays -> (rgb_map, depth) with the hand-written backward."""

    @staticmethod
    def forward(ctx, model, rays, jitter, n_samples, white_bg, *params):
        B = rays.shape[0]
        dev = rays.device
        buf = alloc_samples(B, n_samples, dev)
        cap = buf["xyz"].shape[0]
        buf.update(sigma=torch.empty(cap, device=dev), rgb=torch.empty((cap, 3), device=dev))
        model._feature_buffers(buf, B)
        s = model.render_samples(rays, n_samples, jitter, buf)
        rgb_map = torch.empty((B, 3), device=dev)
        rgb_raw = torch.empty((B, 3), device=dev)
        depth = torch.empty(B, device=dev)
        thres = float(model.cfg.renderer.rayMarch_weight_thres)
        L.check(L.lib().ff_composite_fwd(ctypes.byref(s), B, L.ptr(buf["sigma"]), L.ptr(buf["rgb"]),
                                         int(white_bg), thres, L.ptr(rgb_map), L.ptr(rgb_raw),
                                         L.ptr(depth), None, None, L.stream_of(rays)), "ff_composite_fwd")
        ctx.model, ctx.white_bg, ctx.thres, ctx.buf = model, white_bg, thres, buf
        model._last_render_buf = buf
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
        d_sample = buf["d_sample"] = torch.empty((cap, 4), device=rays.device)
        L.check(L.lib().ff_composite_bwd(ctypes.byref(s), B, L.ptr(buf["sigma"]), L.ptr(buf["rgb"]),
                                         int(ctx.white_bg), ctx.thres, L.ptr(rgb_raw),
                                         L.ptr(d_rgb.contiguous()), None, 1.0, None, L.ptr(d_sample), st),
                "ff_composite_bwd")
        gc = []
        if getattr(model, "_feat_head", False):
            gc = [torch.zeros_like(t) for t in model.renderModule.params_list()]
            model._color_bwd(rays, buf, gc, st)
        d, m = model._field_structs()
        h = model._head(buf)
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
        grads = [gr if nd else None for gr, nd in zip(gp + gt + gc, needs)]
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
