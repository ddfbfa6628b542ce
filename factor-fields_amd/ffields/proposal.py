"""Learned proposal sampler: a 64^3 x 8 density grid + MLP 8 -> 32 -> 1,
trained online on the field's own rendering weights, that places the fine
samples of each ray (sample_pdf semantics, dataLoader/ray_utils.py:132-174).
All arithmetic runs in libffields.so (ff_proposal_sample / ff_proposal_bwd).
"""

import ctypes
import math

import torch
import torch.nn.functional as F
import torch.nn as nn

from . import _lib as L
from .ops import texel_major


class ProposalNet(nn.Module):
    def __init__(self, aabb, reso=64, hidden=32, n_fine=48, density_shift=-5.0, distance_scale=25.0,
                 near=0.05, table_dtype="bf16", device="cuda", seed=0):
        super().__init__()
        if hidden != 32:
            raise NotImplementedError("the proposal kernels are built for 32 hidden units")
        self.device = torch.device(device)
        g = torch.Generator().manual_seed(seed)
        grid = torch.randn((1, 8, reso, reso, reso), generator=g) * 0.1
        self.grid = nn.Parameter(texel_major(grid.to(self.device)))
        self.mlp = nn.ModuleList([nn.Linear(8, hidden), nn.Linear(hidden, 1, bias=False)]).to(self.device)
        self.aabb = torch.as_tensor(aabb, dtype=torch.float32).to(self.device)
        self._aabb_list = torch.as_tensor(aabb, dtype=torch.float32).cpu().tolist()
        self.reso = reso
        self.n_fine = n_fine
        self.density_shift = float(density_shift)
        self.distance_scale = float(distance_scale)
        self.near = float(near)
        self.table_dtype = table_dtype
        self._shadow = None
        self._ws = None
        if table_dtype == "bf16":
            self.refresh_shadow()

    def init_args(self):
        """Constructor arguments that rebuild this network's shape (for
        checkpoints; the trained values come from its state_dict)."""
        return {"reso": self.reso, "n_fine": self.n_fine, "density_shift": self.density_shift,
                "table_dtype": self.table_dtype}

    def refresh_shadow(self):
        t = self.grid
        s = torch.empty(t.shape, dtype=torch.bfloat16, device=t.device, memory_format=torch.channels_last_3d)
        L.check(L.lib().ff_cast_bf16(L.ptr(t), L.ptr(s), t.numel(), L.stream_of(t)), "ff_cast_bf16")
        self._shadow = s
        self._shadow_src = (t, t._version)

    def shadow_of(self, p):
        return self._shadow if (p is self.grid and self.table_dtype == "bf16") else None

    def desc(self):
        """The ProposalDesc, rebuilt only when a pointer or scalar it holds
        changed since the last call."""
        if self.table_dtype == "bf16":
            # the grid was written through torch since the copy was cast (its
            # version moved, or it was replaced): cast again; FusedAdam keeps
            # both current without touching the version
            src, ver = self._shadow_src
            if src is not self.grid or src._version != ver:
                self.refresh_shadow()
        grid = (self._shadow if self.table_dtype == "bf16" else self.grid).data_ptr()
        key = (grid, self.mlp[0].weight.data_ptr(), self.mlp[0].bias.data_ptr(), self.mlp[1].weight.data_ptr(),
               self.reso, self.table_dtype, tuple(map(tuple, self._aabb_list)), self.density_shift,
               self.distance_scale, self.near, self.n_fine)
        c = self.__dict__.get("_desc_cache")
        if c is None or c[0] != key:
            c = (key, self._build_desc(grid))
            self.__dict__["_desc_cache"] = c
        return c[1]

    def _build_desc(self, grid):
        d = L.ProposalDesc()
        d.reso[0] = d.reso[1] = d.reso[2] = self.reso
        d.table_dtype = L.BF16 if self.table_dtype == "bf16" else L.F32
        d.grid = grid
        d.w0 = self.mlp[0].weight.data_ptr()
        d.b0 = self.mlp[0].bias.data_ptr()
        d.w1 = self.mlp[1].weight.data_ptr()
        a = self._aabb_list
        for k in range(3):
            d.aabb_min[k], d.aabb_max[k] = a[0][k], a[1][k]
        d.density_shift = self.density_shift
        d.distance_scale = self.distance_scale
        d.near = self.near
        d.n_fine = self.n_fine
        return d

    def sample(self, rays, samples_struct, prop_u=None, fine_u=None):
        d = self.desc()
        L.check(L.lib().ff_proposal_sample(ctypes.byref(d), L.ptr(rays), L.ptr(prop_u), L.ptr(fine_u),
                                           rays.shape[0], ctypes.byref(samples_struct), L.stream_of(rays)),
                "ff_proposal_sample")

    def backward(self, rays, samples_struct, weights, prop_u=None, loss_scale=1.0, loss_sum=None):
        for p in self.parameters():
            if p.grad is None:
                p.grad = torch.zeros_like(p, memory_format=torch.preserve_format)
        g = L.ProposalGrads()
        g.grid = self.grid.grad.data_ptr()
        g.w0 = self.mlp[0].weight.grad.data_ptr()
        g.b0 = self.mlp[0].bias.grad.data_ptr()
        g.w1 = self.mlp[1].weight.grad.data_ptr()
        g.replicas, g.replica_stride = 1, 0
        d = self.desc()
        nb = L.lib().ff_proposal_bwd_workspace_bytes(rays.shape[0])
        ws = self._ws
        if ws is None or ws.numel() * 4 < nb:
            ws = self._ws = torch.empty((max(nb, 4) + 3) // 4, device=rays.device)
        L.check(L.lib().ff_proposal_bwd(ctypes.byref(d), L.ptr(rays), L.ptr(prop_u), rays.shape[0],
                                        ctypes.byref(samples_struct), L.ptr(weights), float(loss_scale),
                                        ctypes.byref(g), L.ptr(loss_sum), L.ptr(ws), nb, L.stream_of(rays)),
                "ff_proposal_bwd")


def _sq_err_add(a, b, out):
    """out += sum((a - b)^2) on the current stream (ff_sq_err_sum, one launch)."""
    a = a.contiguous()
    b = b.to(a.dtype).contiguous()
    if a.shape != b.shape:
        raise ValueError("shapes differ")
    L.check(L.lib().ff_sq_err_sum(L.ptr(a), L.ptr(b), a.numel(), L.ptr(out), L.stream_of(a)), "ff_sq_err_sum")


class ProposalReferee:
    """The occupancy sampler as fallback and referee of the learned proposal.

    ``compare(model, rays, target)`` renders the first ``n_rays`` rays of a
    training batch with the occupancy sampler and accumulates its squared
    error on the device.  The render only reads tables and weights, so it
    runs on a side stream (default priority) beside the batch's training
    pass.  ``add_proposal(rgb_map, target)`` accumulates the proposal's
    squared error on the same rays from the training pass's own render
    (its first ``n_rays`` rows), so the proposal side costs one reduction
    instead of a second render.  ``join()`` makes the current stream wait for
    the side stream (call it before the optimizer writes the tables).
    ``decide()`` reads the accumulated errors (one host synchronisation) and
    switches the proposal off for good when PSNR(occupancy sampler) -
    PSNR(proposal) exceeds ``max_gap_db``.
    """

    def __init__(self, device, n_rays=256, max_gap_db=0.1, n_samples=-1):
        self.n_rays = n_rays
        self.max_gap_db = max_gap_db
        self.n_samples = n_samples
        self.sq_err = torch.zeros(2, device=device)      # occupancy sampler, proposal
        self.stream = torch.cuda.Stream(device=device)
        self._done = None
        self.active = True
        self.gaps_db = []

    def compare(self, model, rays, target):
        if not self.active or model.alphaMask is None:
            return
        main = torch.cuda.current_stream()
        ready = torch.cuda.Event()
        ready.record(main)
        jr, jt = rays[:self.n_rays], target[:self.n_rays]
        with torch.cuda.stream(self.stream):
            self.stream.wait_event(ready)
            occ = model.render_eval(jr, use_proposal=False, N_samples=self.n_samples)
            _sq_err_add(occ, jt, self.sq_err[0:1])
            self._done = torch.cuda.Event()
            self._done.record(self.stream)
        # the batch tensors are read on the side stream: keep their memory until it is done
        jr.record_stream(self.stream)
        jt.record_stream(self.stream)
        self._pending = True

    def add_proposal(self, rgb_map, target):
        """The proposal's squared error on the rays ``compare`` rendered
        (current stream); counted only when ``compare`` ran for this batch."""
        if not getattr(self, "_pending", False):
            return
        self._pending = False
        n = self.n_rays
        _sq_err_add(rgb_map[:n], target[:n], self.sq_err[1:2])

    def join(self):
        if self._done is not None:
            torch.cuda.current_stream().wait_event(self._done)
            self._done = None

    def decide(self):
        """Returns the gap (dB) of the errors accumulated since the last call,
        or None when there were none; deactivates the proposal above the bar."""
        self.join()
        s0, s1 = self.sq_err.tolist()
        self.sq_err.zero_()
        if not (s0 > 0 and s1 > 0):
            return None
        gap = 10 * math.log10(s1 / s0)      # PSNR(occupancy sampler) - PSNR(proposal)
        self.gaps_db.append(round(gap, 4))
        if gap > self.max_gap_db:
            self.active = False
        return gap
