"""FusedAdam: torch.optim.Adam (betas (0.9, 0.99) in the reference,
train_per_scene.py:132) as one HIP launch over all parameters, with optional
BF16 shadow refresh and in-kernel gradient clearing."""

import ctypes

import torch

from . import _lib as L


def _dense(t):
    """True if t occupies one contiguous block in one of the layouts used here."""
    if t.is_contiguous():
        return True
    if t.dim() == 4 and t.is_contiguous(memory_format=torch.channels_last):
        return True
    return t.dim() == 5 and t.is_contiguous(memory_format=torch.channels_last_3d)


class FusedAdam(torch.optim.Optimizer):
    def __init__(self, params, lr=1e-3, betas=(0.9, 0.99), eps=1e-8, shadow_of=None):
        super().__init__(params, dict(lr=lr, betas=betas, eps=eps))
        self.shadow_of = shadow_of

    @torch.no_grad()
    def step(self, closure=None, zero_grad=False):
        loss = None
        if closure is not None:
            with torch.enable_grad():
                loss = closure()
        buckets = {}
        for group in self.param_groups:
            b1, b2 = group["betas"]
            for p in group["params"]:
                if p.grad is None:
                    continue
                g = p.grad
                if p.dtype != torch.float32 or g.dtype != torch.float32:
                    raise TypeError("FusedAdam needs float32 parameters and gradients")
                if g.stride() != p.stride() or not _dense(p):
                    raise RuntimeError("gradient layout differs from its parameter")
                st = self.state[p]
                if not st:
                    st["step"] = 0
                    st["exp_avg"] = torch.zeros_like(p, memory_format=torch.preserve_format)
                    st["exp_avg_sq"] = torch.zeros_like(p, memory_format=torch.preserve_format)
                st["step"] += 1
                shadow = self.shadow_of(p) if self.shadow_of is not None else None
                R, stride, where = getattr(p, "_ff_rep", (1, 0, None))
                if where != g.data_ptr():
                    R, stride = 1, 0
                t = L.AdamTensor(p.data_ptr(), g.data_ptr(), st["exp_avg"].data_ptr(),
                                 st["exp_avg_sq"].data_ptr(), 0 if shadow is None else shadow.data_ptr(),
                                 p.numel(), float(group["lr"]), R, stride)
                key = (st["step"], float(b1), float(b2), float(group["eps"]), p.device)
                buckets.setdefault(key, []).append(t)
        for (step, b1, b2, eps, dev), ts in buckets.items():
            arr = (L.AdamTensor * len(ts))(*ts)
            if dev.index is None or dev.index == torch.cuda.current_device():
                L.check(L.lib().ff_adam_multi(arr, len(ts), b1, b2, eps, step, 1 if zero_grad else 0,
                                              L.stream_of()), "ff_adam_multi")
                continue
            with torch.cuda.device(dev):
                L.check(L.lib().ff_adam_multi(arr, len(ts), b1, b2, eps, step, 1 if zero_grad else 0,
                                              L.stream_of()), "ff_adam_multi")
        return loss
