"""FactorFields: drop-in for models/FactorFields.py:206-898 on the HIP kernels.

The module keeps the reference's constructor (``FactorFields(cfg, device)``),
config keys, attribute names, parameter names and shapes (``coeffs.*``,
``basises.*``, ``linear_mat.backbone.*``) so checkpoints and callers carry
over.  Tables are stored channels_last (texel-major memory) so one
interpolation corner is one contiguous channel vector; their shapes are the
reference's.

Supported: basis_type grid | none; coeff_type grid | vm | none (``vm`` with a
grid basis is the TensoRF-style sum of plane*line products, the combination
the reference cannot evaluate); basis_mapping sawtooth | triangle | sinc | x;
modes bilinear | nearest; linear_mat with 1 or 2 layers.  Anything else raises
NotImplementedError -- there is no PyTorch or CPU fallback.
"""

import ctypes

import numpy as np
import torch
import torch.nn as nn

from . import _lib as L
from .config import AttrDict, plain_cfg
from . import ops
from .ops import FieldSpec, is_texel_major, mlp_desc, texel_major
from .params import N_to_reso, dct_basis, setup_params, vm_reso
from .render import AlphaGridMask, RenderMixin

MAT_MODE = [[0, 1], [0, 2], [1, 2]]   # models/FactorFields.py:213
VEC_MODE = [2, 1, 0]                  # models/FactorFields.py:214


class MLPMixer(nn.Module):
    """linear_mat (models/FactorFields.py:113-159) with pe = 0; same state_dict
    keys (backbone.{l}.weight / .bias).  With with_dropout, training calls
    (is_train=True) drop input entries with p = 0.1 (:150-151) on the device;
    ``keep`` (bool / 0-1, the shape of x) injects the mask instead of drawing
    it."""

    def __init__(self, in_dim, out_dim=16, num_layers=2, hidden_dim=64, pe=0, with_dropout=False):
        super().__init__()
        if pe:
            raise NotImplementedError("linear_mat positional encoding (pe > 0) has no HIP kernel")
        if num_layers not in (1, 2):
            raise NotImplementedError("linear_mat supports num_layers 1 or 2")
        self.with_dropout = with_dropout
        self.in_dim = in_dim
        self.num_layers = num_layers
        self.hidden_dim = hidden_dim
        self.pe = pe
        layers = []
        for l in range(num_layers):
            li = in_dim if l == 0 else hidden_dim
            lo, bias = (out_dim, False) if l == num_layers - 1 else (hidden_dim, True)
            layers.append(nn.Linear(li, lo, bias=bias))
        self.backbone = nn.ModuleList(layers)

    def params_list(self):
        if self.num_layers == 1:
            return [self.backbone[0].weight]
        return [self.backbone[0].weight, self.backbone[0].bias, self.backbone[1].weight]

    dropout_p = 0.1

    def forward(self, x, is_train=False, keep=None):
        if self.with_dropout and is_train:
            # a fresh draw per call, seeded from torch's host generator (no device sync)
            seed = int(torch.randint(0, 2 ** 62, (1,)).item()) if keep is None else 0
            x = ops.dropout(x, self.dropout_p, seed, keep)
        return ops.linear_mat(x, self.params_list())


class FactorFields(RenderMixin, nn.Module):
    def __init__(self, cfg, device="cuda", table_dtype="fp32"):
        super().__init__()
        cfg = AttrDict.wrap(cfg)
        self.cfg = cfg
        self.device = torch.device(device)
        self.matMode = MAT_MODE
        self.vecMode = VEC_MODE
        self.n_scene, self.scene_idx = 1, 0
        self.alphaMask = None
        self.coeff_type, self.basis_type = cfg.model.coeff_type, cfg.model.basis_type
        if self.basis_type not in ("grid", "none"):
            raise NotImplementedError(f"basis_type {self.basis_type!r} has no HIP kernel")
        if self.coeff_type not in ("grid", "vm", "none"):
            raise NotImplementedError(f"coeff_type {self.coeff_type!r} has no HIP kernel")
        if table_dtype not in ("fp32", "bf16"):
            raise ValueError(table_dtype)
        self.table_dtype = table_dtype
        self.setup_params(cfg.dataset.aabb)
        if self.coeff_type != "none":
            self.coeffs = self.init_coef()
        if self.basis_type != "none":
            self.basises = self.init_basis()
        out_dim = cfg.model.out_dim
        in_dim = sum(cfg.model.basis_dims)
        self.linear_mat = MLPMixer(in_dim, out_dim, num_layers=cfg.model.num_layers,
                                   hidden_dim=cfg.model.hidden_dim,
                                   with_dropout=cfg.model.get("with_dropout", False)).to(self.device)
        self._shadows = None
        if table_dtype == "bf16":
            self.refresh_shadows()
        self._build_spec()
        if "reconstruction" in cfg.defaults.mode:
            self._init_render()

    # ------------------------------------------------ models/FactorFields.py:262-307
    def setup_params(self, aabb):
        s = setup_params(self.cfg, aabb)
        self.in_dim = s.in_dim
        self.aabb = s.aabb.to(self.device)
        self.basis_dims = list(self.cfg.model.basis_dims)
        self.basis_reso = s.basis_reso
        self.coeff_reso = s.coeff_reso
        self.freq_bands = s.freq_bands.to(self.device)
        self.T_basis, self.T_coeff = s.T_basis, s.T_coeff
        self.n_scene = s.n_scene
        self._shapes = s

    # ------------------------------------------------ models/FactorFields.py:311-332
    def init_coef(self):
        C = sum(self.basis_dims)
        init = float(self.cfg.model.coef_init)
        if self.coeff_type == "grid":
            t = init * torch.ones((1, C, *self.coeff_reso), device=self.device)
            return nn.ParameterList([nn.Parameter(texel_major(t))])
        # vm: planes xy, xz, yz then lines z, y, x
        r = vm_reso(self.cfg, self._shapes.aabb)
        self.vm_reso = r
        out = []
        for m in range(3):
            a0, a1 = MAT_MODE[m]
            out.append(nn.Parameter(texel_major(init * torch.ones((1, C, r[a1], r[a0]), device=self.device))))
        for m in range(3):
            out.append(nn.Parameter(texel_major(init * torch.ones((1, C, r[VEC_MODE[m]], 1),
                                                                 device=self.device))))
        return nn.ParameterList(out)

    # ------------------------------------------------ models/FactorFields.py:398-423
    def init_basis(self):
        out = []
        D = self.in_dim
        for d, reso in zip(self.basis_dims, self.basis_reso):
            b = dct_basis(d, reso, D).reshape([1, d] + [reso] * D).to(self.device)
            out.append(nn.Parameter(texel_major(b)))
        return nn.ParameterList(out)

    # ------------------------------------------------ kernel description
    def _tables(self):
        # The ParameterLists' own dicts hold the tables in index order;
        # iterating a ParameterList goes through its string-keyed __getitem__
        # (about 17 us per call here), and every training step's host code
        # asks for the tables many times.
        b = list(self.basises._parameters.values()) if self.basis_type != "none" else []
        c = list(self.coeffs._parameters.values()) if self.coeff_type != "none" else []
        return b, c

    def refresh_shadows(self):
        """(Re)build the BF16 copies the kernels gather from (BF16 mode), and
        remember which master tensors, at which versions, they were cast from."""
        b, c = self._tables()
        self._shadows = []
        for t in b + c:
            if not is_texel_major(t):
                raise RuntimeError("table lost its texel-major layout")
            s = torch.empty(t.shape, dtype=torch.bfloat16, device=t.device,
                            memory_format=torch.channels_last if t.dim() == 4 else torch.channels_last_3d)
            L.check(L.lib().ff_cast_bf16(L.ptr(t), L.ptr(s), t.numel(), L.stream_of(t)), "ff_cast_bf16")
            self._shadows.append(s)
        self._shadow_src = [(t, t._version) for t in b + c]
        self._shadow_map = {id(t): (t, s) for t, s in zip(b + c, self._shadows)}

    def _shadows_current(self, tabs):
        """True while every FP32 master is the tensor, at the version, its BF16
        copy was cast from.  Any in-place write through torch (torch.optim.Adam,
        load_state_dict, copy_) bumps the version.  FusedAdam writes the masters
        and their copies together through raw pointers, which leaves the
        version alone and the copies current."""
        src = getattr(self, "_shadow_src", None)
        if src is None or len(src) != len(tabs):
            return False
        for t, (s, v) in zip(tabs, src):
            if t is not s or t._version != v:
                return False
        return True

    def shadow_of(self, p):
        """The BF16 copy cast from table p, or None.  A table that replaced p
        since the last refresh (shrink, upsample) has no copy until the next
        read refreshes them, so a write through this lookup never lands in a
        copy of another tensor."""
        m = self.__dict__.get("_shadow_map")
        e = m.get(id(p)) if m else None
        return e[1] if e is not None and e[0] is p else None

    def _build_spec(self):
        cfg = self.cfg
        mode = cfg.defaults.mode
        D = self.in_dim
        a0 = self.aabb[0, :D].cpu()
        a1 = self.aabb[1, :D].cpu()
        aabbSize = max(a1 - a0)                                    # FactorFields.py:12
        map_scale = (aabbSize[..., None] / self.freq_bands.cpu()).tolist()
        full = self.aabb.cpu()
        inv = (2.0 / (full[1] - full[0])).tolist()                 # FactorFields.py:636
        origin = full[0].tolist()
        if self.coeff_type == "grid":
            ck, cnd = L.COEFF_GRID, len(self.coeff_reso)
            reso_xyz = list(self.coeff_reso)[::-1]
        elif self.coeff_type == "vm":
            ck, cnd = L.COEFF_VM, 3
            reso_xyz = list(self.vm_reso)
        else:
            ck, cnd = L.COEFF_NONE, 0
            reso_xyz = [1, 1, 1]
        x_dim = len(origin) if self.coeff_type != "none" else D
        if mode == "images":
            x_dim = 3
        self._spec = FieldSpec(
            in_dim=D, x_dim=x_dim, basis_dims=self.basis_dims, basis_reso=self.basis_reso,
            mapping=L.MAP[cfg.model.basis_mapping],
            basis_kind=L.BASIS_GRID if self.basis_type == "grid" else L.BASIS_NONE,
            basis_interp=L.INTERP[cfg.model.basis_mode],
            coeff_kind=ck, coeff_interp=L.INTERP[cfg.model.coef_mode], coeff_ndim=cnd,
            coeff_reso_xyz=reso_xyz, map_origin=a0.tolist(), map_scale=map_scale,
            coeff_origin=origin, coeff_inv=inv,
            table_dtype=L.BF16 if self.table_dtype == "bf16" else L.F32)
        self._fused = None

    def _read_tables(self):
        b, c = self._tables()
        if self.table_dtype == "bf16":
            if not self._shadows_current(b + c):
                self.refresh_shadows()
            return list(self._shadows)
        for t in b + c:
            if not is_texel_major(t):
                raise RuntimeError("table lost its texel-major layout")
        return b + c

    def fused_supported(self):
        if self._fused is None:
            b, c = self._tables()
            d = self._spec.desc(b, c)
            m = mlp_desc(self._layers())
            self._fused = bool(L.lib().ff_field_fused_supported(ctypes.byref(d), ctypes.byref(m)))
        return self._fused

    def _layers(self):
        p = self.linear_mat.params_list()
        return [(p[0], None)] if len(p) == 1 else [(p[0], p[1]), (p[2], None)]

    # ------------------------------------------------ models/FactorFields.py:425-533
    def _coding(self, x, part=None):
        """(feats, coeff, basis) of get_coding; part 'coeff' / 'basis' gathers
        only that part's tables (the other outputs are then that part too)."""
        b, c = self._tables()
        rt = self._read_tables()
        nb = len(b)
        if part == "coeff" and nb and c:
            return ops.get_coding(self._spec.part("coeff"), rt[nb:], x, c)
        if part == "basis" and nb and c and self._spec.x_dim == self._spec.in_dim:
            # (image-set inputs carry the image index, which only the coefficient reads)
            return ops.get_coding(self._spec.part("basis"), rt[:nb], x, b)
        return ops.get_coding(self._spec, rt, x, b + c)

    def get_coeff(self, xyz_sampled):
        return self._coding(xyz_sampled, "coeff")[1]

    def get_basis(self, x):
        return self._coding(x, "basis")[2]

    def get_coding(self, x):
        feats, coeff, _ = self._coding(x)
        return feats, coeff

    def normalize_coord(self, xyz_sampled):
        invaabbSize = 2.0 / (self.aabb[1] - self.aabb[0])
        return (xyz_sampled - self.aabb[0]) * invaabbSize - 1

    def field(self, x):
        """linear_mat(get_coding(x)[0]) in one fused kernel (falls back to the
        two HIP kernels when the fused path has no instantiation)."""
        if not self.fused_supported():
            return self.linear_mat(self.get_coding(x)[0])
        b, c = self._tables()
        p = self.linear_mat.params_list()
        return ops.field(self._spec, self._read_tables(), x, p, b + c)

    def basis2density(self, density_features):
        """models/FactorFields.py:639-643."""
        r = self.cfg.renderer
        if r.fea2denseAct == "softplus":
            return nn.functional.softplus(density_features + r.density_shift)
        return nn.functional.relu(density_features + r.density_shift)

    # ------------------------------------------------ fused training steps
    # Gradient copies the fused training steps add into (one per XCD when 8).
    # Table-gradient atomics are bounded by the chip-wide cache-line request
    # rate, which copies do not raise, while every extra copy costs the
    # optimizer a read and a clear of all table gradients: one copy by default.
    grad_replicas = 1

    def _ensure_grads(self):
        for p in self.parameters():
            if p.grad is None:
                p.grad = torch.zeros_like(p, memory_format=torch.preserve_format)

    def grad_arena(self, replicas=None):
        """FP32 arena [replicas, T] holding every table gradient (T = sum of the
        table sizes, each padded to 64 floats).  The replica-0 views become the
        tables' ``.grad``; the fused training kernels add into the replica of
        the XCD they run on.  Returns (arena, T, replicas)."""
        R = self.grad_replicas if replicas is None else replicas
        b, c = self._tables()
        tabs = b + c
        st = getattr(self, "_arena", None)
        ok = (st is not None and st[2] == R and len(st[3]) == len(tabs)
              and all(t.grad is not None and t.grad.data_ptr() == st[0].data_ptr() + 4 * off
                      for t, off in zip(tabs, st[3])))
        if not ok:
            offs, T = [], 0
            for t in tabs:
                offs.append(T)
                T += (t.numel() + 63) // 64 * 64
            arena = torch.zeros((R, T), device=self.device, dtype=torch.float32)
            for t, off in zip(tabs, offs):
                view = torch.as_strided(arena, t.shape, t.stride(), off)
                if t.grad is not None:
                    view.copy_(t.grad)
                t.grad = view
                # FusedAdam sums the copies only while .grad is still this view
                t._ff_rep = (R, T, view.data_ptr())
            st = (arena, T, R, offs)
            self._arena = st
        return st[0], st[1], st[2]

    def reduce_grad_replicas(self):
        """Sum the per-XCD gradient copies into the tables' ``.grad`` (needed
        before an optimizer other than FusedAdam reads them)."""
        st = getattr(self, "_arena", None)
        if st is None or st[2] <= 1:
            return
        arena, T, R, _ = st
        L.check(L.lib().ff_reduce_replicas(L.ptr(arena), R, T, T, L.stream_of(arena)), "ff_reduce_replicas")

    def _train_grads(self):
        """Grads struct for the fused training kernels (tables in the arena)."""
        self._ensure_grads()
        arena, T, R = self.grad_arena()
        b, c = self._tables()
        p = self.linear_mat.params_list()
        g = L.Grads()
        for l, t in enumerate(b):
            g.basis[l] = t.grad.data_ptr()
        for k, t in enumerate(c):
            g.coeff[k] = t.grad.data_ptr()
        g.w0 = p[0].grad.data_ptr()
        if len(p) == 3:
            g.b0, g.w1 = p[1].grad.data_ptr(), p[2].grad.data_ptr()
        g.replicas, g.replica_stride = R, T
        return g

    def regression_step(self, x, target, loss="mse", loss_scale=1.0, loss_sum=None, reduce=True):
        """Forward, loss and backward of one regression batch in one kernel
        (scripts/sdf_regression.ipynb:369-379).  Gradients are accumulated
        into each parameter's ``.grad``; with reduce=False the table
        gradients stay split over the per-XCD copies for FusedAdam to sum.
        Returns the device scalar holding sum(err^2) (mse) or sum|err| (l1)."""
        if not self.fused_supported():
            raise NotImplementedError("no fused kernel for this configuration; use field() + autograd")
        g = self._train_grads()
        d = self._spec.desc(*self._split(self._read_tables()))
        m = mlp_desc(self._layers())
        x = x.contiguous()
        target = target.contiguous()
        n = x.shape[0]
        ws_bytes = L.lib().ff_field_bwd_workspace_bytes(ctypes.byref(d), ctypes.byref(m), n)
        ws = self._workspace(ws_bytes)
        if loss_sum is None:
            loss_sum = torch.zeros((), device=x.device, dtype=torch.float32)
        L.check(L.lib().ff_field_train_regression(ctypes.byref(d), ctypes.byref(m), L.ptr(x),
                                                  L.ptr(target), n, L.LOSS[loss], float(loss_scale),
                                                  ctypes.byref(g), L.ptr(loss_sum), L.ptr(ws),
                                                  ws_bytes, L.stream_of(x)),
                "ff_field_train_regression")
        if reduce:
            self.reduce_grad_replicas()
        return loss_sum

    def _split(self, tabs):
        nb = self._spec.n_basis
        return tabs[:nb], tabs[nb:]

    def _workspace(self, nbytes):
        ws = getattr(self, "_ws", None)
        if ws is None or ws.numel() * 4 < nbytes:
            ws = torch.empty((max(nbytes, 4) + 3) // 4, device=self.device, dtype=torch.float32)
            self._ws = ws
        return ws

    # ------------------------------------------------ models/FactorFields.py:535-554
    def n_parameters(self):
        total = sum(p.numel() for p in self.parameters())
        if "fix" in self.cfg.model.basis_type:
            total -= self.T_basis
        return total

    def get_optparam_groups(self, lr_small=0.001, lr_large=0.02):
        t = self.cfg.training
        groups = []
        if t.get("linear_mat", True):
            groups += [{"params": self.linear_mat.parameters(), "lr": lr_small}]
        if self.coeff_type != "none" and t.get("coeff", True):
            groups += [{"params": self.coeffs.parameters(), "lr": lr_large}]
        if "fix" not in self.basis_type and self.basis_type != "none" and t.get("basis", True):
            groups += [{"params": self.basises.parameters(), "lr": lr_large}]
        if "reconstruction" in self.cfg.defaults.mode and t.get("renderModule", True):
            rm = getattr(self, "renderModule", None)
            if rm is not None and len(list(rm.parameters())):
                groups += [{"params": rm.parameters(), "lr": lr_small}]
        if getattr(self, "proposal", None) is not None:
            groups += [{"params": [self.proposal.grid], "lr": lr_large},
                       {"params": self.proposal.mlp.parameters(), "lr": lr_small}]
        return groups

    # ------------------------------------------------ models/FactorFields.py:669-691
    def save(self, path):
        """The reference's checkpoint: ``state_dict`` (its keys and shapes
        only), ``cfg`` as plain nested dicts and lists, and the packed
        occupancy mask.  A learned proposal's state goes beside it under
        ``proposal`` (its state_dict and constructor arguments), so the
        reference's strict load_state_dict reads the file.  The reference's
        cal_mean_coef (:645-667) averages the coefficients of several scenes;
        for the one scene here it is the identity and is not applied."""
        ckpt = {"state_dict": {k: v.detach().contiguous().cpu() for k, v in self.state_dict().items()
                               if not k.startswith("proposal.")},
                "cfg": plain_cfg(self.cfg)}
        if self.alphaMask is not None:
            vol = self.alphaMask.alpha_volume.bool().cpu().numpy()
            ckpt["alphaMask.shape"] = vol.shape
            ckpt["alphaMask.mask"] = np.packbits(vol.reshape(-1))
            ckpt["alphaMask.aabb"] = self.alphaMask.aabb.cpu()
        pr = getattr(self, "proposal", None)
        if pr is not None:
            ckpt["proposal"] = {"state_dict": {k: v.detach().contiguous().cpu() for k, v in pr.state_dict().items()},
                                "args": pr.init_args()}
        torch.save(ckpt, path)

    def load(self, ckpt):
        """Strict, like the reference's load_state_dict (:686): a missing or
        unexpected key, or a shape that differs, raises before any value is
        copied.  A ``proposal`` entry (written by save) restores the learned
        proposal, attaching one when the model has none."""
        sd = ckpt["state_dict"]
        own = {k: v for k, v in self.state_dict().items() if not k.startswith("proposal.")}
        missing, unexpected = sorted(set(own) - set(sd)), sorted(set(sd) - set(own))
        if missing or unexpected:
            raise RuntimeError(f"checkpoint state_dict does not match the model: missing {missing}, "
                               f"unexpected {unexpected}")
        bad = [k for k in own if tuple(own[k].shape) != tuple(sd[k].shape)]
        if bad:
            raise RuntimeError("checkpoint shapes differ: " + ", ".join(
                f"{k} {tuple(sd[k].shape)} vs {tuple(own[k].shape)}" for k in bad))
        if "alphaMask.aabb" in ckpt:
            length = int(np.prod(ckpt["alphaMask.shape"]))
            vol = torch.from_numpy(np.unpackbits(ckpt["alphaMask.mask"])[:length].reshape(
                ckpt["alphaMask.shape"]))
            self.alphaMask = AlphaGridMask(self.device, ckpt["alphaMask.aabb"].to(self.device),
                                           vol.float().to(self.device))
        with torch.no_grad():
            for k, v in sd.items():
                own[k].copy_(v)
        if "proposal" in ckpt:
            pr = getattr(self, "proposal", None)
            if pr is None:
                pr = self.enable_proposal(**ckpt["proposal"]["args"])
            pr.load_state_dict(ckpt["proposal"]["state_dict"], strict=True)
            if pr.table_dtype == "bf16":
                pr.refresh_shadow()
        if self.table_dtype == "bf16":
            self.refresh_shadows()
        if "reconstruction" in self.cfg.defaults.mode:
            volumeSize = N_to_reso(self.cfg.training.volume_resoFinal ** self.in_dim, self.aabb.cpu())
            self.update_renderParams(volumeSize)
