"""torch.autograd.Functions over the C ABI.  Every forward has its hand-written
backward in libffields.so; PyTorch provides only memory and streams.

Tables are passed to the kernels in texel-major layout: each parameter keeps
the reference's shape ``[1, C, (D,) H, W]`` but is stored channels_last, so
its memory is ``[(D,) H, W, C]``.
"""

import ctypes

import torch

from . import _lib as L


def _f32c(t):
    if t.dtype != torch.float32:
        raise TypeError(f"expected float32, got {t.dtype}")
    return t.contiguous()


def texel_major(t):
    """The channels_last view (texel-major memory) of a [1, C, ...] table."""
    if t.dim() == 4:
        return t.contiguous(memory_format=torch.channels_last)
    if t.dim() == 5:
        return t.contiguous(memory_format=torch.channels_last_3d)
    raise ValueError(f"table must be 4-D or 5-D, got {tuple(t.shape)}")


def is_texel_major(t):
    if t.dim() == 4:
        return t.is_contiguous(memory_format=torch.channels_last)
    if t.dim() == 5:
        return t.is_contiguous(memory_format=torch.channels_last_3d)
    return False


class FieldSpec:
    """Static description of a field (everything in ff_field_desc except the
    table pointers)."""

    def __init__(self, *, in_dim, x_dim, basis_dims, basis_reso, mapping, basis_kind, basis_interp,
                 coeff_kind, coeff_interp, coeff_ndim, coeff_reso_xyz, map_origin, map_scale,
                 coeff_origin, coeff_inv, table_dtype=L.F32):
        self.in_dim, self.x_dim = in_dim, x_dim
        self.basis_dims = list(basis_dims)
        self.basis_reso = list(basis_reso)
        self.C = sum(self.basis_dims)
        self.mapping, self.basis_kind, self.basis_interp = mapping, basis_kind, basis_interp
        self.coeff_kind, self.coeff_interp, self.coeff_ndim = coeff_kind, coeff_interp, coeff_ndim
        self.coeff_reso_xyz = list(coeff_reso_xyz)
        self.map_origin, self.map_scale = list(map_origin), list(map_scale)
        self.coeff_origin, self.coeff_inv = list(coeff_origin), list(coeff_inv)
        self.table_dtype = table_dtype
        self.n_basis = len(self.basis_dims) if basis_kind == L.BASIS_GRID else 0
        self.n_coeff = {L.COEFF_NONE: 0, L.COEFF_GRID: 1, L.COEFF_VM: 6}[coeff_kind]

    def desc(self, basis_tabs, coeff_tabs):
        d = L.FieldDesc()
        d.in_dim, d.x_dim = self.in_dim, self.x_dim
        d.n_levels = len(self.basis_dims)
        d.n_channels = self.C
        d.mapping, d.basis_kind, d.basis_interp = self.mapping, self.basis_kind, self.basis_interp
        d.coeff_kind, d.coeff_interp, d.coeff_ndim = self.coeff_kind, self.coeff_interp, self.coeff_ndim
        d.table_dtype = self.table_dtype
        for l, (dd, r) in enumerate(zip(self.basis_dims, self.basis_reso)):
            d.basis_dims[l], d.basis_reso[l] = dd, r
        for a, r in enumerate(self.coeff_reso_xyz):
            d.coeff_reso[a] = r
        for a, v in enumerate(self.map_origin):
            d.map_origin[a] = v
        for l, v in enumerate(self.map_scale):
            d.map_scale[l] = v
        for a, v in enumerate(self.coeff_origin):
            d.coeff_origin[a] = v
        for a, v in enumerate(self.coeff_inv):
            d.coeff_inv[a] = v
        for l, t in enumerate(basis_tabs):
            d.basis[l] = t.data_ptr()
        for k, t in enumerate(coeff_tabs):
            d.coeff[k] = t.data_ptr()
        return d


def mlp_desc(layers):
    """layers: list of (weight [out, in], bias or None) of MLPMixer.backbone."""
    m = L.MlpDesc()
    m.n_layers = len(layers)
    w0, b0 = layers[0]
    m.in_dim = w0.shape[1]
    if len(layers) == 1:
        m.hidden, m.out_dim = 0, w0.shape[0]
        m.w0 = w0.data_ptr()
    elif len(layers) == 2:
        w1 = layers[1][0]
        m.hidden, m.out_dim = w0.shape[0], w1.shape[0]
        m.w0, m.b0, m.w1 = w0.data_ptr(), b0.data_ptr(), w1.data_ptr()
    else:
        raise NotImplementedError("linear_mat with num_layers > 2 has no HIP kernel")
    return m


def _grad_like(t, needed):
    if not needed:
        return None
    return torch.zeros_like(t, memory_format=torch.preserve_format)


def _grads_struct(basis_g, coeff_g, w0=None, b0=None, w1=None):
    g = L.Grads()
    for l, t in enumerate(basis_g):
        g.basis[l] = 0 if t is None else t.data_ptr()
    for k, t in enumerate(coeff_g):
        g.coeff[k] = 0 if t is None else t.data_ptr()
    g.w0 = 0 if w0 is None else w0.data_ptr()
    g.b0 = 0 if b0 is None else b0.data_ptr()
    g.w1 = 0 if w1 is None else w1.data_ptr()
    return g


def _scratch_grads(spec, tables, needs):
    """FP32 gradient buffers for every table (kernels need a destination for
    all of them); entries whose input does not need a grad are dropped later."""
    out = []
    for t in tables:
        out.append(torch.zeros_like(t, dtype=torch.float32, memory_format=torch.preserve_format))
    return out


class GetCoding(torch.autograd.Function):
    """models/FactorFields.py:523-533: returns (feats, coeff, basis)."""

    @staticmethod
    def forward(ctx, spec, read_tables, x, *tables):
        x = _f32c(x)
        n = x.shape[0]
        nb = spec.n_basis
        d = spec.desc(read_tables[:nb], read_tables[nb:])
        feats = torch.empty((n, spec.C), device=x.device, dtype=torch.float32)
        coeff = torch.empty_like(feats)
        basis = torch.empty_like(feats)
        L.check(L.lib().ff_get_coding(ctypes.byref(d), L.ptr(x), n, L.ptr(feats), L.ptr(coeff),
                                      L.ptr(basis), L.stream_of(x)), "ff_get_coding")
        ctx.spec = spec
        ctx.read_tables = read_tables
        ctx.save_for_backward(x, *tables)
        return feats, coeff, basis

    @staticmethod
    def backward(ctx, d_feats, d_coeff, d_basis):
        x, *tables = ctx.saved_tensors
        spec = ctx.spec
        nb = spec.n_basis
        needs = ctx.needs_input_grad[3:]
        if not any(needs):
            return (None, None, None) + (None,) * len(tables)
        grads = _scratch_grads(spec, tables, needs)
        d = spec.desc(ctx.read_tables[:nb], ctx.read_tables[nb:])
        g = _grads_struct(grads[:nb], grads[nb:])
        n = x.shape[0]
        df = None if d_feats is None else _f32c(d_feats)
        dc = None if d_coeff is None else _f32c(d_coeff)
        db = None if d_basis is None else _f32c(d_basis)
        L.check(L.lib().ff_get_coding_bwd(ctypes.byref(d), L.ptr(x), n, L.ptr(df), L.ptr(dc), L.ptr(db),
                                          ctypes.byref(g), L.stream_of(x)), "ff_get_coding_bwd")
        return (None, None, None) + tuple(gr if nd else None for gr, nd in zip(grads, needs))


class LinearMat(torch.autograd.Function):
    """models/FactorFields.py:144-159 (MLPMixer.forward, pe = 0, no dropout)."""

    @staticmethod
    def forward(ctx, x, *params):
        x = _f32c(x)
        layers = _layers_from(params)
        m = mlp_desc(layers)
        n = x.shape[0]
        out = torch.empty((n, m.out_dim), device=x.device, dtype=torch.float32)
        L.check(L.lib().ff_linear_mat_fwd(ctypes.byref(m), L.ptr(x), n, L.ptr(out), L.stream_of(x)),
                "ff_linear_mat_fwd")
        ctx.save_for_backward(x, *params)
        return out

    @staticmethod
    def backward(ctx, d_out):
        x, *params = ctx.saved_tensors
        layers = _layers_from(params)
        m = mlp_desc(layers)
        n = x.shape[0]
        d_out = _f32c(d_out)
        gp = [torch.zeros_like(p) for p in params]
        gw0, gb0, gw1 = _split_layer_grads(gp)
        dx = torch.empty_like(x) if ctx.needs_input_grad[0] else None
        ws_bytes = L.lib().ff_linear_mat_bwd_workspace_bytes(ctypes.byref(m), n)
        ws = torch.empty((max(ws_bytes, 4) // 4,), device=x.device, dtype=torch.float32)
        g = _grads_struct([], [], gw0, gb0, gw1)
        L.check(L.lib().ff_linear_mat_bwd(ctypes.byref(m), L.ptr(x), n, L.ptr(d_out), L.ptr(dx),
                                          ctypes.byref(g), L.ptr(ws), ws_bytes, L.stream_of(x)),
                "ff_linear_mat_bwd")
        return (dx,) + tuple(gp)


def _layers_from(params):
    """params = (w0,) | (w0, b0, w1)."""
    if len(params) == 1:
        return [(params[0], None)]
    return [(params[0], params[1]), (params[2], None)]


def _split_layer_grads(gp):
    if len(gp) == 1:
        return gp[0], None, None
    return gp[0], gp[1], gp[2]


class Field(torch.autograd.Function):
    """Fused linear_mat(get_coding(x)[0]) with its fused backward."""

    @staticmethod
    def forward(ctx, spec, read_tables, x, n_mlp, *tensors):
        x = _f32c(x)
        params, tables = tensors[:n_mlp], tensors[n_mlp:]
        m = mlp_desc(_layers_from(params))
        nb = spec.n_basis
        d = spec.desc(read_tables[:nb], read_tables[nb:])
        n = x.shape[0]
        out = torch.empty((n, m.out_dim), device=x.device, dtype=torch.float32)
        L.check(L.lib().ff_field_fwd(ctypes.byref(d), ctypes.byref(m), L.ptr(x), n, L.ptr(out),
                                     L.stream_of(x)), "ff_field_fwd")
        ctx.spec, ctx.read_tables, ctx.n_mlp = spec, read_tables, n_mlp
        ctx.save_for_backward(x, *tensors)
        return out

    @staticmethod
    def backward(ctx, d_out):
        x, *tensors = ctx.saved_tensors
        spec, n_mlp = ctx.spec, ctx.n_mlp
        params, tables = tensors[:n_mlp], tensors[n_mlp:]
        m = mlp_desc(_layers_from(params))
        nb = spec.n_basis
        d = spec.desc(ctx.read_tables[:nb], ctx.read_tables[nb:])
        n = x.shape[0]
        gp = [torch.zeros_like(p) for p in params]
        gw0, gb0, gw1 = _split_layer_grads(gp)
        gt = _scratch_grads(spec, tables, None)
        g = _grads_struct(gt[:nb], gt[nb:], gw0, gb0, gw1)
        ws_bytes = L.lib().ff_field_bwd_workspace_bytes(ctypes.byref(d), ctypes.byref(m), n)
        ws = torch.empty((max(ws_bytes, 4) // 4,), device=x.device, dtype=torch.float32)
        L.check(L.lib().ff_field_bwd(ctypes.byref(d), ctypes.byref(m), L.ptr(x), n, L.ptr(_f32c(d_out)),
                                     ctypes.byref(g), L.ptr(ws), ws_bytes, L.stream_of(x)),
                "ff_field_bwd")
        needs = ctx.needs_input_grad
        out = [None, None, None, None]
        for k, gr in enumerate(gp + gt):
            out.append(gr if needs[4 + k] else None)
        return tuple(out)


class Dropout(torch.autograd.Function):
    """torch.nn.functional.dropout(x, p) as linear_mat applies it to its input
    (models/FactorFields.py:150-151): out = x * keep / (1 - p).  The mask is
    `keep` (0/1 uint8, the shape of x) when given, else drawn on the device by
    ff_dropout's counter-based generator from `seed`; the backward applies
    the same mask to the output gradient (regenerated from the seed, not
    stored)."""

    @staticmethod
    def forward(ctx, x, p, seed, keep=None):
        x = _f32c(x)
        out = torch.empty_like(x)
        k = None if keep is None else keep.to(torch.uint8).contiguous()
        L.check(L.lib().ff_dropout(L.ptr(x), x.numel(), float(p), L.ptr(k), int(seed), 0, L.ptr(out), None,
                                   L.stream_of(x)), "ff_dropout")
        ctx.p, ctx.seed, ctx.keep = p, seed, k
        return out

    @staticmethod
    def backward(ctx, d_out):
        d_out = _f32c(d_out)
        dx = torch.empty_like(d_out)
        L.check(L.lib().ff_dropout(L.ptr(d_out), d_out.numel(), float(ctx.p), L.ptr(ctx.keep), int(ctx.seed), 0,
                                   L.ptr(dx), None, L.stream_of(d_out)), "ff_dropout")
        return dx, None, None, None
