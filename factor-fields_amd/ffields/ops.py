"""PyTorch custom ops over the C ABI (``torch.library.custom_op``, namespace
``ffields``): get_coding, linear_mat, field (the fused pair) and dropout,
each with a fake (meta) implementation for tracing and a registered autograd
formula whose backward is itself a custom op.  Every forward has its
hand-written backward in libffields.so; PyTorch provides only memory and
streams.  The ops run on CUDA (HIP) tensors only: there is no CPU kernel.

Tables are passed to the kernels in texel-major layout: each parameter keeps
the reference's shape ``[1, C, (D,) H, W]`` but is stored channels_last, so
its memory is ``[(D,) H, W, C]``.  A field's static description travels as
two flat lists (``spec_args``); the ctypes descriptors built from it and from
the table pointers are cached, so a call costs no struct rebuild.
"""

import ctypes
from typing import List, Optional, Tuple

import torch
from torch import Tensor

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

    def part(self, which):
        """The same field with only its basis ('basis': coefficient 'none') or
        only its coefficient ('coeff': basis 'none'), for get_basis /
        get_coeff: the kernels then gather only that part's tables."""
        key = "_part_" + which
        if getattr(self, key, None) is None:
            kw = dict(in_dim=self.in_dim, x_dim=self.x_dim, basis_dims=self.basis_dims, basis_reso=self.basis_reso,
                      mapping=self.mapping, basis_kind=self.basis_kind, basis_interp=self.basis_interp,
                      coeff_kind=self.coeff_kind, coeff_interp=self.coeff_interp, coeff_ndim=self.coeff_ndim,
                      coeff_reso_xyz=self.coeff_reso_xyz, map_origin=self.map_origin, map_scale=self.map_scale,
                      coeff_origin=self.coeff_origin, coeff_inv=self.coeff_inv, table_dtype=self.table_dtype)
            if which == "basis":
                kw["coeff_kind"] = L.COEFF_NONE
            else:
                kw["basis_kind"] = L.BASIS_NONE
            setattr(self, key, FieldSpec(**kw))
        return getattr(self, key)

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


def _layers_from(params):
    """params = (w0,) | (w0, b0, w1)."""
    if len(params) == 1:
        return [(params[0], None)]
    return [(params[0], params[1]), (params[2], None)]


def _split_layer_grads(gp):
    if len(gp) == 1:
        return gp[0], None, None
    return gp[0], gp[1], gp[2]


def spec_args(spec):
    """(ints, floats): a FieldSpec as the two flat lists the custom ops take."""
    a = getattr(spec, "_args", None)
    if a is None:
        ints = [spec.in_dim, spec.x_dim, spec.mapping, spec.basis_kind, spec.basis_interp, spec.coeff_kind,
                spec.coeff_interp, spec.coeff_ndim, spec.table_dtype, len(spec.basis_dims)]
        ints += spec.basis_dims + spec.basis_reso + [len(spec.coeff_reso_xyz)] + spec.coeff_reso_xyz
        floats = []
        for v in (spec.map_origin, spec.map_scale, spec.coeff_origin, spec.coeff_inv):
            floats += [float(len(v))] + [float(x) for x in v]
        a = spec._args = (ints, floats)
        _SPECS[(tuple(ints), tuple(floats))] = spec
    return a


_SPECS = {}


def spec_of(ints, floats):
    """The FieldSpec of spec_args' lists (cached)."""
    key = (tuple(ints), tuple(floats))
    s = _SPECS.get(key)
    if s is None:
        it = iter(ints)
        in_dim, x_dim, mapping, bk, bi, ck, ci, cnd, td, nl = (next(it) for _ in range(10))
        dims = [next(it) for _ in range(nl)]
        reso = [next(it) for _ in range(nl)]
        cr = [next(it) for _ in range(next(it))]
        ft = iter(floats)
        vec = []
        for _ in range(4):
            vec.append([next(ft) for _ in range(int(next(ft)))])
        s = FieldSpec(in_dim=in_dim, x_dim=x_dim, basis_dims=dims, basis_reso=reso, mapping=mapping,
                      basis_kind=bk, basis_interp=bi, coeff_kind=ck, coeff_interp=ci, coeff_ndim=cnd,
                      coeff_reso_xyz=cr, map_origin=vec[0], map_scale=vec[1], coeff_origin=vec[2],
                      coeff_inv=vec[3], table_dtype=td)
        s._args = (list(ints), list(floats))
        _SPECS[key] = s
    return s


_DESCS = {}


def _cached(cache, key, make):
    v = cache.get(key)
    if v is None:
        if len(cache) > 256:
            cache.clear()
        v = cache[key] = make()
    return v


def _field_desc(spec, read_tables):
    nb = spec.n_basis
    key = (tuple(spec._args[0]), tuple(spec._args[1]), tuple((t.data_ptr(), t.dtype) for t in read_tables))
    return _cached(_DESCS, key, lambda: spec.desc(read_tables[:nb], read_tables[nb:]))


def _mlp(params):
    key = tuple((p.data_ptr(), tuple(p.shape)) for p in params)
    return _cached(_DESCS, ("mlp",) + key, lambda: mlp_desc(_layers_from(params)))


def _opt(t):
    return None if t is None else _f32c(t)


# ---- ffields::get_coding (models/FactorFields.py:523-533): (feats, coeff, basis)
@torch.library.custom_op("ffields::get_coding", mutates_args=(), device_types="cuda")
def get_coding_op(x: Tensor, read_tables: List[Tensor], tables: List[Tensor], spec_i: List[int],
                  spec_f: List[float]) -> Tuple[Tensor, Tensor, Tensor]:
    spec = spec_of(spec_i, spec_f)
    x = _f32c(x)
    n = x.shape[0]
    d = _field_desc(spec, read_tables)
    feats = torch.empty((n, spec.C), device=x.device, dtype=torch.float32)
    coeff = torch.empty_like(feats)
    basis = torch.empty_like(feats)
    L.check(L.lib().ff_get_coding(ctypes.byref(d), L.ptr(x), n, L.ptr(feats), L.ptr(coeff), L.ptr(basis),
                                  L.stream_of(x)), "ff_get_coding")
    return feats, coeff, basis


@get_coding_op.register_fake
def _(x, read_tables, tables, spec_i, spec_f):
    C = spec_of(spec_i, spec_f).C
    out = x.new_empty((x.shape[0], C), dtype=torch.float32)
    return out, torch.empty_like(out), torch.empty_like(out)


@torch.library.custom_op("ffields::get_coding_backward", mutates_args=(), device_types="cuda")
def get_coding_backward_op(x: Tensor, read_tables: List[Tensor], tables: List[Tensor], d_feats: Optional[Tensor],
                           d_coeff: Optional[Tensor], d_basis: Optional[Tensor], spec_i: List[int],
                           spec_f: List[float]) -> List[Tensor]:
    spec = spec_of(spec_i, spec_f)
    nb = spec.n_basis
    grads = _scratch_grads(spec, tables, None)
    d = _field_desc(spec, read_tables)
    g = _grads_struct(grads[:nb], grads[nb:])
    L.check(L.lib().ff_get_coding_bwd(ctypes.byref(d), L.ptr(x), x.shape[0], L.ptr(_opt(d_feats)),
                                      L.ptr(_opt(d_coeff)), L.ptr(_opt(d_basis)), ctypes.byref(g), L.stream_of(x)),
            "ff_get_coding_bwd")
    return grads


@get_coding_backward_op.register_fake
def _(x, read_tables, tables, d_feats, d_coeff, d_basis, spec_i, spec_f):
    return [torch.empty_like(t, dtype=torch.float32, memory_format=torch.preserve_format) for t in tables]


def _gc_setup(ctx, inputs, output):
    x, read_tables, tables, spec_i, spec_f = inputs
    ctx.n_read = len(read_tables)
    ctx.spec = (spec_i, spec_f)
    ctx.save_for_backward(x, *read_tables, *tables)


def _gc_backward(ctx, d_feats, d_coeff, d_basis):
    x, *rest = ctx.saved_tensors
    rt, tb = rest[:ctx.n_read], rest[ctx.n_read:]
    g = get_coding_backward_op(x, rt, tb, d_feats, d_coeff, d_basis, *ctx.spec)
    return None, [None] * len(rt), g, None, None


get_coding_op.register_autograd(_gc_backward, setup_context=_gc_setup)


# ---- ffields::linear_mat (models/FactorFields.py:144-159, MLPMixer.forward without dropout)
@torch.library.custom_op("ffields::linear_mat", mutates_args=(), device_types="cuda")
def linear_mat_op(x: Tensor, params: List[Tensor]) -> Tensor:
    x = _f32c(x)
    m = _mlp(params)
    n = x.shape[0]
    out = torch.empty((n, m.out_dim), device=x.device, dtype=torch.float32)
    L.check(L.lib().ff_linear_mat_fwd(ctypes.byref(m), L.ptr(x), n, L.ptr(out), L.stream_of(x)),
            "ff_linear_mat_fwd")
    return out


@linear_mat_op.register_fake
def _(x, params):
    return x.new_empty((x.shape[0], params[-1].shape[0]), dtype=torch.float32)


@torch.library.custom_op("ffields::linear_mat_backward", mutates_args=(), device_types="cuda")
def linear_mat_backward_op(x: Tensor, params: List[Tensor], d_out: Tensor) -> Tuple[Tensor, List[Tensor]]:
    x = _f32c(x)
    m = _mlp(params)
    n = x.shape[0]
    gp = [torch.zeros_like(p) for p in params]
    gw0, gb0, gw1 = _split_layer_grads(gp)
    dx = torch.empty_like(x)
    ws_bytes = L.lib().ff_linear_mat_bwd_workspace_bytes(ctypes.byref(m), n)
    ws = torch.empty((max(ws_bytes, 4) // 4,), device=x.device, dtype=torch.float32)
    g = _grads_struct([], [], gw0, gb0, gw1)
    L.check(L.lib().ff_linear_mat_bwd(ctypes.byref(m), L.ptr(x), n, L.ptr(_f32c(d_out)), L.ptr(dx), ctypes.byref(g),
                                      L.ptr(ws), ws_bytes, L.stream_of(x)), "ff_linear_mat_bwd")
    return dx, gp


@linear_mat_backward_op.register_fake
def _(x, params, d_out):
    return torch.empty_like(x, dtype=torch.float32), [torch.empty_like(p) for p in params]


def _lm_setup(ctx, inputs, output):
    x, params = inputs
    ctx.save_for_backward(x, *params)


def _lm_backward(ctx, d_out):
    x, *params = ctx.saved_tensors
    dx, gp = linear_mat_backward_op(x, params, d_out)
    return dx, gp


linear_mat_op.register_autograd(_lm_backward, setup_context=_lm_setup)


# ---- ffields::field: linear_mat(get_coding(x)[0]) fused, with its fused backward
@torch.library.custom_op("ffields::field", mutates_args=(), device_types="cuda")
def field_op(x: Tensor, params: List[Tensor], read_tables: List[Tensor], tables: List[Tensor], spec_i: List[int],
             spec_f: List[float]) -> Tensor:
    spec = spec_of(spec_i, spec_f)
    x = _f32c(x)
    m = _mlp(params)
    d = _field_desc(spec, read_tables)
    n = x.shape[0]
    out = torch.empty((n, m.out_dim), device=x.device, dtype=torch.float32)
    L.check(L.lib().ff_field_fwd(ctypes.byref(d), ctypes.byref(m), L.ptr(x), n, L.ptr(out), L.stream_of(x)),
            "ff_field_fwd")
    return out


@field_op.register_fake
def _(x, params, read_tables, tables, spec_i, spec_f):
    return x.new_empty((x.shape[0], params[-1].shape[0]), dtype=torch.float32)


@torch.library.custom_op("ffields::field_backward", mutates_args=(), device_types="cuda")
def field_backward_op(x: Tensor, params: List[Tensor], read_tables: List[Tensor], tables: List[Tensor],
                      d_out: Tensor, spec_i: List[int], spec_f: List[float]) -> List[Tensor]:
    spec = spec_of(spec_i, spec_f)
    x = _f32c(x)
    m = _mlp(params)
    nb = spec.n_basis
    d = _field_desc(spec, read_tables)
    n = x.shape[0]
    gp = [torch.zeros_like(p) for p in params]
    gw0, gb0, gw1 = _split_layer_grads(gp)
    gt = _scratch_grads(spec, tables, None)
    g = _grads_struct(gt[:nb], gt[nb:], gw0, gb0, gw1)
    ws_bytes = L.lib().ff_field_bwd_workspace_bytes(ctypes.byref(d), ctypes.byref(m), n)
    ws = torch.empty((max(ws_bytes, 4) // 4,), device=x.device, dtype=torch.float32)
    L.check(L.lib().ff_field_bwd(ctypes.byref(d), ctypes.byref(m), L.ptr(x), n, L.ptr(_f32c(d_out)),
                                 ctypes.byref(g), L.ptr(ws), ws_bytes, L.stream_of(x)), "ff_field_bwd")
    return gp + gt


@field_backward_op.register_fake
def _(x, params, read_tables, tables, d_out, spec_i, spec_f):
    return ([torch.empty_like(p) for p in params]
            + [torch.empty_like(t, dtype=torch.float32, memory_format=torch.preserve_format) for t in tables])


def _fd_setup(ctx, inputs, output):
    x, params, read_tables, tables, spec_i, spec_f = inputs
    ctx.n = (len(params), len(read_tables))
    ctx.spec = (spec_i, spec_f)
    ctx.save_for_backward(x, *params, *read_tables, *tables)


def _fd_backward(ctx, d_out):
    x, *rest = ctx.saved_tensors
    npar, nrd = ctx.n
    params, rt, tb = rest[:npar], rest[npar:npar + nrd], rest[npar + nrd:]
    g = field_backward_op(x, params, rt, tb, d_out, *ctx.spec)
    return None, g[:npar], [None] * nrd, g[npar:], None, None


field_op.register_autograd(_fd_backward, setup_context=_fd_setup)


# ---- ffields::dropout (models/FactorFields.py:150-151, F.dropout on linear_mat's input)
@torch.library.custom_op("ffields::dropout", mutates_args=(), device_types="cuda")
def dropout_op(x: Tensor, p: float, seed: int, keep: Optional[Tensor]) -> Tensor:
    """x * keep / (1 - p).  The mask is `keep` (0/1, the shape of x) when
    given, else drawn on the device by ff_dropout's counter-based generator
    from `seed`; the backward applies the same mask to the output gradient
    (regenerated from the seed, not stored)."""
    x = _f32c(x)
    out = torch.empty_like(x)
    k = None if keep is None else keep.to(torch.uint8).contiguous()
    L.check(L.lib().ff_dropout(L.ptr(x), x.numel(), float(p), L.ptr(k), int(seed), 0, L.ptr(out), None,
                               L.stream_of(x)), "ff_dropout")
    return out


@dropout_op.register_fake
def _(x, p, seed, keep):
    return torch.empty_like(x, dtype=torch.float32)


def _do_setup(ctx, inputs, output):
    x, p, seed, keep = inputs
    ctx.p, ctx.seed = p, seed
    ctx.save_for_backward(keep if keep is not None else x.new_empty(0))
    ctx.has_keep = keep is not None


def _do_backward(ctx, d_out):
    (keep,) = ctx.saved_tensors
    return dropout_op(d_out, ctx.p, ctx.seed, keep if ctx.has_keep else None), None, None, None


dropout_op.register_autograd(_do_backward, setup_context=_do_setup)


# ---- ffields::sh_render: the degree-2 SH colour head (models/sh.py:34-85 + sigmoid)
@torch.library.custom_op("ffields::sh_render", mutates_args=(), device_types="cuda")
def sh_render_op(viewdirs: Tensor, sh: Tensor) -> Tensor:
    d = _f32c(viewdirs).reshape(-1, 3)
    f = _f32c(sh).reshape(-1, 27)
    n = d.shape[0]
    rgb = torch.empty((n, 3), device=d.device, dtype=torch.float32)
    L.check(L.lib().ff_sh_rgb(L.ptr(d), 3, L.ptr(f), 27, n, L.ptr(rgb), L.stream_of(d)), "ff_sh_rgb")
    return rgb.reshape(*sh.shape[:-1], 3)


@sh_render_op.register_fake
def _(viewdirs, sh):
    return sh.new_empty((*sh.shape[:-1], 3), dtype=torch.float32)


@torch.library.custom_op("ffields::sh_render_backward", mutates_args=(), device_types="cuda")
def sh_render_backward_op(viewdirs: Tensor, sh: Tensor, d_rgb: Tensor) -> Tensor:
    d = _f32c(viewdirs).reshape(-1, 3)
    f = _f32c(sh).reshape(-1, 27)
    g = _f32c(d_rgb).reshape(-1, 3)
    n = d.shape[0]
    d_sh = torch.empty((n, 27), device=d.device, dtype=torch.float32)
    L.check(L.lib().ff_sh_rgb_bwd(L.ptr(d), 3, L.ptr(f), 27, n, L.ptr(g), L.ptr(d_sh), L.stream_of(d)),
            "ff_sh_rgb_bwd")
    return d_sh.reshape(sh.shape)


@sh_render_backward_op.register_fake
def _(viewdirs, sh, d_rgb):
    return torch.empty_like(sh, dtype=torch.float32)


def _sh_setup(ctx, inputs, output):
    ctx.save_for_backward(*inputs)


def _sh_backward(ctx, d_rgb):
    viewdirs, sh = ctx.saved_tensors
    return None, sh_render_backward_op(viewdirs, sh, d_rgb)


sh_render_op.register_autograd(_sh_backward, setup_context=_sh_setup)


# ---- call forms used by the module
def get_coding(spec, read_tables, x, tables):
    """(feats, coeff, basis) of ffields::get_coding."""
    return get_coding_op(x, list(read_tables), list(tables), *spec_args(spec))


def linear_mat(x, params):
    return linear_mat_op(x, list(params))


def field(spec, read_tables, x, params, tables):
    return field_op(x, list(params), list(read_tables), list(tables), *spec_args(spec))


def dropout(x, p, seed, keep=None):
    return dropout_op(x, float(p), int(seed), keep)
