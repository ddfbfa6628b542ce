"""The PyTorch custom ops (namespace ffields) on the CPU: they are
registered, a field spec survives its flat-list encoding, and their fake
(meta) implementations give the kernels' output shapes without a GPU -- what
torch.compile and other tracers rely on."""
import torch

import ffields.ops as O
from ffields import FactorFields
from helpers import small_cfg


def test_ops_are_registered():
    for name in ("get_coding", "get_coding_backward", "linear_mat", "linear_mat_backward", "field",
                 "field_backward", "dropout"):
        assert hasattr(torch.ops.ffields, name), name


def test_spec_round_trip():
    for kw in (dict(mode="sdf", D=3, ctype="vm"), dict(mode="image", D=2, ctype="grid")):
        m = FactorFields(small_cfg(**kw), "cpu")
        ints, floats = O.spec_args(m._spec)
        s = O.spec_of(list(ints), list(floats))
        for k in ("in_dim", "x_dim", "basis_dims", "basis_reso", "mapping", "basis_kind", "basis_interp",
                  "coeff_kind", "coeff_interp", "coeff_ndim", "coeff_reso_xyz", "table_dtype", "C"):
            assert getattr(s, k) == getattr(m._spec, k), k
        for k in ("map_origin", "map_scale", "coeff_origin", "coeff_inv"):
            assert [float(v) for v in getattr(s, k)] == [float(v) for v in getattr(m._spec, k)], k


def test_fake_implementations_give_output_shapes():
    m = FactorFields(small_cfg(mode="sdf", D=3, ctype="vm", out_dim=1), "cpu")
    meta = lambda t: torch.empty_like(t, device="meta")      # noqa: E731
    b, c = m._tables()
    tabs = [meta(t) for t in b + c]
    params = [meta(p) for p in m.linear_mat.params_list()]
    x = torch.empty((1000, 3), device="meta")
    si, sf = O.spec_args(m._spec)
    feats, coeff, basis = O.get_coding_op(x, tabs, tabs, si, sf)
    assert feats.shape == coeff.shape == basis.shape == (1000, 18)
    assert O.field_op(x, params, tabs, tabs, si, sf).shape == (1000, 1)
    assert O.linear_mat_op(feats, params).shape == (1000, 1)
    assert O.dropout_op(feats, 0.1, 7, None).shape == (1000, 18)
    g = O.field_backward_op(x, params, tabs, tabs, torch.empty((1000, 1), device="meta"), si, sf)
    assert [t.shape for t in g] == [p.shape for p in params] + [t.shape for t in tabs]
    g = O.get_coding_backward_op(x, tabs, tabs, feats, None, None, si, sf)
    assert [t.shape for t in g] == [t.shape for t in tabs]
    dx, gp = O.linear_mat_backward_op(feats, params, torch.empty((1000, 1), device="meta"))
    assert dx.shape == feats.shape and [t.shape for t in gp] == [p.shape for p in params]


def test_no_cpu_kernel():
    """The ops have a CUDA kernel only: CPU tensors fail loudly."""
    import pytest
    m = FactorFields(small_cfg(mode="sdf", D=3, ctype="grid", out_dim=1), "cpu")
    with pytest.raises(Exception):
        O.linear_mat(torch.zeros((4, 18)), m.linear_mat.params_list())
