"""CPU tests of the C ABI: the library loads, exports every symbol declared in
include/*.h, and rejects bad arguments with a status before touching a GPU."""
import ctypes

import pytest

import ffields._lib as L


def test_library_exports_every_declared_symbol():
    lib = L.lib()
    names = L.exported_symbols()
    assert len(names) >= 10
    for n in names:
        assert hasattr(lib, n), n
    assert lib.ff_version() >= 1


def test_status_strings():
    lib = L.lib()
    assert lib.ff_status_str(0) == b"ok"
    assert b"invalid" in lib.ff_status_str(1)
    assert b"unknown" in lib.ff_status_str(99)


def _valid_desc():
    d = L.FieldDesc()
    d.in_dim = d.x_dim = 3
    d.n_levels = 1
    d.n_channels = 4
    d.mapping = 0
    d.basis_kind = L.BASIS_GRID
    d.coeff_kind = L.COEFF_GRID
    d.coeff_ndim = 3
    d.basis_dims[0] = 4
    d.basis_reso[0] = 8
    d.coeff_reso[0] = d.coeff_reso[1] = d.coeff_reso[2] = 4
    d.basis[0] = 0x1000
    d.coeff[0] = 0x2000
    return d


@pytest.mark.parametrize("mutate", [
    lambda d: setattr(d, "in_dim", 4),
    lambda d: setattr(d, "n_channels", 5),
    lambda d: setattr(d, "x_dim", 2),
    lambda d: setattr(d, "mapping", 9),
    lambda d: setattr(d, "basis_interp", 7),
    lambda d: setattr(d, "table_dtype", 3),
    lambda d: d.basis.__setitem__(0, 0),
    lambda d: d.coeff.__setitem__(0, 0),
    lambda d: d.basis_reso.__setitem__(0, 1),
    lambda d: setattr(d, "coeff_kind", L.COEFF_VM),      # VM needs six tables
])
def test_bad_field_desc_rejected(mutate):
    d = _valid_desc()
    mutate(d)
    st = L.lib().ff_get_coding(ctypes.byref(d), ctypes.c_void_p(0x3000), 10, None, None, None, None)
    assert st == L.FF_ERR_ARG


def test_null_arguments_rejected():
    lib = L.lib()
    d = _valid_desc()
    assert lib.ff_get_coding(None, ctypes.c_void_p(0x3000), 10, None, None, None, None) == L.FF_ERR_ARG
    assert lib.ff_get_coding(ctypes.byref(d), None, 10, None, None, None, None) == L.FF_ERR_ARG
    assert lib.ff_get_coding(ctypes.byref(d), None, -1, None, None, None, None) == L.FF_ERR_ARG
    m = L.MlpDesc()
    assert lib.ff_linear_mat_fwd(ctypes.byref(m), None, 0, None, None) == L.FF_ERR_ARG
    m.n_layers, m.in_dim, m.hidden, m.out_dim, m.w0 = 2, 4, 8, 40, 0x10
    assert lib.ff_linear_mat_fwd(ctypes.byref(m), None, 0, None, None) == L.FF_ERR_ARG  # out_dim > 32
    g = L.Grads()
    assert lib.ff_get_coding_bwd(ctypes.byref(d), ctypes.c_void_p(0x3000), 10, None, None, None,
                                 ctypes.byref(g), None) == L.FF_ERR_ARG             # null grad tables
    assert lib.ff_adam_multi(None, 1, 0.9, 0.99, 1e-8, 1, 0, None) == L.FF_ERR_ARG
    t = L.AdamTensor(0, 0, 0, 0, 0, 5, 0.1)
    assert lib.ff_adam_multi(ctypes.byref(t), 1, 0.9, 0.99, 1e-8, 1, 0, None) == L.FF_ERR_ARG
    assert lib.ff_adam_multi(ctypes.byref(t), 1, 0.9, 0.99, 1e-8, 0, 0, None) == L.FF_ERR_ARG


def test_zero_size_is_a_noop():
    d = _valid_desc()
    assert L.lib().ff_get_coding(ctypes.byref(d), None, 0, None, None, None, None) == L.FF_OK


def test_fused_support_query():
    d = _valid_desc()
    m = L.MlpDesc()
    m.n_layers, m.in_dim, m.hidden, m.out_dim, m.w0, m.b0, m.w1 = 2, 4, 64, 3, 0x10, 0x20, 0x30
    assert L.lib().ff_field_fused_supported(ctypes.byref(d), ctypes.byref(m)) == 0   # 4 channels: generic
