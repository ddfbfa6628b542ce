"""Dictionary image-set fast path (image-set mode, basis_dims
[32,32,32,16,16,16] = 144 channels, per-image coefficient grid, decoder
144 -> 64 -> 3 on the matrix cores).  The FP32 form (FP32 tables and decoder
operands) is checked against the FP32 oracle at 1e-5 / 1e-4 (end of file);
the BF16 form against a CPU emulation of its arithmetic: features from the oracle on the BF16-rounded tables (FP32), the
decoder's inputs (features with a constant-1 bias column), weights, hidden
activations, output gradients and hidden gradients rounded to BF16 (round to
nearest even), products accumulated in FP32.

Bars: within 2e-3 of the largest magnitude plus 2e-3 relative, and a median
relative error below 1e-4 for the forward.  A value on a BF16 rounding
boundary can round the other way on one side (one BF16 step of one operand);
a layout or indexing mistake gives errors of the size of the values."""
import sys

import pytest
import torch
import torch.nn.functional as F

from helpers import assert_close, new_model, oracle_of, small_cfg

pytestmark = pytest.mark.gpu

DICT = dict(mode="images", D=2, ctype="grid", dims=(32, 32, 32, 16, 16, 16), resos=(5, 6, 7, 8, 9, 10),
            coeff_reso=4, aabb=[[0.0, 0.0, 0.0], [8.0, 6.0, 5.0]], out_dim=3)


def bf(t):
    return t.to(torch.bfloat16).float()


def image_points(n, seed, W=8, H=6, D=5):
    """Pixel centres and image index + 0.5 (dataLoader/image_set.py:55-60),
    plus a jitter of the pixel position as in its continuous sampling."""
    g = torch.Generator().manual_seed(seed)
    xy = torch.rand((n, 2), generator=g) * torch.tensor([W, H], dtype=torch.float32)
    img = torch.randint(0, D, (n, 1), generator=g).float() + 0.5
    return torch.cat([xy, img], 1)


def emulate(m, feats):
    lin = m.linear_mat.backbone
    w0, b0, w1 = (t.detach().cpu() for t in (lin[0].weight, lin[0].bias, lin[1].weight))
    W0a = bf(torch.cat([w0, b0[:, None]], 1))
    W1 = bf(w1)
    xb = bf(torch.cat([feats, torch.ones((feats.shape[0], 1))], 1))
    pre = (xb.double() @ W0a.double().T).float()
    hb = bf(F.relu(pre))
    out = (hb.double() @ W1.double().T).float()
    return out, pre, hb, xb, W0a, W1


def make():
    print("[case] dict model", file=sys.stderr, flush=True)
    m = new_model(small_cfg(**DICT), dtype="bf16", seed=2)
    assert m.fused_supported()
    return m


def test_dict_forward():
    m = make()
    x = image_points(3001, seed=3)
    out = m.field(x.cuda())
    of = oracle_of(m, bf16_tables=True)
    ref, *_ = emulate(m, of.get_coding(x)[0])
    assert_close(out, ref, 2e-3, "out", atol_frac=2e-3)
    rel = ((out.cpu() - ref).abs() / ref.abs().clamp(min=1e-6)).flatten()
    assert rel.median().item() < 1e-4, rel.median().item()


def repeated_points(n, seed, W=8, H=6, D=5):
    """Pixels drawn from a few pixel centres across all images, ordered by
    position: consecutive pixels of a wave share basis texels (and some
    coefficient texels), so many lanes of one atomic instruction add into the
    same addresses."""
    g = torch.Generator().manual_seed(seed)
    pos = torch.tensor([[0.5, 0.5], [3.5, 2.5], [7.5, 5.5], [4.5, 1.5]])
    k = torch.sort(torch.randint(0, len(pos), (n,), generator=g)).values
    img = torch.randint(0, D, (n, 1), generator=g).float() + 0.5
    return torch.cat([pos[k], img], 1)


def cell_sorted_points(n, seed):
    """image_points ordered by image, then by the 2x2-pixel coefficient cell
    (the 8x6 images on a 4-texel-wide coefficient grid), then by position:
    runs of pixels sharing some or all coefficient corners, of varying length."""
    x = image_points(n, seed)
    cx, cy = torch.floor((x[:, 0] + 1) / 2), torch.floor((x[:, 1] + 1) / 2)
    key = ((x[:, 2].floor() * 8 + cy) * 8 + cx) * 64 + x[:, 1] * 8 + x[:, 0]
    return x[torch.argsort(key)].contiguous()


@pytest.mark.parametrize("how", ["regression_mse", "autograd", "regression_runs", "regression_sorted"])
def test_dict_backward(how):
    m = make()
    n = 2503
    if how == "regression_runs":
        x = repeated_points(n, seed=5)
    elif how == "regression_sorted":
        x = cell_sorted_points(n, seed=5)
    else:
        x = image_points(n, seed=5)
    tgt = torch.rand((n, 3), generator=torch.Generator().manual_seed(6))
    r = torch.randn((n, 3), generator=torch.Generator().manual_seed(7))
    m.zero_grad(set_to_none=True)
    of = oracle_of(m, bf16_tables=True).requires_grad_(True)
    feats = of.get_coding(x)[0]
    out, pre, hb, xb, W0a, W1 = emulate(m, feats.detach())
    if how.startswith("regression"):
        m.regression_step(x.cuda(), tgt.cuda(), loss="mse")
        dout = 2 * (out - tgt) / (n * 3)
    else:
        (m.field(x.cuda()) * r.cuda()).sum().backward()
        dout = r
    db = bf(dout)
    dh = (db.double() @ W1.double()).float() * (pre > 0).float()
    dhb = bf(dh)
    df = (dhb.double() @ W0a.double()).float()[:, :144]
    feats.backward(df)
    lin = m.linear_mat.backbone
    dW1 = (db.double().T @ hb.double()).float()
    dW0a = (dhb.double().T @ xb.double()).float()
    assert_close(lin[1].weight.grad, dW1, 2e-3, "dW1", atol_frac=2e-3)
    assert_close(lin[0].weight.grad, dW0a[:, :144], 2e-3, "dW0", atol_frac=2e-3)
    assert_close(lin[0].bias.grad, dW0a[:, 144], 2e-3, "db0", atol_frac=2e-3)
    for k, (t, o) in enumerate(zip(list(m.basises) + list(m.coeffs), of.basises + of.coeffs)):
        assert_close(t.grad.contiguous(), o.grad, 2e-3, f"table {k}", atol_frac=2e-3)


def test_dict_batch_order_invariance():
    """The same pixels in draw order and ordered by image and coefficient
    cell (the bench's batch order) give the same gradients, up to FP32
    summation order (atomics and per-wave partial sums)."""
    m = make()
    n = 4099
    x = image_points(n, seed=8)
    tgt = torch.rand((n, 3), generator=torch.Generator().manual_seed(9))
    cx, cy = torch.floor((x[:, 0] + 1) / 2), torch.floor((x[:, 1] + 1) / 2)
    perm = torch.argsort(((x[:, 2].floor() * 8 + cy) * 8 + cx) * 64 + x[:, 1] * 8 + x[:, 0])
    params = list(m.basises) + list(m.coeffs) + list(m.linear_mat.parameters())
    grads = []
    for xi, ti in [(x, tgt), (x[perm].contiguous(), tgt[perm].contiguous())]:
        m.zero_grad(set_to_none=True)
        m.regression_step(xi.cuda(), ti.cuda(), loss="mse")
        grads.append([p.grad.detach().clone() for p in params])
    for k, (a, b) in enumerate(zip(*grads)):
        assert_close(a, b, 1e-4, f"param {k}", atol_frac=1e-5)


# ---------------------------------------------------------------- FP32 form
def make_fp32():
    print("[case] dict model, FP32", file=sys.stderr, flush=True)
    m = new_model(small_cfg(**DICT), dtype="fp32", seed=2)
    assert m.fused_supported()       # the dictionary kernels with FP32 tables and FP32 decoder operands
    return m


def test_dict_fp32_forward_matches_oracle():
    """FP32 tables and the FP32 matrix-core decoder against the FP32 oracle
    (models/FactorFields.py:425-533, :144-159) at 1e-5."""
    m = make_fp32()
    x = image_points(3001, seed=13)
    out = m.field(x.cuda())
    ref = oracle_of(m).field(x)
    assert_close(out, ref, 1e-5, "FP32 dictionary forward")


@pytest.mark.parametrize("how", ["regression_mse", "autograd", "regression_sorted"])
def test_dict_fp32_backward_matches_oracle(how):
    """Gradients of every table and decoder weight against the FP32 oracle's
    autograd at 1e-4 (FP64 finite differences pin the oracle)."""
    m = make_fp32()
    n = 2503
    x = cell_sorted_points(n, seed=15) if how == "regression_sorted" else image_points(n, seed=15)
    tgt = torch.rand((n, 3), generator=torch.Generator().manual_seed(16))
    r = torch.randn((n, 3), generator=torch.Generator().manual_seed(17))
    m.zero_grad(set_to_none=True)
    of = oracle_of(m).requires_grad_(True)
    ref = of.field(x)
    if how.startswith("regression"):
        m.regression_step(x.cuda(), tgt.cuda(), loss="mse")
        torch.mean((ref - tgt) ** 2).backward()
    else:
        (m.field(x.cuda()) * r.cuda()).sum().backward()
        (ref * r).sum().backward()
    ours = list(m.basises) + list(m.coeffs) + list(m.linear_mat.parameters())
    theirs = of.basises + of.coeffs + [p for l in of.mlp for p in l if p is not None]
    assert len(ours) == len(theirs)
    for k, (t, o) in enumerate(zip(ours, theirs)):
        assert_close(t.grad.contiguous(), o.grad, 1e-4, f"param {k}", atol_frac=1e-5)
