"""GPU parity of the field path (get_coding, linear_mat, fused field, fused
regression step) against the CPU oracle on the same seeded inputs.

Bars: forward within 1e-5 relative (plus an absolute floor of 1e-6 of the
output's max magnitude), gradients within 1e-4 relative (floor 1e-5), as the
FP32 form must meet against the reference's grid_sample semantics."""
import sys

import pytest
import torch

from helpers import assert_close, assert_close_fp64, new_model, oracle_of, rand_points, small_cfg

pytestmark = pytest.mark.gpu

FWD_RTOL = 1e-5
GRAD_RTOL = 1e-4

CASES = {
    # name: cfg kwargs
    "2d_grid_bilinear": dict(mode="image", D=2, ctype="grid"),
    "2d_grid_nearest": dict(mode="image", D=2, ctype="grid", cmode="nearest", bmode="nearest"),
    "2d_triangle": dict(mode="image", D=2, ctype="grid", mapping="triangle"),
    "3d_grid": dict(mode="sdf", D=3, ctype="grid"),
    "3d_vm": dict(mode="sdf", D=3, ctype="vm"),
    "3d_vm_nearest": dict(mode="sdf", D=3, ctype="vm", cmode="nearest"),
    "3d_sinc": dict(mode="sdf", D=3, ctype="grid", mapping="sinc"),
    "3d_x_zeros_padding": dict(mode="sdf", D=3, ctype="grid", mapping="x"),
    "generic_dims_2d": dict(mode="image", D=2, ctype="grid", dims=(3, 5, 2), resos=(7, 9, 4)),
    "generic_dims_vm": dict(mode="sdf", D=3, ctype="vm", dims=(2, 3), resos=(6, 8)),
    # image-set mode (models/FactorFields.py:286-288, :469-470): 2-D basis on (x, y),
    # 3-D coefficient grid [C, n_images, h, w] sampled at (x, y, image)
    "images_grid": dict(mode="images", D=2, ctype="grid", aabb=[[0.0, 0.0, 0.0], [8.0, 6.0, 5.0]]),
    "images_generic_dims": dict(mode="images", D=2, ctype="grid", dims=(6, 3), resos=(7, 9),
                                aabb=[[0.0, 0.0, 0.0], [8.0, 6.0, 4.0]]),
}


def _case(name):
    print(f"[case] {name}", file=sys.stderr, flush=True)
    return small_cfg(**CASES[name])


@pytest.mark.parametrize("name", list(CASES))
def test_get_coding_forward(name):
    m = new_model(_case(name))
    of = oracle_of(m)
    x = rand_points(m, 4099)
    feats, coeff = m.get_coding(x.cuda())
    basis = m.get_basis(x.cuda())
    rf, rc = of.get_coding(x)
    assert_close(feats, rf, FWD_RTOL, "feats")
    assert_close(coeff, rc, FWD_RTOL, "coeff")
    assert_close(basis, of.get_basis(x), FWD_RTOL, "basis")


@pytest.mark.parametrize("name", list(CASES))
def test_get_coding_backward(name):
    m = new_model(_case(name))
    of = oracle_of(m).requires_grad_(True)
    x = rand_points(m, 3001, seed=4)
    g = torch.Generator().manual_seed(9)
    r1 = torch.randn((3001, m._spec.C), generator=g)
    r2 = torch.randn((3001, m._spec.C), generator=g)
    feats, coeff = m.get_coding(x.cuda())
    ((feats * r1.cuda()).sum() + (coeff * r2.cuda()).sum()).backward()
    rf, rc = of.get_coding(x)
    ((rf * r1).sum() + (rc * r2).sum()).backward()
    tabs = list(m.basises) + list(m.coeffs)
    otabs = of.basises + of.coeffs
    for k, (t, o) in enumerate(zip(tabs, otabs)):
        assert_close(t.grad.contiguous(), o.grad, GRAD_RTOL, f"table {k}", atol_frac=1e-5)


@pytest.mark.parametrize("name", list(CASES))
def test_get_coeff_and_get_basis_alone(name):
    """get_coeff / get_basis gather only their own tables (a field spec with
    the other part 'none'); values and gradients equal the oracle's
    (models/FactorFields.py:425-516), and the other part's tables get no
    gradient."""
    m = new_model(_case(name))
    of = oracle_of(m).requires_grad_(True)
    x = rand_points(m, 2999, seed=6)
    g = torch.Generator().manual_seed(10)
    r1 = torch.randn((2999, m._spec.C), generator=g)
    r2 = torch.randn((2999, m._spec.C), generator=g)
    coeff = m.get_coeff(x.cuda())
    basis = m.get_basis(x.cuda())
    assert_close(coeff, of.get_coeff(x), FWD_RTOL, "coeff alone")
    assert_close(basis, of.get_basis(x), FWD_RTOL, "basis alone")
    m.zero_grad(set_to_none=True)
    (coeff * r1.cuda()).sum().backward()
    (of.get_coeff(x) * r1).sum().backward()
    assert all(t.grad is None for t in m.basises)
    for k, (t, o) in enumerate(zip(m.coeffs, of.coeffs)):
        assert_close(t.grad.contiguous(), o.grad, GRAD_RTOL, f"coeff table {k}", atol_frac=1e-5)
    m.zero_grad(set_to_none=True)
    for t in of.basises:
        t.grad = None
    (basis * r2.cuda()).sum().backward()
    (of.get_basis(x) * r2).sum().backward()
    # image-set inputs (x carries the image index) take the full coding kernel: zero coefficient gradients
    assert all(t.grad is None or not t.grad.any() for t in m.coeffs)
    for k, (t, o) in enumerate(zip(m.basises, of.basises)):
        assert_close(t.grad.contiguous(), o.grad, GRAD_RTOL, f"basis table {k}", atol_frac=1e-5)


@pytest.mark.parametrize("layers,out_dim", [(2, 3), (2, 1), (2, 28), (1, 1), (2, 5)])
def test_linear_mat(layers, out_dim):
    m = new_model(small_cfg(mode="sdf", D=3, out_dim=out_dim, layers=layers))
    of = oracle_of(m).requires_grad_(True)
    g = torch.Generator().manual_seed(2)
    f = torch.randn((2050, 18), generator=g)
    fc = f.cuda().requires_grad_(True)
    out = m.linear_mat(fc)
    fo = f.clone().requires_grad_(True)
    ref = of.linear_mat(fo)
    assert_close(out, ref, FWD_RTOL, "out")
    r = torch.randn(out.shape, generator=g)
    (out * r.cuda()).sum().backward()
    (ref * r).sum().backward()
    assert_close(fc.grad, fo.grad, GRAD_RTOL, "d_feats", atol_frac=1e-5)
    for l, (w, b) in enumerate(of.mlp):
        lin = m.linear_mat.backbone[l]
        assert_close(lin.weight.grad, w.grad, GRAD_RTOL, f"W{l}", atol_frac=1e-5)
        if b is not None:
            assert_close(lin.bias.grad, b.grad, GRAD_RTOL, f"b{l}", atol_frac=1e-5)


FUSED = ["2d_grid_bilinear", "3d_grid", "3d_vm", "2d_grid_nearest"]


@pytest.mark.parametrize("name", FUSED)
@pytest.mark.parametrize("out_dim,layers", [(3, 2), (1, 2), (28, 2), (1, 1)])
def test_field_fused(name, out_dim, layers):
    kw = dict(CASES[name], out_dim=out_dim, layers=layers)
    print(f"[case] fused {name} {out_dim} {layers}", file=sys.stderr, flush=True)
    m = new_model(small_cfg(**kw))
    assert m.fused_supported()
    of = oracle_of(m).requires_grad_(True)
    x = rand_points(m, 5000, seed=7)
    out = m.field(x.cuda())
    ref = of.field(x)
    assert_close(out, ref, FWD_RTOL, "out")
    g = torch.Generator().manual_seed(3)
    r = torch.randn(out.shape, generator=g)
    (out * r.cuda()).sum().backward()
    (ref * r).sum().backward()
    for k, (t, o) in enumerate(zip(list(m.basises) + list(m.coeffs), of.basises + of.coeffs)):
        assert_close(t.grad.contiguous(), o.grad, GRAD_RTOL, f"table {k}", atol_frac=1e-5)
    for l, (w, b) in enumerate(of.mlp):
        lin = m.linear_mat.backbone[l]
        assert_close(lin.weight.grad, w.grad, GRAD_RTOL, f"W{l}", atol_frac=1e-5)
        if b is not None:
            assert_close(lin.bias.grad, b.grad, GRAD_RTOL, f"b{l}", atol_frac=1e-5)


@pytest.mark.parametrize("loss", ["mse", "l1"])
@pytest.mark.parametrize("name", ["2d_grid_bilinear", "3d_vm", "3d_vm_nearest", "3d_vm_long_lines"])
def test_regression_step(name, loss):
    """The fused training pass against the oracle.  VM: lines of up to 128
    texels are summed in LDS per block (bilinear and nearest); 3d_vm_long_lines
    (coefficient lines of 140+ texels) takes the global-atomic path."""
    print(f"[case] regression step {name} {loss}", file=sys.stderr, flush=True)
    kw = dict(CASES["3d_vm"], coeff_reso=140) if name == "3d_vm_long_lines" else CASES[name]
    m = new_model(small_cfg(**dict(kw, out_dim=3 if name.startswith("2d") else 1)))
    if name == "3d_vm_long_lines":
        assert max(m.vm_reso) > 128
    of = oracle_of(m).requires_grad_(True)
    n = 6000
    x = rand_points(m, n, seed=8)
    tgt = torch.randn((n, m.cfg.model.out_dim), generator=torch.Generator().manual_seed(5))
    ls = m.regression_step(x.cuda(), tgt.cuda(), loss=loss, loss_scale=0.7)
    ref = of.field(x)
    e = ref - tgt
    lr = (e ** 2).mean() if loss == "mse" else e.abs().mean()
    (lr * 0.7).backward()
    s = (e ** 2).sum() if loss == "mse" else e.abs().sum()
    assert abs(ls.item() - s.item()) <= 1e-4 * abs(s.item())
    for k, (t, o) in enumerate(zip(list(m.basises) + list(m.coeffs), of.basises + of.coeffs)):
        assert_close(t.grad.contiguous(), o.grad, GRAD_RTOL, f"table {k}", atol_frac=1e-5)
    for l, (w, b) in enumerate(of.mlp):
        lin = m.linear_mat.backbone[l]
        assert_close(lin.weight.grad, w.grad, GRAD_RTOL, f"W{l}", atol_frac=1e-5)


def test_regression_step_morton_batch():
    """The SDF workload's model (VM 128, six basis levels) on its batch order:
    a sorted draw from a Morton-ordered pool, with repeats (pool as large as
    the batch), puts long runs of equal texels in every wave, so the scatter
    sums runs of many points before each atomic.  Gradients against the FP64
    oracle, within the FP32 oracle's own distance from it."""
    from ffields import workloads as W
    print("[case] regression step, SDF model, Morton batch", file=sys.stderr, flush=True)
    cfg = W.sdf_cfg()
    m = new_model(cfg, seed=2)
    n = 1 << 15
    pts, sdf = W.SdfScene().training_pool(n)
    perm = W.morton_order(pts, cfg.dataset.aabb)
    pts, sdf = pts[perm], sdf[perm]
    idx = W.sorted_batch(n, n, torch.Generator().manual_seed(4), "cpu").long()
    assert len(torch.unique(idx)) < n             # repeated points: runs across every corner slot
    x, t = pts[idx], sdf[idx]
    ls = m.regression_step(x.cuda(), t.cuda(), loss="l1")
    ofs = []
    for dt in (torch.float32, torch.float64):
        of = oracle_of(m, dtype=dt).requires_grad_(True)
        e = of.field(x.to(dt)) - t.to(dt)
        e.abs().mean().backward()
        ofs.append(of)
    s = (ofs[1].field(x.double()) - t.double()).abs().sum().item()
    assert abs(ls.item() - s) <= 1e-4 * abs(s)
    for k, (tb, o32, o64) in enumerate(zip(list(m.basises) + list(m.coeffs), ofs[0].basises + ofs[0].coeffs,
                                           ofs[1].basises + ofs[1].coeffs)):
        assert_close_fp64(tb.grad.contiguous(), o32.grad, o64.grad, GRAD_RTOL, f"table {k}", atol_frac=1e-5)
    for l, ((w32, _), (w64, _)) in enumerate(zip(ofs[0].mlp, ofs[1].mlp)):
        assert_close_fp64(m.linear_mat.backbone[l].weight.grad, w32.grad, w64.grad, GRAD_RTOL, f"W{l}",
                          atol_frac=1e-5)


BF16_CASES = {
    "3d_grid": dict(mode="sdf", D=3, ctype="grid"),
    # the dictionary image-set layout: 144 channels, image-set coefficient grid
    "images_dict_c144": dict(mode="images", D=2, ctype="grid", dims=(32, 32, 32, 16, 16, 16),
                             resos=(5, 6, 7, 8, 9, 10), coeff_reso=4, aabb=[[0.0, 0.0, 0.0], [8.0, 6.0, 3.0]]),
}


@pytest.mark.parametrize("name", list(BF16_CASES))
def test_bf16_tables_forward(name):
    """BF16 tables: features equal the oracle's on the BF16-rounded tables
    (FP32 arithmetic).  The field output does too, except on the dictionary
    layout, whose fused kernel runs the decoder on BF16 matrix cores: there it
    is checked against the emulation of that decoder's rounding
    (test_gpu_dict.emulate, same bars as there)."""
    print(f"[case] bf16 {name}", file=sys.stderr, flush=True)
    cfg = small_cfg(**BF16_CASES[name])
    m = new_model(cfg, dtype="bf16")
    of = oracle_of(m, bf16_tables=True)
    x = rand_points(m, 4000, seed=12)
    feats, coeff = m.get_coding(x.cuda())
    rf, rc = of.get_coding(x)
    assert_close(feats, rf, FWD_RTOL, "feats")
    if name == "images_dict_c144":
        from test_gpu_dict import emulate
        assert m.fused_supported()
        ref, *_ = emulate(m, rf)
        assert_close(m.field(x.cuda()), ref, 2e-3, "field", atol_frac=2e-3)
    else:
        assert_close(m.field(x.cuda()), of.field(x), FWD_RTOL, "field")


def test_empty_batch():
    m = new_model(small_cfg(mode="sdf", D=3, ctype="vm"))
    x = torch.zeros((0, 3), device="cuda")
    feats, coeff = m.get_coding(x)
    assert feats.shape == (0, 18)
    assert m.field(x).shape == (0, 3)


def test_adam_matches_torch():
    from ffields import FusedAdam
    torch.manual_seed(0)
    p1 = torch.nn.Parameter(torch.randn(10007, device="cuda"))
    p2 = torch.nn.Parameter(torch.randn((1, 4, 9, 9), device="cuda").contiguous(memory_format=torch.channels_last))
    q1 = torch.nn.Parameter(p1.detach().clone())
    q2 = torch.nn.Parameter(p2.detach().clone())
    opt = FusedAdam([{"params": [p1], "lr": 0.02}, {"params": [p2], "lr": 0.001}], betas=(0.9, 0.99))
    ref = torch.optim.Adam([{"params": [q1], "lr": 0.02}, {"params": [q2], "lr": 0.001}],
                           betas=(0.9, 0.99), foreach=False)
    for it in range(5):
        g1 = torch.randn(p1.shape, device="cuda")
        g2 = torch.randn(p2.shape, device="cuda").contiguous(memory_format=torch.channels_last)
        p1.grad, p2.grad = g1.clone(), g2.clone()
        q1.grad, q2.grad = g1.clone(), g2.clone()
        opt.step()
        ref.step()
    assert_close(p1, q1, 1e-6, "p1", atol_frac=1e-7)
    assert_close(p2, q2, 1e-6, "p2", atol_frac=1e-7)


def test_grad_replicas_summed_by_adam():
    """Table gradients split over per-XCD copies: FusedAdam summing them in its
    pass equals reducing first, and both equal the single-copy result."""
    from ffields import FusedAdam
    cfg = small_cfg(**dict(CASES["3d_vm"], out_dim=1))
    x = rand_points(new_model(cfg), 5000, seed=21).cuda()
    tgt = torch.randn((5000, 1), generator=torch.Generator().manual_seed(22)).cuda()
    params = []
    for reps, reduce in ((1, True), (8, True), (8, False)):
        m = new_model(cfg, seed=3)
        m.grad_replicas = reps
        opt = FusedAdam(m.get_optparam_groups(0.01, 0.02))
        for _ in range(3):
            m.regression_step(x, tgt, loss="l1", reduce=reduce)
            opt.step(zero_grad=True)
        params.append([p.detach().clone() for p in m.parameters()])
    for a, b, c in zip(*params):
        assert_close(b, a, 1e-5, "reduce-first vs single copy", atol_frac=1e-6)
        assert_close(c, a, 1e-5, "adam-summed vs single copy", atol_frac=1e-6)
