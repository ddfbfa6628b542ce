"""Shared test helpers: build product models and matching oracle copies."""
import copy

import torch

from ffields import FactorFields
from ffields.config import AttrDict
from oracle import OracleField


def small_cfg(mode="sdf", D=3, ctype="grid", cmode="bilinear", bmode="bilinear", mapping="sawtooth",
              dims=(4, 4, 4, 2, 2, 2), resos=(9, 11, 13, 7, 6, 5), coeff_reso=6, out_dim=3, layers=2,
              hidden=64, aabb=None):
    if aabb is None:
        aabb = [[0.0, 0.0], [8.0, 6.0]] if D == 2 else [[-1.0, -0.5, -1.25], [1.0, 1.5, 0.75]]
    freq = [1.0, 1.7, 2.9, 3.3, 4.1, 5.0][:len(dims)]
    return AttrDict.wrap(dict(
        defaults=dict(mode=mode),
        model=dict(basis_dims=list(dims), basis_resos=list(resos), freq_bands=freq, coeff_reso=coeff_reso,
                   coeff_type=ctype, basis_type="grid", coef_mode=cmode, basis_mode=bmode,
                   basis_mapping=mapping, param_derivation="literal", out_dim=out_dim,
                   hidden_dim=hidden, num_layers=layers, coef_init=1.0, total_params=0, T_basis=0,
                   T_coeff=0, with_dropout=False),
        dataset=dict(aabb=aabb, scene_reso=1024, is_unbound=False),
        renderer=dict(fea2denseAct="softplus", density_shift=-10, distance_scale=25.0),
        training=dict()))


def randomize(model, seed=0, scale=1.0):
    g = torch.Generator().manual_seed(seed)
    with torch.no_grad():
        for p in model.parameters():
            p.copy_((torch.randn(p.shape, generator=g) * scale).to(p.device))
    if model.table_dtype == "bf16":
        model.refresh_shadows()


def oracle_of(model, dtype=torch.float32, bf16_tables=False):
    cfg = copy.deepcopy(model.cfg)
    of = OracleField(cfg, dtype=dtype)

    def cv(t):
        t = t.detach().float().cpu().contiguous()
        if bf16_tables:
            t = t.to(torch.bfloat16).float()
        return t.to(dtype)

    if model.coeff_type != "none":
        of.coeffs = [cv(t) for t in model.coeffs]
    if model.basis_type != "none":
        of.basises = [cv(t) for t in model.basises]
    lin = model.linear_mat.backbone
    of.mlp = [(lin[i].weight.detach().cpu().to(dtype),
               None if lin[i].bias is None else lin[i].bias.detach().cpu().to(dtype))
              for i in range(len(lin))]
    return of


def rand_points(model, n, seed=1, margin=0.1):
    g = torch.Generator().manual_seed(seed)
    a = model.aabb.cpu()
    if model.cfg.defaults.mode == "images":
        lo, hi = a[0], a[1]
    else:
        lo, hi = a[0], a[1]
    ext = hi - lo
    return lo - margin * ext + (1 + 2 * margin) * ext * torch.rand((n, len(lo)), generator=g)


def assert_close(a, b, rtol, what="", atol_frac=1e-6):
    a = a.detach().double().cpu()
    b = b.detach().double().cpu()
    assert a.shape == b.shape, (what, a.shape, b.shape)
    scale = b.abs().max().item() or 1.0
    err = (a - b).abs()
    tol = rtol * b.abs() + atol_frac * scale
    bad = err > tol
    assert not bad.any(), (f"{what}: {int(bad.sum())}/{bad.numel()} outside tol; max abs err "
                           f"{err.max().item():.3e}, max rel {(err / (b.abs() + 1e-30)).max().item():.3e}")


def new_model(cfg, dtype="fp32", seed=0):
    m = FactorFields(cfg, "cuda", table_dtype=dtype)
    randomize(m, seed)
    return m


def assert_close_fp64(a, b32, b64, rtol, what="", atol_frac=1e-6, slack=2.0):
    """Element-wise: the kernel result a is within the usual bar of the FP64
    result b64 (rtol * |b64| + atol_frac * max|b64|), or no further from it
    than `slack` times the FP32 oracle b32 is.  For sums whose terms cancel,
    where two FP32 summation orders may differ by more than the bar while
    both are as close to the exact value as FP32 allows."""
    a = a.detach().double().cpu()
    b32 = b32.detach().double().cpu()
    b64 = b64.detach().double().cpu()
    assert a.shape == b64.shape == b32.shape, (what, a.shape, b32.shape, b64.shape)
    scale = b64.abs().max().item() or 1.0
    err = (a - b64).abs()
    tol = torch.maximum(rtol * b64.abs() + atol_frac * scale, slack * (b32 - b64).abs())
    bad = err > tol
    assert not bad.any(), (f"{what}: {int(bad.sum())}/{bad.numel()} outside tol vs FP64; max abs err "
                           f"{err.max().item():.3e}, FP32 oracle max abs err {(b32 - b64).abs().max().item():.3e}")
