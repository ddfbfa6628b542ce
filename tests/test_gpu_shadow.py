"""BF16 table copies follow their FP32 masters when something other than
FusedAdam writes the masters (torch.optim.Adam in train_per_scene.py:132/218,
load_state_dict, in-place copies).  The kernels must then read the new
tables, never the copies cast before the write."""
import pytest
import torch

from helpers import assert_close, oracle_of, rand_points, randomize, small_cfg
from ffields import FactorFields

pytestmark = pytest.mark.gpu


def test_bf16_forward_follows_torch_adam():
    m = FactorFields(small_cfg(mode="sdf", D=3, ctype="grid", out_dim=2), "cuda", table_dtype="bf16")
    randomize(m, 5, scale=0.5)
    x = rand_points(m, 4096, seed=6).cuda()
    opt = torch.optim.Adam(m.parameters(), lr=0.05, betas=(0.9, 0.99))
    out0 = m.field(x)
    out0.square().mean().backward()
    opt.step()                                  # writes the FP32 masters through torch
    with torch.no_grad():
        out1 = m.field(x)
    assert not torch.equal(out0.detach(), out1)
    ref = oracle_of(m, bf16_tables=True).field(x.cpu())
    assert_close(out1, ref, 1e-4, "bf16 forward after torch.optim.Adam", atol_frac=1e-5)


def test_bf16_forward_follows_load_state_dict():
    m = FactorFields(small_cfg(mode="sdf", D=3, ctype="grid", out_dim=2), "cuda", table_dtype="bf16")
    randomize(m, 7, scale=0.5)
    src = FactorFields(small_cfg(mode="sdf", D=3, ctype="grid", out_dim=2), "cuda")
    randomize(src, 8, scale=0.5)
    m.load_state_dict(src.state_dict())
    x = rand_points(m, 2048, seed=9).cuda()
    with torch.no_grad():
        out = m.field(x)
    ref = oracle_of(src, bf16_tables=True).field(x.cpu())
    assert_close(out, ref, 1e-4, "bf16 forward after load_state_dict", atol_frac=1e-5)


def test_fused_adam_leaves_copies_of_replaced_tables_alone():
    """A table replaced by a new tensor (as shrink and upsample_volume_grid
    do) has no BF16 copy until the next read casts one: FusedAdam updates
    its master only, never the copy cast from the old tensor, and the next
    forward reads a fresh copy of the new master."""
    from ffields import FusedAdam
    m = FactorFields(small_cfg(mode="sdf", D=3, ctype="grid", out_dim=2), "cuda", table_dtype="bf16")
    randomize(m, 11, scale=0.5)
    x = rand_points(m, 2048, seed=12).cuda()
    with torch.no_grad():
        m.field(x)                                  # casts the copies
    old = m.coeffs[0]
    old_copy = m.shadow_of(old)
    assert old_copy is not None
    kept = old_copy.clone()
    new = torch.nn.Parameter(old.detach().clone() * 1.5)
    m.coeffs[0] = new
    assert m.shadow_of(new) is None
    new.grad = torch.full_like(new, 0.25)
    opt = FusedAdam([new], lr=0.05, betas=(0.9, 0.99), shadow_of=m.shadow_lookup)
    opt.step(zero_grad=True)
    assert torch.equal(old_copy, kept)              # the old tensor's copy was not written
    with torch.no_grad():
        out = m.field(x)
    ref = oracle_of(m, bf16_tables=True).field(x.cpu())
    assert_close(out, ref, 1e-4, "bf16 forward after a table replacement", atol_frac=1e-5)
