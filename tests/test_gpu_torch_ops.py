"""The ffields custom ops on the GPU: torch.library.opcheck (schema, fake
implementation, autograd registration), and a torch.compile'd training step
(aot_eager: traced through dynamo and AOT autograd, no code generation)
whose loss and gradients equal eager mode's."""
import sys

import pytest
import torch

import ffields.ops as O
from helpers import new_model, rand_points, small_cfg

pytestmark = pytest.mark.gpu


def _model(**kw):
    m = new_model(small_cfg(**kw))
    with torch.no_grad():
        for p in m.parameters():
            p.mul_(0.5)
    return m


@pytest.mark.parametrize("ctype", ["grid", "vm"])
def test_opcheck(ctype):
    print(f"[case] opcheck {ctype}", file=sys.stderr, flush=True)
    m = _model(mode="sdf", D=3, ctype=ctype, out_dim=1)
    x = rand_points(m, 256, seed=3).cuda()
    b, c = m._tables()
    tabs = list(b + c)
    params = list(m.linear_mat.params_list())
    si, sf = O.spec_args(m._spec)
    # the atomics' summation order differs between runs, so compare with a tolerance, not bitwise
    utils = ("test_schema", "test_autograd_registration", "test_faketensor")
    torch.library.opcheck(O.get_coding_op, (x, tabs, tabs, si, sf), test_utils=utils)
    torch.library.opcheck(O.field_op, (x, params, tabs, tabs, si, sf), test_utils=utils)
    feats = O.get_coding(m._spec, tabs, x, tabs)[0].detach()
    torch.library.opcheck(O.linear_mat_op, (feats, params), test_utils=utils)
    torch.library.opcheck(O.dropout_op, (feats.requires_grad_(True), 0.1, 11, None), test_utils=utils)


def test_compiled_training_step_matches_eager():
    print("[case] torch.compile aot_eager", file=sys.stderr, flush=True)
    m = _model(mode="sdf", D=3, ctype="vm", out_dim=1)
    x = rand_points(m, 4096, seed=4).cuda()
    t = torch.randn((4096, 1), generator=torch.Generator().manual_seed(5)).cuda()

    def step(x, t):
        out = m.field(x)
        feats, coeff = m.get_coding(x)
        return (out - t).abs().mean() + 1e-3 * m.linear_mat(feats).square().mean() + 1e-3 * coeff.square().mean()

    def grads(fn):
        m.zero_grad(set_to_none=True)
        loss = fn(x, t)
        loss.backward()
        return loss.detach(), [p.grad.detach().clone() for p in m.parameters()]

    l0, g0 = grads(step)
    compiled = torch.compile(step, backend="aot_eager", fullgraph=True)
    l1, g1 = grads(compiled)
    assert torch.allclose(l0, l1, rtol=1e-6, atol=0)
    for a, b in zip(g0, g1):
        assert torch.allclose(a, b, rtol=1e-4, atol=1e-6 * float(a.abs().max()))


def test_stream_of_is_the_current_stream():
    """The entry points run on torch's current stream of the tensor's device,
    also inside a stream context."""
    import ffields._lib as L
    t = torch.zeros(4, device="cuda")
    assert (L.stream_of(t).value or 0) == torch.cuda.current_stream().cuda_stream
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        assert (L.stream_of(t).value or 0) == s.cuda_stream
        assert (L.stream_of().value or 0) == s.cuda_stream
    assert (L.stream_of().value or 0) == torch.cuda.current_stream().cuda_stream
