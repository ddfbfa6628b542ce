"""Pins the oracle's render functions (oracle/render.py) against closed forms,
identities and brute-force loops that share no code with it."""
import math

import numpy as np
import torch

import npref
from oracle import alpha_grid_sample, composite, eval_sh, raw2alpha, sample_pdf, sample_point


def test_sample_point_closed_form():
    g = torch.Generator().manual_seed(0)
    B, Ns = 50, 40
    o = torch.randn((B, 3), generator=g) * 3
    d = torch.randn((B, 3), generator=g)
    d = d / d.norm(dim=-1, keepdim=True)
    aabb = torch.tensor([[-1.0, -1.2, -0.8], [1.1, 0.9, 1.3]])
    step = torch.tensor(0.05)
    u = torch.rand(B, generator=g)
    pts, t, inner = sample_point(o, d, aabb, step, Ns, u)
    on, dn, an = o.double().numpy(), d.double().numpy(), aabb.double().numpy()
    ta = (an[1] - on) / dn
    tb = (an[0] - on) / dn
    tmin = np.clip(np.minimum(ta, tb).max(-1), 0.05, 1e3)
    t_ref = tmin[:, None] + 0.05 * (np.arange(Ns)[None] + u.double().numpy()[:, None])
    assert np.allclose(t.double().numpy(), t_ref, atol=1e-5)
    p_ref = on[:, None] + dn[:, None] * t_ref[..., None]
    assert np.allclose(pts.double().numpy(), p_ref, atol=1e-4)
    inside = ((p_ref >= an[0]) & (p_ref <= an[1])).all(-1)
    margin = np.minimum(np.abs(p_ref - an[0]), np.abs(p_ref - an[1])).min(-1) > 1e-4
    assert np.array_equal(inner.numpy()[margin], inside[margin])
    # eval mode: no jitter, integer multiples of the step
    _, t0, _ = sample_point(o, d, aabb, step, Ns, None)
    assert np.allclose(t0.double().numpy(), tmin[:, None] + 0.05 * np.arange(Ns)[None], atol=1e-5)


def test_alpha_grid_sample_vs_bruteforce():
    g = torch.Generator().manual_seed(1)
    vol = (torch.rand((5, 6, 7), generator=g) < 0.5).float()
    aabb = torch.tensor([[-1.0, -0.5, -1.5], [1.0, 1.5, 0.5]])
    x = aabb[0] + (aabb[1] - aabb[0]) * (torch.rand((300, 3), generator=g) * 1.2 - 0.1)
    got = alpha_grid_sample(vol, aabb, x.double().float())
    c = ((x.double() - aabb[0].double()) / (aabb[1] - aabb[0]).double() * 2 - 1).numpy()
    ref = npref.grid_sample(vol[None].numpy(), c, align_corners=True, padding="zeros")[:, 0]
    assert np.allclose(got.numpy(), ref, atol=1e-5)


def test_raw2alpha_telescoping_and_special_cases():
    g = torch.Generator().manual_seed(2)
    sigma = torch.rand((20, 33), generator=g, dtype=torch.float64) * 5
    sigma[:, ::4] = 0
    dist = torch.rand((20, 33), generator=g, dtype=torch.float64) * 0.3
    alpha, w, T_last = raw2alpha(sigma, dist)
    f = 1 - alpha + 1e-10
    T = torch.cumprod(torch.cat([torch.ones_like(f[:, :1]), f], -1), -1)
    # sum(w) + T_last = 1 + 1e-10 * sum_i T_i
    assert torch.allclose(w.sum(-1) + T_last[:, 0], 1 + 1e-10 * T[:, :-1].sum(-1), atol=1e-12)
    # w_i = alpha_i * prod_{j<i} (1 - alpha_j + 1e-10), by an explicit loop
    for r in range(3):
        run = 1.0
        for i in range(33):
            a = 1 - math.exp(-float(sigma[r, i]) * float(dist[r, i]))
            assert abs(float(w[r, i]) - a * run) < 1e-12
            run *= 1 - a + 1e-10
    a1, w1, T1 = raw2alpha(torch.tensor([[2.0]], dtype=torch.float64), torch.tensor([[0.5]], dtype=torch.float64))
    assert torch.allclose(w1, a1) and torch.allclose(T1, 1 - a1 + 1e-10)
    a0, w0, T0 = raw2alpha(torch.zeros((1, 5), dtype=torch.float64), torch.ones((1, 5), dtype=torch.float64))
    assert torch.all(w0 == 0) and torch.all(a0 == 0)


def test_composite_white_background():
    sigma = torch.zeros((4, 10), dtype=torch.float64)
    rgb = torch.rand((4, 10, 3), dtype=torch.float64)
    z = torch.linspace(0, 1, 10, dtype=torch.float64).expand(4, 10)
    dists = torch.full((4, 10), 0.1, dtype=torch.float64)
    rgb_map, depth, w, acc = composite(sigma, rgb, z, dists, 25.0, True)
    assert torch.allclose(rgb_map, torch.ones_like(rgb_map)) and torch.all(depth == 0)
    rgb_map, _, _, _ = composite(sigma, rgb, z, dists, 25.0, False)
    assert torch.all(rgb_map == 0)
    # one opaque sample returns its colour
    sigma[:, 3] = 1e6
    rgb_map, depth, _, acc = composite(sigma, rgb, z, dists, 25.0, True)
    assert torch.allclose(rgb_map, rgb[:, 3].clamp(0, 1), atol=1e-8)
    assert torch.allclose(depth, z[:, 3], atol=1e-8)


def test_composite_gradient_fp64_central_differences():
    """d rgb_map / d sigma and d rgb_map / d rgb of the composite (the
    quantities the hand-written composite backward produces) against FP64
    central differences of an explicit-loop composite."""
    g = torch.Generator().manual_seed(9)
    sigma = torch.rand((3, 12), generator=g, dtype=torch.float64) * 3
    rgb = torch.rand((3, 12, 3), generator=g, dtype=torch.float64) * 0.8 + 0.1
    z = torch.cumsum(torch.rand((3, 12), generator=g, dtype=torch.float64) * 0.05 + 0.01, -1)
    dists = torch.cat([z[:, 1:] - z[:, :-1], torch.zeros_like(z[:, :1])], -1)
    wt = torch.rand((3, 3), generator=g, dtype=torch.float64)

    def loop_loss(sg, cl):
        tot = 0.0
        for r in range(3):
            run, acc, out = 1.0, 0.0, [0.0, 0.0, 0.0]
            for i in range(12):
                a = 1 - math.exp(-float(sg[r, i]) * float(dists[r, i]) * 25.0)
                w = a * run
                acc += w
                for c in range(3):
                    out[c] += w * float(cl[r, i, c])
                run *= 1 - a + 1e-10
            for c in range(3):
                tot += float(wt[r, c]) * min(max(out[c] + (1 - acc), 0.0), 1.0)
        return tot

    s_ = sigma.clone().requires_grad_(True)
    c_ = rgb.clone().requires_grad_(True)
    rgb_map, _, _, _ = composite(s_, c_, z, dists, 25.0, True)
    gs, gc = torch.autograd.grad((rgb_map * wt).sum(), [s_, c_])
    eps = 1e-6
    for (r, i) in [(0, 0), (1, 5), (2, 11), (0, 7)]:
        sp, sm = sigma.clone(), sigma.clone()
        sp[r, i] += eps
        sm[r, i] -= eps
        fd = (loop_loss(sp, rgb) - loop_loss(sm, rgb)) / (2 * eps)
        assert abs(fd - float(gs[r, i])) <= 1e-6 * max(1.0, abs(fd)), (r, i, fd, float(gs[r, i]))
        cp, cm = rgb.clone(), rgb.clone()
        cp[r, i, 1] += eps
        cm[r, i, 1] -= eps
        fd = (loop_loss(sigma, cp) - loop_loss(sigma, cm)) / (2 * eps)
        assert abs(fd - float(gc[r, i, 1])) <= 1e-6 * max(1.0, abs(fd)), (r, i)


def test_sh_degree2_orthonormal():
    # Gauss-Legendre in cos(theta) x uniform phi quadrature
    nt, nphi = 24, 48
    xs, ws = np.polynomial.legendre.leggauss(nt)
    phi = (np.arange(nphi) + 0.5) * 2 * np.pi / nphi
    ct, ph = np.meshgrid(xs, phi, indexing="ij")
    st = np.sqrt(1 - ct ** 2)
    dirs = torch.tensor(np.stack([st * np.cos(ph), st * np.sin(ph), ct], -1).reshape(-1, 3))
    wq = torch.tensor((ws[:, None] * np.full(nphi, 2 * np.pi / nphi)[None]).reshape(-1))
    Y = torch.stack([eval_sh(2, torch.eye(9, dtype=torch.float64)[k].expand(len(dirs), 1, 9), dirs)[:, 0]
                     for k in range(9)], -1)
    G = (Y * wq[:, None]).T @ Y
    assert torch.allclose(G, torch.eye(9, dtype=torch.float64), atol=1e-6)


def test_sample_pdf_uniform_and_bruteforce():
    bins = torch.linspace(2.0, 6.0, 9, dtype=torch.float64).expand(3, 9).contiguous()
    w = torch.ones((3, 8), dtype=torch.float64)
    s = sample_pdf(bins, w, 17, det=True)
    assert torch.allclose(s, torch.linspace(2.0, 6.0, 17, dtype=torch.float64).expand(3, 17), atol=1e-9)
    g = torch.Generator().manual_seed(4)
    bins = torch.sort(torch.rand((5, 11), generator=g, dtype=torch.float64) * 4, -1)[0]
    w = torch.rand((5, 10), generator=g, dtype=torch.float64)
    w[:, 3] = 0
    u = torch.sort(torch.rand((5, 13), generator=g, dtype=torch.float64), -1)[0]
    got = sample_pdf(bins, w, 13, u=u)
    for r in range(5):
        ww = [float(v) + 1e-5 for v in w[r]]
        tot = sum(ww)
        cdf = [0.0]
        for v in ww:
            cdf.append(cdf[-1] + v / tot)
        for k in range(13):
            uu = float(u[r, k])
            above = 0
            while above < len(cdf) and cdf[above] <= uu:
                above += 1
            below = max(above - 1, 0)
            above = min(above, len(cdf) - 1)
            den = cdf[above] - cdf[below]
            den = 1.0 if den < 1e-5 else den
            t = (uu - cdf[below]) / den
            ref = float(bins[r, below]) + t * (float(bins[r, above]) - float(bins[r, below]))
            assert abs(float(got[r, k]) - ref) < 1e-9


# ---------------------------------------------------------------- MLPRender_Fea
def test_positional_encoding_layout_and_values():
    """positional_encoding (models/FactorFields.py:74-79): for D channels and F
    frequencies the output is [sin block | cos block], entry c*F + f of each
    block at channel c, frequency 2^f.  A single non-zero channel isolates its
    entries, so a transposed (f*D + c) layout fails."""
    import math
    from oracle import positional_encoding
    D, F = 5, 3
    for k in range(D):
        a = 0.37 + 0.1 * k
        x = torch.zeros((1, D), dtype=torch.float64)
        x[0, k] = a
        pe = positional_encoding(x, F)[0]
        assert pe.shape == (2 * D * F,)
        for c in range(D):
            for f in range(F):
                v = a * 2 ** f if c == k else 0.0
                assert abs(pe[c * F + f].item() - math.sin(v)) < 1e-15
                assert abs(pe[D * F + c * F + f].item() - math.cos(v)) < 1e-15


def _np_render_fea(viewdirs, feat, layers, viewpe, feape):
    """MLPRender_Fea.forward (:188-203) in numpy FP64 with the input assembled
    entry by entry: features, viewdir, sin/cos of features * 2^f, sin/cos of
    viewdir * 2^f."""
    import numpy as np
    def pe(x, F):
        s = [np.sin(x[:, c] * 2.0 ** f) for c in range(x.shape[1]) for f in range(F)]
        co = [np.cos(x[:, c] * 2.0 ** f) for c in range(x.shape[1]) for f in range(F)]
        return np.stack(s + co, 1)
    h = np.concatenate([feat, viewdirs, pe(feat, feape), pe(viewdirs, viewpe)], 1)
    for l, (w, b) in enumerate(layers):
        h = h @ w.T + (0 if b is None else b)
        if l != len(layers) - 1:
            h = np.maximum(h, 0)
    return 1 / (1 + np.exp(-h))


def test_mlp_render_fea_against_numpy_and_shapes():
    import numpy as np
    from oracle import mlp_render_fea
    g = torch.Generator().manual_seed(1)
    N, C, H = 64, 31, 128
    for viewpe in (6, 2):
        in_mlpC = 3 + C + 2 * viewpe * 3 + 2 * 2 * C          # models/FactorFields.py:166
        assert in_mlpC == (194 if viewpe == 6 else 170)
        layers = [(torch.randn((H, in_mlpC), generator=g, dtype=torch.float64) * 0.1,
                   torch.randn(H, generator=g, dtype=torch.float64) * 0.1),
                  (torch.randn((H, H), generator=g, dtype=torch.float64) * 0.1,
                   torch.randn(H, generator=g, dtype=torch.float64) * 0.1),
                  (torch.randn((3, H), generator=g, dtype=torch.float64) * 0.1, None)]
        feat = torch.randn((N, C), generator=g, dtype=torch.float64)
        d = torch.randn((N, 3), generator=g, dtype=torch.float64)
        d = d / d.norm(dim=-1, keepdim=True)
        got = mlp_render_fea(d, feat, layers, viewpe=viewpe, feape=2).numpy()
        ref = _np_render_fea(d.numpy(), feat.numpy(), [(w.numpy(), None if b is None else b.numpy())
                                                       for w, b in layers], viewpe, 2)
        assert np.abs(got - ref).max() < 1e-13
        assert (got > 0).all() and (got < 1).all()


def test_mlp_render_fea_feature_gradient_fp64_central_differences():
    """The feature gradient through the positional encoding (the chain the
    colour kernels' backward implements), against FP64 central differences."""
    from oracle import mlp_render_fea
    g = torch.Generator().manual_seed(2)
    N, C, H = 3, 31, 128
    layers = [(torch.randn((H, 194), generator=g, dtype=torch.float64) * 0.1,
               torch.randn(H, generator=g, dtype=torch.float64) * 0.1),
              (torch.randn((H, H), generator=g, dtype=torch.float64) * 0.1,
               torch.randn(H, generator=g, dtype=torch.float64) * 0.1),
              (torch.randn((3, H), generator=g, dtype=torch.float64) * 0.1, None)]
    feat = torch.randn((N, C), generator=g, dtype=torch.float64)
    d = torch.nn.functional.normalize(torch.randn((N, 3), generator=g, dtype=torch.float64), dim=-1)
    R = torch.randn((N, 3), generator=g, dtype=torch.float64)
    f = feat.clone().requires_grad_(True)
    (mlp_render_fea(d, f, layers) * R).sum().backward()
    eps = 1e-6
    for i in range(N):
        for c in range(0, C, 5):
            fp, fm = feat.clone(), feat.clone()
            fp[i, c] += eps
            fm[i, c] -= eps
            num = ((mlp_render_fea(d, fp, layers) * R).sum() - (mlp_render_fea(d, fm, layers) * R).sum()) / (2 * eps)
            assert abs(num.item() - f.grad[i, c].item()) < 1e-7 * max(1.0, abs(num.item()))


def test_linear_mat_dropout_with_injected_mask():
    """OracleField.linear_mat(h, keep): F.dropout(h, p=0.1) with the mask keep
    is h * keep / 0.9 entering the MLP; numpy FP64 of that closed form."""
    import numpy as np
    from oracle import OracleField, make_cfg
    cfg = make_cfg("sdf", dict(out_dim=3, num_layers=2, hidden_dim=64, coeff_reso=4,
                               basis_resos=[5, 6, 7, 8, 9, 10], total_params=0),
                   dict(aabb=[[0.0, 0.0, 0.0], [4.0, 4.0, 4.0]]))
    of = OracleField(cfg, dtype=torch.float64)
    g = torch.Generator().manual_seed(3)
    of.mlp = [(torch.randn((64, 18), generator=g, dtype=torch.float64), torch.randn(64, generator=g,
                                                                                   dtype=torch.float64)),
              (torch.randn((3, 64), generator=g, dtype=torch.float64), None)]
    h = torch.randn((50, 18), generator=g, dtype=torch.float64)
    keep = torch.rand((50, 18), generator=g) < 0.9
    got = of.linear_mat(h, keep=keep).numpy()
    (w0, b0), (w1, _) = [(w.numpy(), None if b is None else b.numpy()) for w, b in of.mlp]
    x = h.numpy() * keep.numpy() / 0.9
    ref = np.maximum(x @ w0.T + b0, 0) @ w1.T
    assert np.abs(got - ref).max() < 1e-12
    # no mask: the plain decoder
    assert torch.equal(of.linear_mat(h), of.linear_mat(h, keep=None))
