"""Synthetic workload generators: the batched image-set generator reproduces
the single-image generator exactly (same draws, same accumulation order)."""
import numpy as np
import torch

from ffields import workloads as W


def test_image_set_matches_single_images():
    a = W.synth_image_set(3, 48, "cpu", chunk=2)
    for i in range(3):
        b = W.synth_image(48, seed=W.BASE_SEED + 4000 + i).reshape(-1, 3)
        assert torch.equal(a[i], b), i


def test_dict_cfg_parameter_count():
    """1,024 images: 1024 x 32 x 32 x 144 coefficients + the 144-channel basis."""
    from ffields import FactorFields
    m = FactorFields(W.dict_cfg(), "cpu")
    basis = sum(d * r * r for d, r in zip([32, 32, 32, 16, 16, 16], W.RESOS))
    lin = 144 * 64 + 64 + 64 * 3
    assert m.n_parameters() == 1024 * 32 * 32 * 144 + basis + lin


def test_cell_order_groups_coefficient_cells():
    """cell_order is a permutation of the pixels in which every coefficient
    cell (the pixels interpolating the same four texels: (p + 4) // 8 on a
    256-pixel image with 32 texels) is one contiguous block."""
    o = W.cell_order(256, 32)
    assert torch.equal(torch.sort(o).values, torch.arange(256 * 256))
    y, x = o // 256, o % 256
    cell = ((y + 4) // 8) * 64 + (x + 4) // 8
    assert bool((cell[1:] >= cell[:-1]).all())
    # texel corners of align_corners=False (pixel centre p + 0.5 -> p / 8 - 7/16)
    x0 = torch.clamp((x.double() + 0.5) / 8 - 0.5, 0, 31).floor()
    assert bool(((x + 4) // 8 - x0).abs().le(1).all())


def test_ordered_batch_distribution():
    """The ordered batch is sorted by (image, cell) and has the distribution
    of independent uniform draws: chi-square of the image and column
    histograms within 5 standard deviations of their degrees of freedom, and
    the mean key within 5 standard errors of (K - 1) / 2."""
    o = W.cell_order(64, 8)
    n, N, P = 1 << 18, 32, 64 * 64
    g = torch.Generator().manual_seed(3)
    img, pix = W.image_set_batch(n, N, 64, g, "cpu", order=o)
    rank = torch.argsort(o)
    key = img * P + rank[pix]
    assert bool((key[1:] >= key[:-1]).all())
    assert int(img.min()) >= 0 and int(img.max()) < N and int(pix.min()) >= 0 and int(pix.max()) < P
    for h, k in [(torch.bincount(img, minlength=N), N), (torch.bincount(pix % 64, minlength=64), 64)]:
        e = n / k
        chi2 = float(((h.double() - e) ** 2 / e).sum())
        assert abs(chi2 - (k - 1)) < 5 * (2 * (k - 1)) ** 0.5, chi2
    K = N * P
    flat = img.double() * P + pix.double()
    se = ((K * K - 1) / 12 / n) ** 0.5
    assert abs(float(flat.mean()) - (K - 1) / 2) < 5 * se
    # unordered: the reference's independent draws
    g2 = torch.Generator().manual_seed(3)
    i2, p2 = W.image_set_batch(16, N, 64, g2, "cpu")
    g3 = torch.Generator().manual_seed(3)
    assert torch.equal(i2, torch.randint(0, N, (16,), generator=g3))
    assert torch.equal(p2, torch.randint(0, P, (16,), generator=g3))


def _morton_key(q, bits):
    """Bit-by-bit Z-order key of integer cell coordinates (x, y, z)."""
    k = 0
    for b in range(bits):
        for a in range(3):
            k |= ((int(q[a]) >> b) & 1) << (3 * b + a)
    return k


def test_morton_order_sorts_by_z_order_key():
    g = torch.Generator().manual_seed(3)
    p = torch.rand((500, 3), generator=g) * 4 - 2          # box [-2, 2]^3, some points on its faces
    p[:5] = torch.tensor([[-2.0, -2, -2], [2, 2, 2], [2, -2, 0], [0, 0, 0], [-2, 2, 2]])
    bits = 4
    perm = W.morton_order(p, [[-2, -2, -2], [2, 2, 2]], bits=bits)
    assert sorted(perm.tolist()) == list(range(500))
    q = ((p + 2) / 4 * (1 << bits)).floor().clamp(0, (1 << bits) - 1).long()
    keys = [_morton_key(q[i], bits) for i in perm.tolist()]
    assert keys == sorted(keys)
    # ties keep the pool order (stable)
    for a, b in zip(perm.tolist(), perm.tolist()[1:]):
        if _morton_key(q[a], bits) == _morton_key(q[b], bits):
            assert a < b


def test_sorted_batch_is_a_sorted_uniform_draw():
    """Ascending, in range, and each order statistic distributed as that of
    sorted torch.randint draws (two-sample KS, fixed seeds)."""
    from scipy import stats
    g = torch.Generator().manual_seed(0)
    idx = W.sorted_batch(4000, 100000, g, "cpu")
    assert idx.dtype == torch.int32 and bool((idx[1:] >= idx[:-1]).all())
    assert int(idx.min()) >= 0 and int(idx.max()) < 4000
    assert stats.chisquare(np.bincount(idx.numpy(), minlength=4000)).pvalue > 1e-3
    g = torch.Generator().manual_seed(1)
    a = np.stack([W.sorted_batch(100, 5, g, "cpu").numpy() for _ in range(4000)])
    gr = torch.Generator().manual_seed(2)
    b = np.sort(torch.randint(0, 100, (4000, 5), generator=gr).numpy(), 1)
    for k in range(5):
        assert stats.ks_2samp(a[:, k], b[:, k]).pvalue > 1e-3, k


def test_sorted_batch_large_pool_uses_int64():
    g = torch.Generator().manual_seed(0)
    idx = W.sorted_batch(3 * 2 ** 31, 1000, g, "cpu")
    assert idx.dtype == torch.int64 and int(idx.max()) < 3 * 2 ** 31 and bool((idx[1:] >= idx[:-1]).all())
    assert int(idx.max()) > 2 ** 31                   # the upper part of the pool is reached
