"""GPU tests of occupancy-grid maintenance (models/FactorFields.py:710-793):
pooling and threshold, the dense alpha volume, and the small-object filter
(ff_remove_small_components) against the oracle's (scipy labelling):
the same kept voxels, the same component of every voxel (its smallest flat
index) and the same component sizes."""
import sys

import numpy as np
import pytest
import torch
import torch.nn.functional as F

from helpers import assert_close, oracle_of
from test_gpu_render import make_model
from oracle import alpha_grid_sample
from oracle.render import remove_small_objects
from ffields.render import remove_small_components
import ffields._lib as L

pytestmark = pytest.mark.gpu


def test_pool_threshold_matches_max_pool3d():
    g = torch.Generator().manual_seed(0)
    a = torch.rand((13, 17, 11), generator=g) * 3 - 1     # values outside [0,1] exercise the clamp
    out = torch.empty((13, 17, 11), dtype=torch.uint8, device="cuda")
    ac = a.cuda()
    L.check(L.lib().ff_occupancy_pool(L.ptr(ac), 13, 17, 11, 2.0, 0.3, L.ptr(out), L.stream_of(ac)))
    ref = F.max_pool3d((a / 2.0).clamp(0, 1)[None, None], kernel_size=3, padding=1, stride=1)[0, 0] >= 0.3
    assert torch.equal(out.cpu().bool(), ref)


def test_dense_alpha_matches_oracle():
    m = make_model(seed=3)
    torch.manual_seed(0)
    alpha, dense = m.getDenseAlpha([9, 10, 11], times=1)
    of = oracle_of(m)
    x = dense.reshape(-1, 3).cpu()
    sig = of.basis2density(of.field(x)[:, 0])
    am = m.alphaMask
    keep = alpha_grid_sample(am.alpha_volume.cpu()[0, 0], am.aabb.cpu(), x) > 0
    units = (m.aabb[1] - m.aabb[0]).cpu() / (torch.tensor([9, 10, 11]) - 1)
    length = float(units.mean() * 25.0)                # getDenseAlpha's own step (FactorFields.py:735-737)
    ref = torch.where(keep, 1 - torch.exp(-sig * length), torch.zeros_like(sig))
    assert_close(alpha.reshape(-1), ref, 1e-5, "alpha")


def test_update_alpha_mask_runs_and_shrinks_box():
    m = make_model(seed=5, occupancy=False)
    box = m.updateAlphaMask((16, 16, 16), is_update_alphaMask=True, times=2)
    assert m.alphaMask is not None
    assert box.shape == (2, 3)
    occ = m.alphaMask.occ
    assert occ.shape == (16, 16, 16) and int(occ.sum()) > 0


def _components(mask_np, min_size):
    """The device filter with its labels and sizes, beside the oracle's."""
    m = torch.from_numpy(mask_np.astype(np.uint8)).cuda()
    D, H, W = m.shape
    labels = torch.empty((D, H, W), dtype=torch.int32, device="cuda")
    sizes = torch.empty((D, H, W), dtype=torch.int32, device="cuda")
    out = torch.empty_like(m)
    L.check(L.lib().ff_remove_small_components(L.ptr(m), D, H, W, float(min_size), L.ptr(labels), L.ptr(sizes),
                                               L.ptr(out), L.stream_of(m)), "ff_remove_small_components")
    lab = labels.cpu().numpy().astype(np.int64)
    size = np.where(lab >= 0, sizes.cpu().numpy().reshape(-1)[np.maximum(lab, 0)], 0)
    return out.cpu().numpy().astype(bool), lab, size, remove_small_objects(mask_np, min_size)


def _spiral(n):
    """A one-voxel-wide serpentine through a [n, n, n] box: one component of
    about n^3 / 4 voxels whose path visits the box's indices back and forth."""
    m = np.zeros((n, n, n), bool)
    for z in range(0, n, 2):
        for y in range(0, n, 2):
            m[z, y, :] = True
            xe = n - 1 if (y // 2) % 2 == 0 else 0
            if y + 2 < n:
                m[z, y + 1, xe] = True
        ye = n - 1 if n % 2 else n - 2
        if z + 2 < n:
            m[z + 1, ye if (z // 2) % 2 == 0 else 0, (n - 1) if ((ye // 2) % 2 == 0) else 0] = True
    return m


@pytest.mark.parametrize("case", ["sparse", "percolating", "dense", "thin", "ragged", "column", "all", "none",
                                  "spiral"])
def test_small_components_match_oracle(case):
    print(f"[case] remove_small_components: {case}", file=sys.stderr, flush=True)
    rng = np.random.default_rng(7)
    shape = {"thin": (1, 1, 300), "ragged": (3, 41, 17), "column": (40, 30, 1)}.get(case, (29, 23, 31))
    if case == "spiral":
        m = _spiral(33)
    elif case == "all":
        m = np.ones(shape, bool)
    elif case == "none":
        m = np.zeros(shape, bool)
    else:
        p = {"sparse": 0.15, "percolating": 0.32, "dense": 0.7, "thin": 0.8, "ragged": 0.45, "column": 0.6}[case]
        m = rng.random(shape) < p
    ms = float(np.mean(m.shape))
    out, lab, size, (r_out, r_first, r_size) = _components(m, ms)
    assert np.array_equal(lab.reshape(m.shape), r_first), case
    assert np.array_equal(size.reshape(m.shape), r_size), case
    assert np.array_equal(out, r_out), case


def test_small_components_size_boundary():
    """A component of exactly min_size voxels is kept, one voxel fewer is not."""
    print("[case] remove_small_components: size boundary", file=sys.stderr, flush=True)
    m = np.zeros((4, 5, 12), bool)
    m[1, 1, 0:7] = True          # 7 voxels
    m[3, 3, 0:6] = True          # 6 voxels
    out = _components(m, 7.0)[0]
    assert out[1, 1, 0:7].all() and not out[3].any()
    out = _components(m, 6.5)[0]
    assert out[1, 1, 0:7].all() and not out[3].any()
    out = _components(m, 6.0)[0]
    assert out[1, 1, 0:7].all() and out[3, 3, 0:6].all()


def test_small_components_default_grid():
    """The default update grid (200^3) with a solid ball, a hollow shell and
    scattered floaters: identical to the oracle."""
    print("[case] remove_small_components: 200^3", file=sys.stderr, flush=True)
    n = 200
    g = np.linspace(-1, 1, n)
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    r = np.sqrt(x * x + y * y + z * z)
    m = (r < 0.3) | ((r > 0.6) & (r < 0.62)) | (np.random.default_rng(3).random((n, n, n)) < 0.02)
    out, lab, size, (r_out, r_first, r_size) = _components(m, float(n))
    assert np.array_equal(out, r_out)
    assert np.array_equal(lab.reshape(m.shape), r_first)
    assert np.array_equal(size.reshape(m.shape), r_size)


def test_update_alpha_mask_filter_matches_oracle():
    """updateAlphaMask's mask equals the oracle filter of the pooled,
    thresholded dense alpha volume of the same jitter."""
    print("[case] updateAlphaMask vs oracle filter", file=sys.stderr, flush=True)
    m = make_model(seed=6, occupancy=False)
    grid = (20, 18, 22)
    alpha, _ = m.getDenseAlpha(grid, times=2, generator=torch.Generator(device="cuda").manual_seed(4))
    m.updateAlphaMask(grid, is_update_alphaMask=True, times=2,
                      generator=torch.Generator(device="cuda").manual_seed(4))
    pooled = F.max_pool3d((alpha / 1.0).clamp(0, 1)[None, None], kernel_size=3, padding=1, stride=1)[0, 0]
    thres = float(m.cfg.renderer.alphaMask_thres)
    want = remove_small_objects((pooled >= thres).cpu().numpy(), float(np.mean(pooled.shape)))[0]
    got = m.alphaMask.alpha_volume[0, 0].cpu().numpy() > 0.5
    assert want.any() and np.array_equal(got, want)


def test_small_components_host_tensor_refused():
    with pytest.raises(ValueError):
        remove_small_components(torch.zeros((3, 3, 3), dtype=torch.uint8), 2.0)
