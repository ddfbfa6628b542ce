"""Pins the proposal oracle (oracle/proposal.py): bin bookkeeping of the
interlevel bound loss against an explicit interval-overlap loop, the loss's
zero set, and its gradient against FP64 finite differences."""
import torch

from oracle.proposal import interlevel_loss, ray_box


def _overlap_bound(bins, w, t0, t1):
    tot = 0.0
    for j in range(len(w)):
        if bins[j] < t1 and bins[j + 1] > t0:
            tot += float(w[j])
    return tot


def test_bound_matches_explicit_overlap():
    g = torch.Generator().manual_seed(0)
    for _ in range(20):
        bins = torch.sort(torch.rand(9, generator=g, dtype=torch.float64))[0]
        w = torch.rand(8, generator=g, dtype=torch.float64)
        t = torch.sort(bins[0] + (bins[-1] - bins[0]) * torch.rand(6, generator=g, dtype=torch.float64))[0]
        wf = torch.rand(6, generator=g, dtype=torch.float64)
        got = interlevel_loss(bins[None], w[None], t[None], wf[None])[0]
        ref = 0.0
        for k in range(5):
            b = _overlap_bound(bins, w, float(t[k]), float(t[k + 1]))
            if b == 0.0:                      # interval inside one bin edge: the containing bin
                j = int((bins[1:] <= t[k]).sum())
                b = float(w[min(j, 7)])
            gap = max(0.0, float(wf[k]) - b)
            ref += gap * gap / (float(wf[k]) + 1e-7)
        assert abs(float(got) - ref) < 1e-12


def test_loss_zero_when_proposal_covers():
    bins = torch.linspace(0, 1, 9, dtype=torch.float64)[None]
    w = torch.full((1, 8), 0.5, dtype=torch.float64)
    t = torch.linspace(0.05, 0.95, 7, dtype=torch.float64)[None]
    wf = torch.full((1, 7), 0.3, dtype=torch.float64)
    assert float(interlevel_loss(bins, w, t, wf)[0]) == 0.0


def test_loss_gradient_fd():
    g = torch.Generator().manual_seed(1)
    bins = torch.sort(torch.rand((2, 9), generator=g, dtype=torch.float64), -1)[0]
    w = (torch.rand((2, 8), generator=g, dtype=torch.float64) * 0.1).requires_grad_(True)
    t = torch.sort(torch.rand((2, 6), generator=g, dtype=torch.float64), -1)[0] * 0.8 + 0.1
    wf = torch.rand((2, 6), generator=g, dtype=torch.float64)
    interlevel_loss(bins, w, t, wf).sum().backward()
    eps = 1e-7
    for r in range(2):
        for j in range(8):
            wp, wm = w.detach().clone(), w.detach().clone()
            wp[r, j] += eps
            wm[r, j] -= eps
            fd = (interlevel_loss(bins, wp, t, wf).sum() - interlevel_loss(bins, wm, t, wf).sum()) / (2 * eps)
            assert abs(float(fd) - float(w.grad[r, j])) < 1e-5


def test_ray_box():
    rays = torch.tensor([[0.0, 0.0, -5.0, 0.0, 0.0, 1.0], [0.0, 0.0, -5.0, 1.0, 0.0, 0.0]])
    tn, tf, hit = ray_box(rays, torch.tensor([[-1.0, -1.0, -1.0], [1.0, 1.0, 1.0]]))
    assert hit.tolist() == [True, False]
    assert abs(float(tn[0]) - 4.0) < 1e-6 and abs(float(tf[0]) - 6.0) < 1e-6
