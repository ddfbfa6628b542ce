"""The committed golden fixtures (tests/golden/*.npz, written by
tests/golden/make_golden.py) against the oracle recomputed here, and the
fixtures' own consistency checks that do not depend on the oracle's code:
sample counts equal the mask sizes, fine samples are sorted and inside their
bins, field outputs are the MLP applied to the stored features."""
import json
import os
import sys

import numpy as np
import pytest
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import make_golden as MG  # noqa: E402
from oracle import grid_mapping, sample_pdf  # noqa: E402
from oracle.proposal import proposal_forward  # noqa: E402


def load(name):
    with np.load(os.path.join(HERE, "golden", name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.mark.parametrize("name", list(MG.FIELD_CASES))
def test_field_fixture(name):
    fx = load(name)
    cfg = json.loads(str(fx["cfg"]))
    assert cfg == MG.small_cfg(**MG.FIELD_CASES[name])
    new = MG.make_field(name, MG.FIELD_CASES[name])
    for k in fx:
        if k == "cfg":
            continue
        tol = 1e-6 if k.endswith("32") else 1e-12
        np.testing.assert_allclose(new[k], fx[k], rtol=tol, atol=tol, err_msg=f"{name}:{k}")
    # independent of the gather code: out = W1 relu(W0 feats + b0)
    h = np.maximum(fx["feats"] @ fx["w0"].astype(np.float64).T + fx["b0"], 0)
    np.testing.assert_allclose(h @ fx["w1"].astype(np.float64).T, fx["out"], rtol=1e-12, atol=1e-12)
    # the FP32 run of the same operations stays within FP32 rounding of the FP64 one
    scale = np.abs(fx["feats"]).max()
    assert np.abs(fx["feats32"] - fx["feats"]).max() <= 1e-4 * scale


def test_mapping_fixture():
    fx = load("mapping")
    x, freq, aabb = (torch.from_numpy(fx[k]) for k in ("pos", "freq", "aabb"))
    for kind in ("sawtooth", "triangle", "sinc", "trigonometric", "x"):
        np.testing.assert_array_equal(grid_mapping(x, freq, aabb, kind).numpy(), fx[kind], err_msg=kind)
    # closed form of the x mapping: (p - aabb0) / (max extent / f)
    scale = (aabb[1] - aabb[0]).max() / freq
    np.testing.assert_allclose(fx["x"], ((x - aabb[0])[..., None] / scale).numpy(), rtol=1e-15)
    assert np.all(np.abs(fx["sawtooth"]) <= 1.0)


def test_render_fixture():
    fx = load("render")
    new = MG.make_render()
    for k in fx:
        if k == "cfg":
            assert json.loads(str(fx[k])) == json.loads(str(new[k]))
            continue
        np.testing.assert_allclose(new[k], fx[k], rtol=1e-6, atol=1e-7, err_msg=k)
    assert int(fx["counts"].sum()) == len(fx["idx"]) == len(fx["sigma"])
    assert 0 < len(fx["idx"]) < 96 * int(fx["n_samples"])
    assert np.all((fx["rgb_map"] >= 0) & (fx["rgb_map"] <= 1))


def test_sample_pdf_fixture():
    fx = load("sample_pdf")
    bins, w, u = (torch.from_numpy(fx[k]) for k in ("bins", "weights", "u"))
    np.testing.assert_array_equal(sample_pdf(bins, w, 48, u=u).numpy(), fx["samples"])
    np.testing.assert_array_equal(sample_pdf(bins, w, 48, det=True).numpy(), fx["det"])
    s = fx["samples"]
    assert np.all(s >= fx["bins"][:, :1] - 1e-6) and np.all(s <= fx["bins"][:, -1:] + 1e-6)
    assert np.all(np.diff(s, axis=-1) >= 0)


def test_proposal_fixture():
    fx = load("proposal")
    net = dict(grid=torch.from_numpy(fx["grid"]), w0=torch.from_numpy(fx["w0"]), b0=torch.from_numpy(fx["b0"]),
               w1=torch.from_numpy(fx["w1"]), aabb=torch.from_numpy(fx["aabb"]), near=0.05, shift=-1.0,
               distance_scale=25.0, n_fine=48)
    rays = torch.from_numpy(fx["rays"])
    ref = proposal_forward(net, rays, torch.from_numpy(fx["prop_u"]), torch.from_numpy(fx["fine_u"]))
    for k in ("hit", "bins", "w_prop", "t_fine"):
        np.testing.assert_allclose(ref[k].numpy(), fx[k], rtol=1e-6, atol=1e-7, err_msg=k)
    assert not fx["hit"][:2].any() and fx["hit"].sum() > 48
    assert np.all(np.diff(fx["t_fine"][fx["hit"]], axis=-1) >= 0)
    w = fx["w_prop"][fx["hit"]]                       # rays that miss the box are never sampled
    assert np.all(w >= 0) and np.all(w.sum(-1) <= 1 + 1e-6)


def test_dataset_fixture():
    fx = load("datasets")
    new = MG.make_datasets()
    for k in fx:
        np.testing.assert_allclose(new[k], fx[k], rtol=1e-5, atol=1e-6, err_msg=k)
