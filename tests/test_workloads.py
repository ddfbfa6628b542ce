"""Synthetic workload generators: the batched image-set generator reproduces
the single-image generator exactly (same draws, same accumulation order)."""
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
