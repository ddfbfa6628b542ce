"""CPU test of the small-object filter used by updateAlphaMask
(models/FactorFields.py:774-775, skimage remove_small_objects with
connectivity=1), checked against a brute-force 6-neighbour flood fill."""
import numpy as np

from ffields.render import _remove_small_objects


def _flood(mask, min_size):
    out = mask.copy()
    seen = np.zeros_like(mask)
    D, H, W = mask.shape
    for z in range(D):
        for y in range(H):
            for x in range(W):
                if not mask[z, y, x] or seen[z, y, x]:
                    continue
                stack, comp = [(z, y, x)], []
                seen[z, y, x] = True
                while stack:
                    c = stack.pop()
                    comp.append(c)
                    for dz, dy, dx in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)):
                        n = (c[0] + dz, c[1] + dy, c[2] + dx)
                        if 0 <= n[0] < D and 0 <= n[1] < H and 0 <= n[2] < W and mask[n] and not seen[n]:
                            seen[n] = True
                            stack.append(n)
                if len(comp) < min_size:
                    for c in comp:
                        out[c] = False
    return out


def test_remove_small_objects_matches_flood_fill():
    rng = np.random.default_rng(0)
    for trial in range(5):
        m = rng.random((9, 11, 7)) < 0.35
        ms = float(np.mean(m.shape))
        assert np.array_equal(_remove_small_objects(m, ms), _flood(m, ms))


def test_diagonal_contact_is_not_connected():
    m = np.zeros((3, 3, 3), bool)
    m[0, 0, 0] = m[1, 1, 1] = True          # corner contact only
    out = _remove_small_objects(m, 2)
    assert not out.any()
