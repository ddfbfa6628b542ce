"""CPU test of the small-object filter used by updateAlphaMask
(models/FactorFields.py:774-775, skimage remove_small_objects with
connectivity=1): the oracle's restatement checked against a brute-force
6-neighbour flood fill.  The device filter is compared with the oracle in
tests/test_gpu_occupancy.py."""
import numpy as np

from oracle.render import remove_small_objects


def _remove_small_objects(mask, min_size):
    return remove_small_objects(mask, min_size)[0]


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


def _component(mask, seed):
    seen, stack = {seed}, [seed]
    while stack:
        c = stack.pop()
        for d in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)):
            n = tuple(int(c[k] + d[k]) for k in range(3))
            if all(0 <= n[k] < mask.shape[k] for k in range(3)) and mask[n] and n not in seen:
                seen.add(n)
                stack.append(n)
    return seen


def test_remove_small_objects_matches_flood_fill():
    rng = np.random.default_rng(0)
    for trial in range(5):
        m = rng.random((9, 11, 7)) < 0.35
        ms = float(np.mean(m.shape))
        assert np.array_equal(_remove_small_objects(m, ms), _flood(m, ms))
        _, first, size = remove_small_objects(m, ms)
        # the component labelling: smallest flat index of the component, its size
        kept = _flood(m, 1)
        assert np.array_equal(first >= 0, kept) and np.array_equal(size > 0, kept)
        z, y, x = np.nonzero(m)
        for a, b, c in zip(z[:20], y[:20], x[:20]):
            comp = _component(m, (a, b, c))
            assert first[a, b, c] == min(np.ravel_multi_index(v, m.shape) for v in comp)
            assert size[a, b, c] == len(comp)


def test_diagonal_contact_is_not_connected():
    m = np.zeros((3, 3, 3), bool)
    m[0, 0, 0] = m[1, 1, 1] = True          # corner contact only
    out = _remove_small_objects(m, 2)
    assert not out.any()
