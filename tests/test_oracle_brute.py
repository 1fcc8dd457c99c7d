"""CPU checks of the oracle's FP64 exhaustive nearest-neighbour search against an independent
numpy formulation (all query/database pairs at once), on small images."""
import numpy as np

from conftest import seeded
from oracle_ffi import brute_f64


def xy(x, y):
    return (y << 16) | x


def patches(img):
    h, w, _ = img.shape
    out = np.empty((h - 7, w - 7, 192), np.float64)
    for dy in range(8):
        for dx in range(8):
            out[:, :, dy * 24 + dx * 3:dy * 24 + dx * 3 + 3] = img[dy:dy + h - 7, dx:dx + w - 7]
    return out


def test_brute_f64_matches_numpy():
    a = seeded((20, 23, 3), 1)
    b = seeded((18, 26, 3), 2)
    qx, qy = np.meshgrid(np.arange(23 - 7), np.arange(20 - 7))
    qpos = np.array([xy(x, y) for x, y in zip(qx.ravel(), qy.ravel())], np.int32)
    pos, dist = brute_f64(a, b, qpos)
    A = patches(a.astype(np.float64)).reshape(-1, 192)
    B = patches(b.astype(np.float64)).reshape(-1, 192)
    D = ((A[:, None, :] - B[None, :, :]) ** 2).sum(-1)
    best = D.argmin(1)
    dew = 26 - 7
    np.testing.assert_array_equal(pos, [xy(i % dew, i // dew) for i in best])
    np.testing.assert_allclose(dist, D.min(1), rtol=1e-12)


def test_brute_f64_finds_planted_copies():
    """A query patch copied into the database is found at distance 0."""
    a = seeded((30, 30, 3), 4)
    b = seeded((40, 44, 3), 5)
    plants = [((3, 5), (20, 11)), ((10, 0), (0, 30)), ((22, 22), (36, 32))]
    for (qx, qy), (bx, by) in plants:
        b[by:by + 8, bx:bx + 8] = a[qy:qy + 8, qx:qx + 8]
    qpos = np.array([xy(qx, qy) for (qx, qy), _ in plants], np.int32)
    pos, dist = brute_f64(a, b, qpos)
    assert list(pos) == [xy(bx, by) for _, (bx, by) in plants]
    assert not dist.any()
