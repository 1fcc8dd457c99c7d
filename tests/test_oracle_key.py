"""CPU checks of the oracle's learned-key functions (forward pass and triplet-loss gradients).

The key network has no counterpart in the reference; these checks pin the oracle by known
answers (selection weights reproduce patch pixels, so the patch order x[dy*24 + dx*3 + c] and the
plane layout are fixed), by the BF16 round-to-nearest-even cases, and by central finite
differences of the loss for the gradients.
"""
import numpy as np
import pytest

from oracle_ffi import key_forward, key_gradients, key_loss, key_round_weights

KI, KH, KO = 192, 64, 8


def xy(x, y):
    return (y << 16) | x


def selection_net(sel, w1_value=1.0, offset=0.5):
    """W1 row o picks input sel[o] (weight w1_value), b1 = offset (0.5 keeps the ReLU open for
    inputs in [-0.5, 0.5]), W2 picks hidden o, b2 = -offset: key o = w1_value * x[sel[o]]. The
    hidden unit w1_value * x + offset enters layer 2 rounded to BF16, so the known answers use
    values that BF16 holds exactly."""
    w1 = np.zeros((KH, KI), np.float32)
    for o, k in enumerate(sel):
        w1[o, k] = w1_value
    b1 = np.full(KH, offset, np.float32)
    w2 = np.zeros((KO, KH), np.float32)
    for o in range(KO):
        w2[o, o] = 1.0
    b2 = np.full(KO, -offset, np.float32)
    return w1, b1, w2, b2


def test_forward_selection_known_answer():
    rng = np.random.default_rng(5)
    h, w = 21, 30
    img = (rng.integers(0, 256, (h, w, 3)) / 256.0).astype(np.float32)   # exact in BF16 after - 0.5
    taps = [(0, 0, 0), (7, 7, 2), (3, 5, 1), (0, 7, 2), (7, 0, 0), (2, 2, 2), (6, 1, 1), (4, 4, 0)]
    sel = [dy * 24 + dx * 3 + c for dy, dx, c in taps]
    planes = key_forward(*selection_net(sel), img)
    for o, (dy, dx, c) in enumerate(taps):
        expect = img[dy:dy + h - 7, dx:dx + w - 7, c] - 0.5
        np.testing.assert_array_equal(planes[o, 7:h, 7:w], expect)
    mask = np.ones((h + 7, w + 7), bool)
    mask[7:h, 7:w] = False
    assert not planes[:, mask].any()


def test_forward_rounds_weights_to_bf16_nearest_even():
    img = np.full((8, 8, 3), 0.75, np.float32)   # x = 0.25 everywhere: the ReLU is open without an offset
    sel = [0] * KO
    # 1 + 2^-8 is halfway between BF16 1 and 1 + 2^-7: ties to even -> 1.
    k = key_forward(*selection_net(sel, 1.0 + 2.0 ** -8, 0.0), img)[0, 7, 7]
    assert k == np.float32(0.25)
    # 1 + 3 * 2^-8 is halfway between 1 + 2^-7 and 1 + 2^-6: ties to even -> 1 + 2^-6.
    k = key_forward(*selection_net(sel, 1.0 + 3 * 2.0 ** -8, 0.0), img)[0, 7, 7]
    assert k == np.float32(0.25 * (1 + 2.0 ** -6))
    # 1 + 2^-8 + 2^-12 is above the half: rounds up to 1 + 2^-7.
    k = key_forward(*selection_net(sel, 1.0 + 2.0 ** -8 + 2.0 ** -12, 0.0), img)[0, 7, 7]
    assert k == np.float32(0.25 * (1 + 2.0 ** -7))


def test_forward_rounds_hidden_units_to_bf16_nearest_even():
    """The hidden units enter layer 2 in BF16: with b1 = 0.5 the hidden unit of x = 0.25 and
    weight 1 + 2^-7 is 0.75 + 2^-9, halfway between the BF16 values 0.75 and 0.75 + 2^-8: ties to
    even give 0.75, so the key is 0.25; weight 1 + 2^-6 gives 0.75 + 2^-8, exact."""
    img = np.full((8, 8, 3), 0.75, np.float32)
    sel = [0] * KO
    k = key_forward(*selection_net(sel, 1.0 + 2.0 ** -7), img)[0, 7, 7]
    assert k == np.float32(0.25)
    k = key_forward(*selection_net(sel, 1.0 + 2.0 ** -6), img)[0, 7, 7]
    assert k == np.float32(0.25 + 2.0 ** -8)


def random_case(seed, nt=48, size=40):
    rng = np.random.default_rng(seed)
    a = rng.random((size, size, 3), dtype=np.float32)
    b = rng.random((size, size, 3), dtype=np.float32)
    w1 = ((rng.random((KH, KI)) * 2 - 1) * np.sqrt(6 / KI)).astype(np.float32)
    b1 = (rng.standard_normal(KH) * 0.05).astype(np.float32)
    w2 = ((rng.random((KO, KH)) * 2 - 1) * np.sqrt(6 / KH)).astype(np.float32)
    b2 = (rng.standard_normal(KO) * 0.05).astype(np.float32)
    n = size - 7
    pos = np.array([[xy(*rng.integers(0, n, 2)) for _ in range(3)] for _ in range(nt)], np.int32)
    return (w1, b1, w2, b2), a, b, pos


@pytest.mark.parametrize("seed", [1, 2])
def test_gradients_match_finite_differences(seed):
    ws, a, b, pos = random_case(seed)
    margin = 1.0
    key_round_weights(False)
    try:
        grads, loss = key_gradients(*ws, a, b, pos, margin)
        wd = [x.astype(np.float64) for x in ws]
        assert loss == pytest.approx(key_loss(*wd, a, b, pos, margin), rel=1e-6)
        assert loss > 0
        rng = np.random.default_rng(seed + 100)
        eps = 1e-6
        # the loss only sees key differences, so the output bias has no gradient
        assert np.abs(grads[3]).max() < 1e-9
        for ti, g in enumerate(grads[:3]):
            scale = np.abs(g).max()
            assert scale > 0
            flat = g.reshape(-1)
            idx = list(rng.choice(flat.size, size=min(12, flat.size), replace=False)) + [int(np.abs(flat).argmax())]
            for i in idx:
                plus = [x.copy() for x in wd]
                minus = [x.copy() for x in wd]
                plus[ti].reshape(-1)[i] += eps
                minus[ti].reshape(-1)[i] -= eps
                fd = (key_loss(*plus, a, b, pos, margin) - key_loss(*minus, a, b, pos, margin)) / (2 * eps)
                assert abs(fd - flat[i]) <= 1e-4 * scale + 1e-3 * abs(fd), (ti, i, fd, flat[i])
    finally:
        key_round_weights(True)


def test_inactive_triplets_have_no_gradient():
    ws, a, b, pos = random_case(3, nt=8)
    pos[:, 1] = pos[:, 0]          # positive = anchor position, different image
    grads, loss = key_gradients(*ws, a, a, pos, margin=-1e6)
    assert loss == 0
    assert all(not g.any() for g in grads)
