"""GPU: the learned patch key. The MFMA forward pass and the training kernels' gradients against
the oracle; a learned-key table (key table + GCK fallback) bit-exact against the oracle's
construction and parallel lookup from the same key; training lowers the loss; and what a
trained key gains over the plain GCK table on a held-out pair."""
import numpy as np
import pytest

import oracle_ffi as O
from conftest import speed1

pytestmark = pytest.mark.gpu

KI, KH, KO = 192, 64, 8


def xy(x, y):
    return (y << 16) | x


def test_key_forward_selection_known_answer():
    """Selection weights make each key a pixel of the patch: pins the MFMA operand layouts, the
    patch order and the plane layout without the oracle."""
    import torch
    from patchtable_amd import PatchKey
    from test_oracle_key import selection_net
    rng = np.random.default_rng(11)
    h, w = 37, 150                       # 143 patches a row: two full 64-patch blocks and a tail
    img = (rng.integers(0, 256, (h, w, 3)) / 256.0).astype(np.float32)
    taps = [(0, 0, 0), (7, 7, 2), (3, 5, 1), (0, 7, 2), (7, 0, 0), (2, 2, 2), (6, 1, 1), (4, 4, 0)]
    key = PatchKey(1)
    key.set_weights(*selection_net([dy * 24 + dx * 3 + c for dy, dx, c in taps]))
    planes = key.forward(torch.from_numpy(img).cuda()).cpu().numpy()
    for o, (dy, dx, c) in enumerate(taps):
        np.testing.assert_array_equal(planes[o, 7:h, 7:w], img[dy:dy + h - 7, dx:dx + w - 7, c] - 0.5)
    mask = np.ones((h + 7, w + 7), bool)
    mask[7:h, 7:w] = False
    assert not planes[:, mask].any()


@pytest.mark.parametrize("shape", [(8, 8), (70, 133), (300, 400)])
def test_key_forward_matches_oracle(shape):
    import torch
    from patchtable_amd import PatchKey
    h, w = shape
    img = np.random.default_rng(h * 1000 + w).random((h, w, 3), dtype=np.float32)
    key = PatchKey(7)
    (w1, b1, w2, b2), _ = key.weights()
    got = key.forward(torch.from_numpy(img).cuda()).cpu().numpy()
    # the oracle rounds the masters, the inputs and the hidden units to BF16 as the GPU does and
    # accumulates in double; the GPU accumulates in FP32, so a hidden unit within an FP32 rounding
    # of a BF16 tie may round the other way (one BF16 ulp of that unit, 2^-7 of it at most, times
    # its W2 weight): nearly every key matches to FP32 accuracy, every key to that bound
    ref = O.key_forward(w1, b1, w2, b2, img)
    scale = np.abs(ref).max() + 1.0
    diff = np.abs(got - ref)
    err = diff.max()
    close = (diff <= 2e-5 * scale).mean()
    print(f"{shape}: max |key| {scale - 1:.4f}, max error {err:.3g}, within 2e-5 of the scale: {close:.5f}",
          flush=True)
    assert close >= 0.999
    # the bound: one BF16 ulp (at most 2^-7 of the unit) of every hidden unit, through |W2|
    bf = lambda v: ((v.astype(np.float32).view(np.uint32).astype(np.uint64) + 0x7FFF
                     + ((v.astype(np.float32).view(np.uint32) >> 16) & 1)) >> 16 << 16).astype(np.uint32).view(np.float32)
    win = np.lib.stride_tricks.sliding_window_view(img, (8, 8, 3))[:, :, 0]   # (h-7, w-7, 8, 8, 3)
    x = bf(win.reshape(-1, 192) - np.float32(0.5)).astype(np.float64)
    hid = np.maximum(x @ bf(w1).astype(np.float64).T + b1, 0.0)
    bound = (hid * 2.0 ** -7) @ np.abs(bf(w2).astype(np.float64)).T
    assert err <= 2e-5 * scale + bound.max()


def test_bf16_copies_round_to_nearest_even():
    from patchtable_amd import PatchKey
    key = PatchKey(3)
    (w1, _, w2, _), (w1b, w2b) = key.weights()
    for m, bf in ((w1, w1b), (w2, w2b)):
        u = m.view(np.uint32).astype(np.uint64)
        expect = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)
        assert np.array_equal(bf, expect)


def triplets(rng, n, npos, nt):
    return np.array([[xy(*rng.integers(0, n, 2)) for _ in range(npos)] for _ in range(nt)], np.int32)


@pytest.mark.parametrize("seed", [1, 2])
def test_key_gradients_match_oracle(seed):
    import torch
    from patchtable_amd import PatchKey, synth_pair
    a, b = synth_pair(300 + seed, 96, 96)
    rng = np.random.default_rng(seed)
    pos = triplets(rng, 96 - 7, 3, 384)
    key = PatchKey(seed)
    (w1, b1, w2, b2), _ = key.weights()
    margin = 0.5
    g_gpu, l_gpu = key.gradients(torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda(),
                                 torch.from_numpy(pos).cuda(), margin)
    g_ref, l_ref = O.key_gradients(w1, b1, w2, b2, a, b, pos, margin)
    print(f"loss gpu {l_gpu:.6f} oracle {l_ref:.6f}", flush=True)
    assert l_ref > 0
    assert l_gpu == pytest.approx(l_ref, rel=1e-4)
    for name, g, r in zip(("W1", "b1", "W2"), g_gpu, g_ref):
        rel = np.linalg.norm(g - r) / np.linalg.norm(r)
        print(f"{name}: relative error {rel:.3g}", flush=True)
        assert rel < 1e-3, name
    assert np.abs(g_gpu[3]).max() < 1e-5   # output bias: the loss sees key differences only


def keyed_case(p, size, seed, key_seed=5):
    import torch
    from patchtable_amd import PatchKey, PatchTable, synth_pair
    a, b = synth_pair(seed, *size)
    key = PatchKey(key_seed)
    t = PatchTable.keyed(p, key, torch.from_numpy(b).cuda())
    return a, b, key, t


@pytest.mark.parametrize("extra", ["", "-kcoherence_iter -1 -limit 0.2"])
def test_keyed_table_bit_exact(extra):
    """Given the GPU's keys, the key table, the GCK fallback, the k-coherence lists (in key
    space) and the lookup all equal the oracle's."""
    import torch
    p = speed1(extra)
    a, b, key, t = keyed_case(p, (300, 400), 41)
    assert t.info()["grid_ndims"] == KO
    kb = t.descriptor()
    o = O.OracleTable(p.raw, b, key=kb)
    assert t.info()["nslices"] == o.nslices
    assert t.info()["table_add_count"] == o.table_add_count
    assert np.array_equal(t.cells(), o.table())
    ns, add, cells = o.extra(0)
    assert t.extra_info(0)["table_add_count"] == add
    assert np.array_equal(t.extra_cells(0), cells)
    kp, kd = t.kcoherence()
    op, od = o.kcoherence()
    assert np.array_equal(kp, op) and np.array_equal(kd, od)
    ann = t.lookup(torch.from_numpy(a).cuda()).cpu().numpy()
    qk = t.query_descriptor(*a.shape[:2])
    ref = o.lookup_keyed(a, qk)
    assert np.array_equal(ann, ref)
    m, d = O.recompute(a, b, ann)
    assert np.allclose(ann[..., 2], d, rtol=1e-5, atol=1e-6)   # NNF_DIST is the match's exact SSD
    # the key planes the table holds are the network's forward pass
    planes = key.forward(torch.from_numpy(b).cuda()).cpu().numpy()
    assert np.array_equal(np.moveaxis(planes, 0, -1), kb)


def test_keyed_table_mfma_form_bit_exact():
    """A key table in the MFMA lookup form: its candidates come from the key table and the GCK
    fallback, then the MFMA rounds; bit-exact with the oracle given the GPU's keys."""
    import torch
    p = speed1("-limit 0.5 -mfma_rounds 4 -mfma_anchors 5")
    a, b, key, t = keyed_case(p, (180, 260), 44)
    assert t.info()["kcoherence"] == 0
    o = O.OracleTable(p.raw, b, key=t.descriptor())
    ann = t.lookup(torch.from_numpy(a).cuda()).cpu().numpy()
    assert np.array_equal(ann, o.lookup_keyed(a, t.query_descriptor(*a.shape[:2])))
    m, d = O.recompute(a, b, ann)
    assert np.allclose(ann[..., 2], d, rtol=1e-5, atol=1e-6)


def test_training_on_exact_nearest_neighbours():
    """Positives from the exhaustive search (positives="exact"): the loss falls, and the trained
    key's table candidates alone beat the untrained network's on a held-out pair."""
    import torch
    from patchtable_amd import PatchKey, PatchTable, synth_pair
    pairs = [tuple(torch.from_numpy(x).cuda() for x in synth_pair(2000 + i, 192, 256)) for i in range(4)]
    key = PatchKey(1)
    losses = key.train(None, pairs, tpp=2048, steps=200, batch=512, lr=2e-3, margin=1.0, seed=9, positives="exact")
    first, last = losses[:10].mean(), losses[-10:].mean()
    print(f"exact positives: loss {first:.4f} -> {last:.4f}", flush=True)
    assert np.isfinite(losses).all() and last < 0.7 * first
    a, b = synth_pair(3000, 256, 320)
    A, B = torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda()
    p = speed1("-kcoherence 0 -do_prop 0")
    trained = PatchTable.keyed(p, key, B).lookup(A).cpu().numpy()[..., 2].mean()
    untrained = PatchTable.keyed(p, PatchKey(1), B).lookup(A).cpu().numpy()[..., 2].mean()
    print(f"table candidates: trained {trained:.4f} untrained {untrained:.4f}", flush=True)
    assert trained < untrained


def test_keyed_table_rejects_save_and_descriptor_mode(tmp_path):
    import torch
    from patchtable_amd import PatchKey, PatchTable, PatchTableError, synth_pair
    p = speed1()
    a, b, key, t = keyed_case(p, (60, 80), 3)
    with pytest.raises(PatchTableError, match="no file format"):
        t.save(str(tmp_path / "k.table"))
    p2 = speed1()
    p2.is_descriptor = 1
    with pytest.raises(PatchTableError, match="is_descriptor"):
        PatchTable.keyed(p2, key, torch.from_numpy(b).cuda())
    p3 = speed1("-ntables 2")
    with pytest.raises(PatchTableError, match="ntables"):
        PatchTable.keyed(p3, key, torch.from_numpy(b).cuda())


def test_training_lowers_the_loss():
    import torch
    from patchtable_amd import PatchKey, synth_pair
    pairs = [tuple(torch.from_numpy(x).cuda() for x in synth_pair(2000 + i, 192, 256)) for i in range(4)]
    key = PatchKey(1)
    losses = key.train(speed1(), pairs, tpp=2048, steps=200, batch=512, lr=2e-3, margin=1.0, seed=9)
    first, last = losses[:10].mean(), losses[-10:].mean()
    print(f"loss {first:.4f} -> {last:.4f}", flush=True)
    assert np.isfinite(losses).all()
    assert last < 0.7 * first


def test_trained_key_gains_over_gck_table_candidates():
    """What the trained network gains: with table candidates alone (no k-coherence, no
    propagation) the key table + GCK fallback finds closer matches on a held-out pair than two
    GCK tables (the same number of table probes without the network) and than the same network
    untrained. Every distance is exact SSD, so a lower mean is a real gain."""
    import torch
    from patchtable_amd import PatchKey, PatchTable, synth_pair
    pairs = [tuple(torch.from_numpy(x).cuda() for x in synth_pair(2000 + i, 512, 512)) for i in range(8)]
    key = PatchKey(1)
    key.train(speed1(), pairs, tpp=8192, steps=400, batch=2048, lr=2e-3, margin=1.0, seed=9)
    a, b = synth_pair(3000, 768, 768)
    A, B = torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda()
    only = "-kcoherence 0 -do_prop 0"

    def mean_ssd(t):
        ann = t.lookup(A).cpu().numpy()
        m, d = O.recompute(a, b, ann)
        assert np.allclose(ann[..., 2], d, rtol=1e-5, atol=1e-6)
        return m

    gck2 = mean_ssd(PatchTable(speed1(only + " -ntables 2"), B))
    keyed = mean_ssd(PatchTable.keyed(speed1(only), key, B))
    untrained = mean_ssd(PatchTable.keyed(speed1(only), PatchKey(1), B))
    print(f"table candidates only: two GCK tables {gck2:.5f}, untrained key {untrained:.5f}, "
          f"trained key {keyed:.5f}", flush=True)
    assert keyed < gck2
    assert keyed < untrained
