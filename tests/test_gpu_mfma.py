"""GPU: the MFMA lookup (mfma_rounds > 0): dense int8 candidate scoring on matrix cores with exact
rescoring. Bit-exact against the oracle's restatement (lookup_parallel dispatches to it) on
partial groups, bands that leave b, several tables, previous fields, few and many anchors and
every jump schedule; every NNF_DIST is the exact SSD of its match; quality within 1% of the
raster-order restatement at 1024x768."""
import numpy as np
import pytest

import oracle_ffi as O
from conftest import speed1

pytestmark = pytest.mark.gpu


def _both(p, b):
    from patchtable_amd import PatchTable
    return PatchTable(p, b), O.OracleTable(p.raw, b)


@pytest.mark.parametrize("qh,qw,bh,bw,extra", [
    (96, 150, 110, 160, "-mfma_rounds 6"),                    # 143 columns: a partial last group
    (60, 75, 40, 38, "-mfma_rounds 6"),                       # a small b: bands leave the image
    (80, 119, 90, 100, "-mfma_rounds 1"),
    (80, 119, 90, 100, "-mfma_rounds 2 -mfma_anchors 1"),
    (80, 119, 90, 100, "-mfma_rounds 3 -mfma_anchors 16"),
    (70, 90, 120, 130, "-mfma_rounds 5 -ntables 2"),
    (40, 519, 60, 300, "-mfma_rounds 4"),                     # 512 columns: 8 tile columns (XCD bands)
    (80, 119, 90, 100, "-mfma_rounds 4 -mfma_anchors 8 -mfma_jumps 16,4,1,1"),
    (96, 150, 110, 160, "-mfma_rounds 6 -mfma_jumps 24,8,3"),  # rounds past the list: unit jumps
])
def test_mfma_lookup_bit_exact(qh, qw, bh, bw, extra):
    from patchtable_amd import synth_pair
    print(f"mfma: query {qh}x{qw} database {bh}x{bw} {extra}", flush=True)
    a, _ = synth_pair(500 + qh, qh, qw)
    _, b = synth_pair(600 + bw, bh, bw)
    p = speed1("-limit 1 " + extra)
    t, o = _both(p, b)
    assert t.info()["kcoherence"] == 0 == o.K   # no lists: the MFMA lookup walks none
    ann = t.lookup(a)
    ref = o.lookup_parallel(a)
    assert np.array_equal(ann, ref)
    m, d64 = O.recompute(a, b, ann)
    assert np.abs(d64 - ann[..., 2]).max() <= 1e-4


def test_mfma_item_counts_match_oracle():
    """The instrumentation's per-round item counts (pt_table_mfma_items) equal the oracle's, and
    are off until asked for."""
    from patchtable_amd import synth_pair
    a, _ = synth_pair(530, 90, 140)
    _, b = synth_pair(531, 100, 120)
    p = speed1("-limit 1 -mfma_rounds 4")
    t, o = _both(p, b)
    t.lookup(a)
    assert t.mfma_items() == []
    t.set_mfma_stats(True)
    ann = t.lookup(a)
    assert np.array_equal(ann, o.lookup_parallel(a))
    items = t.mfma_items()
    assert len(items) == 4 and items == O.mfma_items(), (items, O.mfma_items())
    groups = (140 - 7 + 15) // 16 * (90 - 7)
    assert all(groups <= n <= 6 * groups for n in items)


def test_mfma_lookup_previous_field():
    """Starting from a previous field (with temporal weighting and recomputed distances)."""
    from patchtable_amd import synth_pair
    a, b = synth_pair(510, 90, 120)
    a2, _ = synth_pair(511, 90, 120)
    p = speed1("-limit 1 -mfma_rounds 4 -coherence_temporal 0.5 -recalc_dist_temporal 1")
    t, o = _both(p, b)
    prev = t.lookup(a)
    assert np.array_equal(prev, o.lookup_parallel(a))
    ann = t.lookup(a2, ann_prev=prev)
    assert np.array_equal(ann, o.lookup_parallel(a2, prev=prev))


def test_mfma_lookup_quality_cfg2():
    """1024x768 (seed 2), speed 1 + 6 MFMA rounds: bit-exact with the oracle; mean SSD and
    mean sqrt SSD within 1% of the raster-order restatement (the reference's algorithm)."""
    from patchtable_amd import synth_pair
    print("mfma cfg2 1024x768", flush=True)
    a, b = synth_pair(2, 768, 1024)
    p = speed1("-mfma_rounds 6")
    t, o = _both(p, b)
    ann = t.lookup(a)
    assert np.array_equal(ann, o.lookup_parallel(a))
    praster = speed1()
    oraster = O.OracleTable(praster.raw, b)
    ref = oraster.lookup_raster(a)
    r_ssd = ann[..., 2].mean() / ref[..., 2].mean()
    r_dist = O.mean_dist(ann) / O.mean_dist(ref)
    print(f"mfma cfg2: mean SSD {ann[..., 2].mean():.6f} vs raster {ref[..., 2].mean():.6f} ({r_ssd:.5f}); "
          f"mean_dist ratio {r_dist:.5f}", flush=True)
    assert 0.99 <= r_ssd <= 1.01 and 0.99 <= r_dist <= 1.01


def test_mfma_lookup_rejects_unsupported():
    from patchtable_amd import PatchTable, PatchTableError, synth_pair
    a, b = synth_pair(520, 40, 50)
    t = PatchTable(speed1("-limit 0.1 -mfma_rounds 2 -query_step 2"), b)
    with pytest.raises(PatchTableError, match="query_step 1"):
        t.lookup(a)
    mask = np.ones((33, 43), np.int32)
    t = PatchTable(speed1("-limit 0.1 -mfma_rounds 2"), b, allowed_patches=mask)
    with pytest.raises(PatchTableError, match="mask"):
        t.lookup(a)
    p = speed1("-limit 0.1 -mfma_rounds 2 -mfma_anchors 17")
    t = PatchTable(p, b)
    with pytest.raises(PatchTableError, match="mfma_anchors"):
        t.lookup(a)
