"""Pins the oracle's restatement of the MFMA lookup (lookup_parallel with mfma_rounds > 0) by
brute force on fields whose anchors are known by construction.

One round (jump 1) from a previous field with exact distances (recalc_dist_temporal 1,
no table candidates): every 16-patch group of a row searches, for each of its anchors (ox, oy),
the database patches x0 + ox - 8 + j (j < 32) on rows y + oy - 1 .. y + oy + 1; each member
takes the candidate of smallest int8 SSD -- sum over the 8 x 8 x 3 values of
(round(255 v_a) - round(255 v_b))^2, computed here directly from the differences, not from
norms and dot products -- ties to the earlier (anchor, row, column); it replaces the match when
its int8 SSD is strictly below the int8 SSD of the match, and the replacing match's distance is
its exact FP32 SSD (the reference's lane order, pinned elsewhere); a patch that keeps its match
keeps the previous distance.

Fields:
  constant offset (0, 0): one distinct offset, so one anchor per group;
  offset (0, 0) left of column 24 and (12, 0) from column 24 on (b 20 pixels wider): groups 0
  sees only (0, 0), group 1 sees both with equal counts (the first slot, member 0, holds
  (0, 0)), groups 2 and 3 only (12, 0); 12 > 8 columns apart, so with two anchors group 1
  searches both, with one anchor only (0, 0).
"""
import numpy as np
import pytest

import oracle_ffi as O
from conftest import speed1


def q8(img):
    v = np.clip(img.astype(np.float32), np.float32(0), np.float32(1)) * np.float32(255)
    return np.rint(v).astype(np.int64)


def expected(a, b, prev, anchors_of_group):
    qa, qb = q8(a), q8(b)
    ah, aw = a.shape[0] - 7, a.shape[1] - 7
    beh, bew = b.shape[0] - 7, b.shape[1] - 7
    out = prev.copy()
    for y in range(ah):
        for x in range(aw):
            x0 = 16 * (x // 16)
            pa = qa[y:y + 8, x:x + 8]
            best = None
            for ai, (ox, oy) in enumerate(anchors_of_group(x0)):
                for dy in (-1, 0, 1):
                    for j in range(32):
                        xs, ys = x0 + ox - 8 + j, y + oy + dy
                        if not (0 <= xs < bew and 0 <= ys < beh):
                            continue
                        s = int(((pa - qb[ys:ys + 8, xs:xs + 8]) ** 2).sum())
                        key = (s, ai, dy, j)
                        if best is None or key < best[0]:
                            best = (key, xs, ys)
            cx, cy = int(prev[y, x, 0]), int(prev[y, x, 1])
            score = int(((pa - qb[cy:cy + 8, cx:cx + 8]) ** 2).sum())
            (s, *_), xs, ys = best
            if (xs, ys) != (cx, cy) and s < score:
                d = np.float32(O.lib().orc_patch_ssd(O.f32(np.ascontiguousarray(a)), a.shape[1], x, y,
                                                     O.f32(np.ascontiguousarray(b)), b.shape[1], xs, ys))
                out[y, x] = (xs, ys, d)
    return out


def field(a, b, offset_of_x):
    ah, aw = a.shape[0] - 7, a.shape[1] - 7
    prev = np.zeros((ah, aw, 3), np.float64)
    for y in range(ah):
        for x in range(aw):
            ox = offset_of_x(x)
            xs = x + ox
            prev[y, x] = (xs, y, O.lib().orc_patch_ssd(O.f32(np.ascontiguousarray(a)), a.shape[1], x, y,
                                                        O.f32(np.ascontiguousarray(b)), b.shape[1], xs, y))
    return prev


def run(a, b, prev, anchors):
    p = speed1(f"-limit 0.1 -mfma_rounds 1 -mfma_anchors {anchors} -do_table_lookup 0 -recalc_dist_temporal 1")
    o = O.OracleTable(p.raw, b)
    return o.lookup_parallel(a, prev=prev)


def test_mfma_restatement_constant_offset():
    from patchtable_amd import synth_pair
    a, b = synth_pair(41, 30, 60)
    _, b2 = synth_pair(42, 30, 60)
    prev = field(a, b2, lambda x: 0)
    got = run(a, b2, prev, 6)
    assert np.array_equal(got, expected(a, b2, prev, lambda x0: [(0, 0)]))
    assert (got[..., :2] != prev[..., :2]).any() and (got[..., :2] == prev[..., :2]).all(axis=2).any()


@pytest.mark.parametrize("anchors", [1, 2])
def test_mfma_restatement_two_offsets(anchors):
    from patchtable_amd import synth_pair
    a, _ = synth_pair(43, 30, 60)
    _, b = synth_pair(44, 30, 80)
    off = lambda x: 0 if x < 24 else 12
    prev = field(a, b, off)

    def anchors_of_group(x0):
        if x0 == 16:
            return [(0, 0), (12, 0)][:anchors]
        return [(0, 0)] if x0 == 0 else [(12, 0)]
    got = run(a, b, prev, anchors)
    assert np.array_equal(got, expected(a, b, prev, anchors_of_group))


@pytest.mark.parametrize("switches,prev", [
    ("-mfma_rounds 4 -mfma_anchors 7 -mfma_jumps 24,6,1,1", None),
    ("-mfma_rounds 6 -mfma_anchors 3", None),
    ("-mfma_rounds 5 -mfma_anchors 6 -mfma_jumps 8,2,1,1", "temporal"),
    ("-mfma_rounds 4 -mfma_anchors 5 -coherence_temporal -0.5", "temporal"),   # weights above 1: no skip
    ("-mfma_rounds 4 -mfma_anchors 5", "sv"),                                  # per-patch weights: no skip
])
def test_mfma_skip_of_searched_anchors_changes_nothing(switches, prev):
    """A round after the first does not search the anchors a group searched in its last round
    with anchors, when every score is at most the int8 SSD of its match: those bands hold no
    candidate strictly below any member's score. The fields with and without the skip are equal
    bit for bit; with the skip the rounds after the first score fewer items (and exactly as many
    where it is off: weights above 1 or per-patch weights)."""
    from patchtable_amd import synth_pair
    a, _ = synth_pair(61, 70, 110)
    _, b = synth_pair(62, 80, 120)
    ah, aw = a.shape[0] - 7, a.shape[1] - 7
    p = speed1("-limit 0.1 " + switches + ("" if prev is None else " -coherence_temporal 1.5"
                                            if "coherence_temporal" not in switches else ""))
    o = O.OracleTable(p.raw, b)
    prev_f = sv = None
    if prev is not None:
        prev_f = o.lookup_parallel(a[:, ::-1].copy())   # a poor previous field
        if prev == "sv":
            sv = np.random.default_rng(5).uniform(0.5, 2.0, (ah, aw)).astype(np.float32)
    out = {}
    for on in (0, 1):
        O.set_mfma_seen_skip(on)
        try:
            out[on] = (o.lookup_parallel(a, prev=prev_f, sv=sv), O.mfma_items())
        finally:
            O.set_mfma_seen_skip(1)
    assert np.array_equal(out[0][0], out[1][0])
    full, skip = out[0][1], out[1][1]
    assert skip[0] == full[0]
    if prev is None or ("coherence_temporal" not in switches and prev != "sv"):
        assert all(s <= f for s, f in zip(skip, full)) and sum(skip) < sum(full)
    else:
        assert skip == full
