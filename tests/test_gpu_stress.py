"""GPU counterpart of the reference's stress.py (stress.py:62-154): seeded random cases mixing
image sizes, speed presets, descriptor mode, a previous field with temporal weighting, allowed
masks and several tables. Each case checks stress.py's invariants -- every match in bounds,
NNF_DIST equal to the recomputed distance of its match (|d| <= 1e-4), no match on a disallowed
patch -- and, beyond them, that the table and the lookup equal the oracle's bit for bit."""
import numpy as np
import pytest

import oracle_ffi as O
from test_gpu_modes import assert_tables_equal, params, random_descriptors, shifted

pytestmark = pytest.mark.gpu

NCASE = 160
# Further seeds whose k-coherence lists meet an exact distance tie in a way the first NCASE do not:
# a later candidate with the same distance and a smaller index than a listed one. A compiled form
# of the selection's tie test kept a duplicate there instead of the smaller index.
TIE_SEEDS = [913, 1578, 1915, 3090, 3105, 3120, 3662]


def case(seed):
    rng = np.random.default_rng(1000 + seed)
    desc = rng.random() < 0.25
    speed = int(rng.integers(0, 10))
    sw = f"-speed {speed} -limit {rng.choice([0.2, 1, 5])}"
    if rng.random() < 0.3:
        sw += f" -ntables {int(rng.integers(2, 4))}"
    use_prev = rng.random() < 0.35
    if use_prev and rng.random() < 0.6:
        sw += f" -coherence_temporal {rng.choice([0.5, 1, 2])}"
    use_mask = rng.random() < 0.35
    h, w = int(rng.integers(40, 170)), int(rng.integers(40, 170))
    return dict(rng=rng, desc=desc, sw=sw, use_prev=use_prev, use_mask=use_mask, h=h, w=w)


def random_mask(rng, bh, bw):
    m = np.ones((bh, bw), np.int32)
    for _ in range(int(rng.integers(1, 4))):
        y0, x0 = int(rng.integers(0, bh)), int(rng.integers(0, bw))
        m[y0:y0 + int(rng.integers(4, bh // 2 + 5)), x0:x0 + int(rng.integers(4, bw // 2 + 5))] = 0
    m[0, 0] = 1   # never empty
    return m


@pytest.mark.parametrize("seed", list(range(NCASE)) + TIE_SEEDS)
def test_stress_case(seed):
    from patchtable_amd import PatchTable, synth_pair
    c = case(seed)
    rng, h, w = c["rng"], c["h"], c["w"]
    print(f"stress {seed}: desc={c['desc']} {c['sw']} prev={c['use_prev']} mask={c['use_mask']} {h}x{w}", flush=True)
    if c["desc"]:
        C = int(rng.choice([4, 6, 8, 12]))
        a, b = random_descriptors(h, w, C, int(rng.integers(0, 1000)))
        p = params(c["sw"], is_descriptor=1)
        bh, bw = h, w
    else:
        a, _ = synth_pair(2000 + seed, h + int(rng.integers(-10, 11)), w + int(rng.integers(-10, 11)))
        _, b = synth_pair(3000 + seed, h, w)
        p = params(c["sw"])
        bh, bw = h - 7, w - 7
    mask = random_mask(rng, bh, bw) if c["use_mask"] else None
    t = PatchTable(p, b, mask)
    o = O.OracleTable(p.raw, b, mask)
    assert_tables_equal(t, o)
    prev = None
    if c["use_prev"]:
        # the previous frame (the query moved by (1, 2)) looked up under the same mask, as
        # stress.py makes its prev_nnf
        prev = t.lookup(shifted(a, 1, 2))
    ann = t.lookup(a, ann_prev=prev)
    assert np.array_equal(ann, o.lookup_parallel(a, prev))
    xs, ys = ann[..., 0].astype(int), ann[..., 1].astype(int)
    assert (xs >= 0).all() and (ys >= 0).all() and (xs < bw).all() and (ys < bh).all()
    if c["desc"]:
        _, d = O.recompute_desc(a, b, ann)
    else:
        _, d = O.recompute(a, b, ann)
    tsc = 1.0 / (1 + p.coherence_temporal) if prev is not None else 1.0
    if prev is None or p.coherence_temporal == 0:
        assert np.abs(d - ann[..., 2]).max() <= 1e-4
    else:
        # the temporal rescaling (patchtable.h:4178-4193) multiplies the distances that came from
        # the previous field; every distance is one of the two
        ok = (np.abs(d - ann[..., 2]) <= 1e-4) | (np.abs(d * tsc - ann[..., 2]) <= 1e-4)
        assert ok.all()
    if mask is not None:
        assert (mask[ys, xs] == p.allowed_index).all()
