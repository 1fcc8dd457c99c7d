"""GPU parity of the lookup and table build beyond the default path: every speed preset
(query_step > 1, random search, k-coherence in every pass), allowed-patch masks, several tables,
a previous field with temporal weighting, and descriptor mode.

Bit-exact: the table build (cells, partitions, k-coherence lists, further tables) against the
oracle's restatement of the reference, and the lookup against the oracle's model of the parallel
lookup. Quality: mean distance within 1% of the oracle's raster-order restatement of the reference
(patchtable.h:3468-4217) at the same parameters, and the acceptance test of test_descriptor
(patchtable_main.cpp:174-240) for descriptor mode.
"""
import numpy as np
import pytest

import oracle_ffi as O

pytestmark = pytest.mark.gpu


def params(switches, **fields):
    from patchtable_amd import PatchTableParams
    p = PatchTableParams()
    if switches:
        p.set_from_switches("-x 0 " + switches)
    for k, v in fields.items():
        setattr(p, k, v)
    return p


@pytest.fixture(scope="module")
def pair_small():
    from patchtable_amd import synth_pair
    return synth_pair(6, 160, 200)


def mask_for(b, seed=0):
    """Allowed-patch mask of b: a disallowed disc and a strip with another label."""
    bh, bw = b.shape[0] - 7, b.shape[1] - 7
    yy, xx = np.mgrid[0:bh, 0:bw]
    m = np.ones((bh, bw), np.int32)
    m[(yy - bh // 2) ** 2 + (xx - bw // 3) ** 2 < (min(bh, bw) // 4) ** 2] = 0
    m[:, int(bw * 0.75):] = 2
    return m


def build_pair(p, b, allowed=None):
    from patchtable_amd import PatchTable
    return PatchTable(p, b, allowed), O.OracleTable(p.raw, b, allowed)


def assert_tables_equal(t, o):
    info = t.info()
    assert info["nslices"] == o.nslices and info["table_add_count"] == o.table_add_count
    assert np.array_equal(t.cells(), o.table())
    for d in range(len(o.nslices)):
        gm, gx, gc, gb = t.partition(d)
        om, ox, oc, ob = o.partition(d)
        assert (gm, gx) == (om, ox)
        assert np.array_equal(gc.view(np.uint32), oc.view(np.uint32)) and np.array_equal(gb, ob)
    if o.K:
        gp, gd = t.kcoherence()
        op, od = o.kcoherence()
        assert np.array_equal(gp, op) and np.array_equal(gd.view(np.uint32), od.view(np.uint32))


@pytest.mark.parametrize("speed", range(10))
def test_speed_presets_bit_exact_and_quality(pair_small, speed):
    a, b = pair_small
    print(f"speed {speed}", flush=True)
    p = params(f"-speed {speed} -limit 1")
    t, o = build_pair(p, b)
    assert_tables_equal(t, o)
    ann = t.lookup(a)
    assert np.array_equal(ann, o.lookup_parallel(a))
    ref = o.lookup_raster(a)
    print(f"speed {speed}: mean_dist gpu {O.mean_dist(ann):.5f} raster {O.mean_dist(ref):.5f}")
    assert O.mean_dist(ann) <= 1.01 * O.mean_dist(ref)
    m, d64 = O.recompute(a, b, ann)
    assert np.abs(d64 - ann[..., 2]).max() <= 1e-4


@pytest.mark.parametrize("switches", ["-speed 1 -limit 1", "-speed 3 -limit 1", "-speed 5 -limit 1",
                                      "-speed 1 -limit 1 -partition_step 2 -allowed_index 2"])
def test_allowed_mask(pair_small, switches):
    a, b = pair_small
    print(f"allowed {switches}", flush=True)
    p = params(switches)
    m = mask_for(b)
    t, o = build_pair(p, b, m)
    assert_tables_equal(t, o)
    assert np.array_equal(t.populated(), o.populated())
    ann = t.lookup(a)
    assert np.array_equal(ann, o.lookup_parallel(a))
    idx = p.allowed_index
    assert (m[ann[..., 1].astype(int), ann[..., 0].astype(int)] == idx).all()
    ref = o.lookup_raster(a)
    assert O.mean_dist(ann) <= 1.01 * O.mean_dist(ref)


@pytest.mark.parametrize("switches", ["-speed 1 -limit 2 -ntables 3", "-speed 3 -limit 1 -ntables 4 -table_ratio 1.5"])
def test_multiple_tables(pair_small, switches):
    a, b = pair_small
    print(f"ntables {switches}", flush=True)
    p = params(switches)
    t, o = build_pair(p, b)
    assert_tables_equal(t, o)
    for i in range(p.ntables - 1):
        ns, add, cells = o.extra(i)
        e = t.extra_info(i)
        assert e["nslices"] == ns and e["table_add_count"] == add
        assert np.array_equal(t.extra_cells(i), cells)
    ann = t.lookup(a)
    assert np.array_equal(ann, o.lookup_parallel(a))
    assert O.mean_dist(ann) <= 1.01 * O.mean_dist(o.lookup_raster(a))


def shifted(img, dy, dx):
    """The previous video frame: img moved by (dy, dx) pixels, edges replicated."""
    h, w = img.shape[:2]
    return np.ascontiguousarray(np.pad(img, ((dy, 0), (dx, 0), (0, 0)), mode="edge")[:h, :w])


@pytest.mark.parametrize("switches", ["-speed 1 -limit 1 -coherence_temporal 1",
                                      "-speed 1 -limit 1 -coherence_temporal 0.5 -recalc_dist_temporal 0",
                                      "-speed 3 -limit 1 -coherence_temporal 2",
                                      "-speed 1 -limit 1 -do_table_lookup 0 -coherence_temporal 1",
                                      "-speed 5 -limit 1 -coherence_temporal 1"])
@pytest.mark.parametrize("with_sv", [False, True])
def test_previous_field(pair_small, switches, with_sv):
    """The field of the previous frame (the query shifted by (2, 3)) seeds the lookup: bit-exact
    against the model and within 1% of the raster restatement. Rows 5..8 of the previous field
    point outside b and are clamped (sanitize_input)."""
    a, b = pair_small
    print(f"prev {switches} sv={with_sv}", flush=True)
    p0 = params("-speed 1 -limit 1")
    t0, _ = build_pair(p0, b)
    prev = t0.lookup(shifted(a, 2, 3))
    prev[5:9, :, 0] = -3
    p = params(switches)
    t, o = build_pair(p, b)
    sv = None
    if with_sv:
        sv = np.random.default_rng(1).uniform(0.3, 1.0, prev.shape[:2]).astype(np.float32)
    ann = t.lookup(a, ann_prev=prev, coherence_temporal_sv=sv)
    ref = o.lookup_parallel(a, prev, sv)
    assert np.array_equal(ann, ref)
    rast = o.lookup_raster(a, prev, sv)
    print(f"prev {switches}: gpu {O.mean_dist(ann):.5f} raster {O.mean_dist(rast):.5f}")
    assert O.mean_dist(ann) <= 1.01 * O.mean_dist(rast)


def test_previous_field_unrelated_is_bit_exact(pair_small):
    """A previous field from an unrelated image, without table candidates: still bit-exact."""
    from patchtable_amd import synth_pair
    a, b = pair_small
    a2, _ = synth_pair(7, 160, 200)
    t0, _ = build_pair(params("-speed 1 -limit 1"), b)
    prev = t0.lookup(a2)
    p = params("-speed 1 -limit 1 -do_table_lookup 0 -coherence_temporal 1")
    t, o = build_pair(p, b)
    assert np.array_equal(t.lookup(a, ann_prev=prev), o.lookup_parallel(a, prev))


def test_previous_field_out_of_range_without_sanitize(pair_small):
    from patchtable_amd import PatchTable, PatchTableError
    a, b = pair_small
    p = params("-speed 1 -limit 1 -sanitize_input 0")
    t = PatchTable(p, b)
    prev = t.lookup(a)
    prev[0, 0, 0] = 10 ** 6
    with pytest.raises(PatchTableError, match="in bounds"):
        t.lookup(a, ann_prev=prev)


def random_descriptors(h, w, k, seed):
    rng = np.random.default_rng(seed)
    return rng.random((h, w, k), dtype=np.float32), rng.random((h, w, k), dtype=np.float32)


def test_descriptor_mode_test_descriptor():
    """test_descriptor (patchtable_main.cpp:174-240): 400x300 random 8-channel descriptors,
    default parameters; grid_ndims 8, then 6, then 6 of 6 channels; the mean distance over the
    full 8 channels must grow in that order and start below 0.77. Each stage is also bit-exact
    against the oracle's model and within 1% of its raster-order restatement."""
    from patchtable_amd import PatchTable, PatchTableParams
    h, w, k = 300, 400, 8
    a0, b0 = random_descriptors(h, w, k, 0)
    dL = []
    for reduce in range(3):
        p = PatchTableParams()
        p.is_descriptor = 1
        A, B = a0, b0
        if reduce >= 1:
            p.grid_ndims = k - 2
        if reduce == 2:
            p.ndims = k - 2
            A, B = np.ascontiguousarray(a0[..., :k - 2]), np.ascontiguousarray(b0[..., :k - 2])
        t = PatchTable(p, B)
        ann = t.lookup(A)
        assert ann.shape == (h, w, 3)
        assert p.ndims == A.shape[2] and p.nchroma == 0   # written back (patchtable.h:4398-4401)
        o = O.OracleTable(p.raw, B)
        assert_tables_equal(t, o)
        assert np.array_equal(ann, o.lookup_parallel(A))
        m, d = O.recompute_desc(A, B, ann)
        assert np.abs(d - ann[..., 2]).max() <= 1e-4
        assert O.mean_dist(ann) <= 1.01 * O.mean_dist(o.lookup_raster(A))
        full, _ = O.recompute_desc(a0, b0, ann)
        dL.append(full)
        print(f"test_descriptor reduce={reduce}: mean_dist {full:.5f}")
    assert dL[0] < dL[1] < dL[2] and dL[0] < 0.77


@pytest.mark.parametrize("k,switches", [(5, "-speed 1"), (12, "-speed 3"), (16, "-speed 5"), (12, "-speed 1 -descriptor_padded 1")])
def test_descriptor_mode_other_widths(k, switches):
    from patchtable_amd import PatchTable
    print(f"descriptor k={k} {switches}", flush=True)
    a, b = random_descriptors(90, 120, k, k)
    p = params(switches, is_descriptor=1)
    t = PatchTable(p, b)
    o = O.OracleTable(p.raw, b)
    assert_tables_equal(t, o)
    ann = t.lookup(a)
    assert np.array_equal(ann, o.lookup_parallel(a))


def test_descriptor_mode_with_mask_and_tables():
    from patchtable_amd import PatchTable
    a, b = random_descriptors(100, 130, 8, 3)
    m = np.ones((100, 130), np.int32)
    m[20:60, 30:90] = 0
    p = params("-speed 1 -ntables 2", is_descriptor=1)
    t = PatchTable(p, b, m)
    o = O.OracleTable(p.raw, b, m)
    assert_tables_equal(t, o)
    ann = t.lookup(a)
    assert np.array_equal(ann, o.lookup_parallel(a))
    assert (m[ann[..., 1].astype(int), ann[..., 0].astype(int)] == 1).all()


def test_quantiles_zeros_of_both_signs():
    """A grid dimension whose quantile ranks fall inside a run of -0.0 and +0.0. The radix sort
    the order statistics follow gives both zeros the same digit and keeps input order, so the
    value at such a rank is the zero at that position; the build's radix select cannot see it and
    sorts that dimension instead. The slice centres must match a stable sort by the same keys bit
    for bit (the sign of a zero centre included); the other dimensions, the oracle's."""
    from patchtable_amd import PatchTable
    rng = np.random.default_rng(5)
    h, w, k = 90, 120, 6
    b = rng.random((h, w, k), dtype=np.float32) - np.float32(0.5)
    zero = rng.random((h, w)) < 0.6
    neg = rng.random((h, w)) < 0.5
    c0 = b[..., 0]
    c0[zero & neg] = np.float32(-0.0)
    c0[zero & ~neg] = np.float32(0.0)
    p = params("-speed 1 -limit 1", is_descriptor=1)
    t = PatchTable(p, b)
    o = O.OracleTable(p.raw, b)
    ns = t.info()["nslices"]
    assert ns == o.nslices and ns[0] + 1 <= 48   # dimension 0 goes through the select (48 slots)
    H, W = t.b_shape[0], t.b_shape[1]
    vals = np.ascontiguousarray(t.descriptor()[7:H, 7:W, 0]).ravel()
    u = vals.view(np.uint32).copy()
    u[u == 0x80000000] = 0
    key = np.where(u & 0x80000000, ~u, u | np.uint32(0x80000000))
    srt = vals[np.argsort(key, kind="stable")]
    n, m = len(vals), ns[0]
    q = np.array([srt[(n - 1) * j // m] for j in range(m + 1)], np.float32)
    want = ((q[:-1] + q[1:]) / np.float32(2)).astype(np.float32)
    _, _, gc, _ = t.partition(0)
    assert np.array_equal(gc.view(np.uint32), want.view(np.uint32))
    zc = want[want == 0]
    assert len(zc) >= 2 and np.signbit(zc).any() and (~np.signbit(zc)).any()
    _, _, oc, _ = o.partition(0)
    assert np.array_equal(gc, oc)   # equal as numbers: std::sort orders the zeros arbitrarily
    for d in range(1, len(ns)):
        gm, gx, gcd, gb = t.partition(d)
        om, ox, ocd, ob = o.partition(d)
        assert (gm, gx) == (om, ox)
        assert np.array_equal(gcd.view(np.uint32), ocd.view(np.uint32)) and np.array_equal(gb, ob)


def test_wide_database_build():
    """A database wider than 16,383 patches (17,993 here): the quantile select counts in 16-bit LDS
    halves, so its blocks take fewer rows of such a grid. Table and lookup must still match the
    oracle bit for bit."""
    from patchtable_amd import PatchTable
    rng = np.random.default_rng(11)
    b = rng.random((24, 18000, 3), dtype=np.float32)
    a = rng.random((24, 700, 3), dtype=np.float32)
    p = params("-speed 1 -limit 1")
    print("wide database 24x18000", flush=True)
    t, o = build_pair(p, b)
    assert_tables_equal(t, o)
    assert np.array_equal(t.lookup(a), o.lookup_parallel(a))


@pytest.mark.parametrize("bshape,ashape,switches", [
    ((8, 8), (8, 8), "-speed 0"),                    # one database patch, one query patch
    ((8, 8), (40, 50), "-speed 1 -limit 1"),         # every query patch has the one candidate
    ((9, 9), (8, 8), "-speed 1"),                    # 4 patches; dims of 61 and 45 slices
    ((8, 300), (30, 20), "-speed 0"),                # database of one patch row
    ((300, 8), (20, 30), "-speed 1 -limit 1"),       # database of one patch column
    ((60, 60), (8, 300), "-speed 9 -limit 1"),       # one query row, query_step 10
    ((12, 15), (300, 8), "-speed 5 -limit 1"),       # one query column, query_step 2
    ((9, 9), (8, 8), "-speed 9 -limit 1"),
])
def test_smallest_and_thinnest_images(bshape, ashape, switches):
    """Images at the 8x8 minimum and one patch thick: the table (cells, partitions, k-coherence
    lists shorter than kcoherence) and the lookup equal the oracle's bit for bit, and every
    NNF_DIST is the recomputed SSD of its in-bounds match."""
    from patchtable_amd import synth_pair
    print(f"database {bshape}, query {ashape}, {switches}", flush=True)
    a, _ = synth_pair(11, *ashape)
    _, b = synth_pair(12, *bshape)
    p = params(switches)
    t, o = build_pair(p, b)
    assert_tables_equal(t, o)
    ann = t.lookup(a)
    assert np.array_equal(ann, o.lookup_parallel(a))
    xs, ys = ann[..., 0].astype(int), ann[..., 1].astype(int)
    assert (xs >= 0).all() and (ys >= 0).all() and (xs < bshape[1] - 7).all() and (ys < bshape[0] - 7).all()
    _, d = O.recompute(a, b, ann)
    assert np.abs(d - ann[..., 2]).max() <= 1e-4


@pytest.mark.parametrize("desc", [False, True])
@pytest.mark.parametrize("switches", ["-speed 1 -limit 1", "-speed 0 -limit 1"])
def test_kcoherence_ties_on_a_repeated_database(desc, switches):
    """A database made of one tile repeated 3 x 3 times: every patch has exact copies, so its
    k-coherence candidates tie in descriptor distance and in exact distance (0 for each copy, 6
    to 8 such entries per list). The lists must equal the oracle's, which orders equal distances
    by raster index when selecting and by XY_TO_INT when sorting."""
    from patchtable_amd import synth_pair
    rng = np.random.default_rng(5)
    if desc:
        b = np.tile(rng.random((20, 20, 6), dtype=np.float32), (3, 3, 1))
        a = rng.random((50, 70, 6), dtype=np.float32)
        p = params(switches, is_descriptor=1)
    else:
        _, tile = synth_pair(21, 30, 30)
        b = np.tile(tile, (3, 3, 1))
        a, _ = synth_pair(22, 60, 80)
        p = params(switches)
    print(f"repeated database {b.shape}, desc={desc}, {switches}", flush=True)
    t, o = build_pair(p, b)
    assert_tables_equal(t, o)
    assert np.array_equal(t.lookup(a), o.lookup_parallel(a))


def test_descriptor_distance_lane_order(lib):
    """pt_desc_dist (descriptor-mode distance) equals the oracle's restatement bit for bit for the
    SSE-chunked widths and the scalar-tail widths."""
    import ctypes
    import torch
    rng = np.random.default_rng(9)
    for C in (4, 5, 7, 8, 12, 16, 20, 24, 40):
        n = 300
        a = rng.normal(size=(n, C)).astype(np.float32) * 3
        b = rng.normal(size=(n, C)).astype(np.float32) * 3
        ref = np.array([O.lib().orc_desc_dist(O.f32(a[i].copy()), O.f32(b[i].copy()), C) for i in range(n)], np.float32)
        da, db = torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda()
        out = torch.empty(n, dtype=torch.float32, device="cuda")
        assert lib.pt_desc_dist_dev(ctypes.c_void_p(da.data_ptr()), ctypes.c_void_p(db.data_ptr()), C, n,
                                    ctypes.c_void_p(out.data_ptr())) == 0
        assert np.array_equal(out.cpu().numpy().view(np.uint32), ref.view(np.uint32)), C
        d64 = ((a.astype(np.float64) - b) ** 2).sum(1)
        assert np.allclose(ref, d64, rtol=1e-5)
