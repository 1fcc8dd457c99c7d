"""GPU parity: every stage of the HIP path against the CPU oracle on the same seeded inputs.

Bit-exact for GCK, exact SSD, quantiser, populate, distance transforms and k-coherence lists;
the lookup is bit-exact against the oracle's model of the parallel algorithm and within 1% mean
SSD of the oracle's restatement of the reference's raster-order lookup.
"""
import ctypes

import numpy as np
import pytest

import oracle_ffi as O
from conftest import seeded, speed1

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def dev(x):
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


def vp(t):
    return ctypes.c_void_p(t.data_ptr())


def oracle_gck(img, n, nchroma, convert=True):
    h, w, _ = img.shape
    src = img
    if convert:
        src = np.empty_like(img)
        O.lib().orc_rgb2yuv(O.f32(img), h, w, O.f32(src))
    out = np.empty((h + 7, w + 7, n + 2 * nchroma), np.float32)
    O.lib().orc_gck(O.f32(np.ascontiguousarray(src)), h, w, n, nchroma, O.f32(out))
    return out


@pytest.mark.parametrize("h,w,n,nchroma", [(300, 400, 4, 1), (37, 53, 6, 1), (64, 40, 8, 0), (20, 30, 2, 2),
                                           (9, 8, 10, 1), (90, 70, 4, 3), (1030, 2100, 4, 1), (130, 257, 10, 1)])
def test_gck_bit_exact(lib, h, w, n, nchroma):
    print(f"gck {h}x{w} n={n} nchroma={nchroma}", flush=True)
    img = seeded((h, w, 3), 100 + h)
    C = n + 2 * nchroma
    out = torch.empty((C, h + 7, w + 7), dtype=torch.float32, device="cuda")
    rc = lib.pt_gck_dev(vp(dev(img)), h, w, n, nchroma, 1, vp(out))
    assert rc == 0
    got = out.cpu().numpy().transpose(1, 2, 0)
    ref = oracle_gck(img, n, nchroma)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


def test_patch_ssd_bit_exact(lib):
    rng = np.random.default_rng(5)
    a = seeded((40, 50, 3), 1)
    b = seeded((30, 60, 3), 2)
    n = 500
    ax = rng.integers(0, 50 - 7, n).astype(np.int32)
    ay = rng.integers(0, 40 - 7, n).astype(np.int32)
    bx = rng.integers(0, 60 - 7, n).astype(np.int32)
    by = rng.integers(0, 30 - 7, n).astype(np.int32)
    out = torch.empty(n, dtype=torch.float32, device="cuda")
    da, db = dev(a), dev(b)
    tx, ty, ux, uy = dev(ax), dev(ay), dev(bx), dev(by)
    assert lib.pt_patch_ssd_dev(vp(da), 50, vp(db), 60, vp(tx), vp(ty), vp(ux), vp(uy), n, vp(out)) == 0
    got = out.cpu().numpy()
    ref = np.array([O.lib().orc_patch_ssd(O.f32(a), 50, int(ax[i]), int(ay[i]), O.f32(b), 60, int(bx[i]), int(by[i]))
                    for i in range(n)], np.float32)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    # and within float rounding of a float64 recompute
    d64 = np.array([((a[ay[i]:ay[i] + 8, ax[i]:ax[i] + 8].astype(np.float64) -
                      b[by[i]:by[i] + 8, bx[i]:bx[i] + 8]) ** 2).sum() for i in range(n)])
    assert np.allclose(got, d64, rtol=1e-5, atol=0)


def _build_both(p, b):
    from patchtable_amd import PatchTable
    t = PatchTable(p, b)
    o = O.OracleTable(p.raw, b)
    return t, o


@pytest.mark.parametrize("extra", ["-limit 1", "-limit 1 -dt_algo downsample", "-limit 1 -dt_algo euclidean",
                                   "-limit 0.5 -dt_iters 3", "-limit 1 -run_dt 0", "-limit 2 -partition_step 3",
                                   "-limit 1 -populate_random 0",
                                   # DT line passes of every shape: single-slice dimensions (the first
                                   # dimension's pass then has stride 1), lines longer than 64, a
                                   # contiguous dimension longer than 256, and a distance cap
                                   "-nslices 40,1,1,1,1,1", "-nslices 100,3,3,3,2,2",
                                   "-nslices 3,3,3,3,2,300", "-nslices 50,4,4,4,4,4 -dt_iters 2"])
def test_table_build_bit_exact(pair_cfg1, extra):
    print(f"table build cfg1 {extra}", flush=True)
    a, b = pair_cfg1
    p = speed1(extra)
    t, o = _build_both(p, b)
    info = t.info()
    assert info["nslices"] == o.nslices
    assert info["ncells"] == o.ncells
    assert info["table_add_count"] == o.table_add_count
    assert np.array_equal(t.descriptor().view(np.uint32), o.descriptor().view(np.uint32))
    for d in range(len(o.nslices)):
        gm, gx, gc, gb = t.partition(d)
        om, ox, oc, ob = o.partition(d)
        assert (gm, gx) == (om, ox)
        assert np.array_equal(gc.view(np.uint32), oc.view(np.uint32))
        assert np.array_equal(gb, ob)
    assert np.array_equal(t.populated(), o.populated())
    assert np.array_equal(t.cells(), o.table())
    gp, gd = t.kcoherence()
    op, od = o.kcoherence()
    assert np.array_equal(gp, op)
    assert np.array_equal(gd.view(np.uint32), od.view(np.uint32))


@pytest.mark.parametrize("extra", ["-nslices 40,40,40,2,2,1", "-nslices 30,30,30,30,1,1",
                                   "-nslices 40,40,40,2,2,1 -dt_iters 10", "-nslices 40,40,40,2,2,1 -dt_iters 40"])
def test_table_build_sparse_far_cells(extra):
    """A 24x24 database (289 patches) in a lattice of about 1e8 cells: cells up to 54-58 steps
    (L1, in cells) from the nearest filled cell, beyond what a 32-bit DT key holds (30), so the
    fill takes the 64-bit passes after the 32-bit ones; with a distance cap of 11 the 32-bit
    passes finish alone, with a cap of 41 they do not."""
    from patchtable_amd import synth_pair
    print(f"table build sparse {extra}", flush=True)
    _, b = synth_pair(5, 24, 24)
    t, o = _build_both(speed1(extra), b)
    assert t.info()["nslices"] == o.nslices
    assert np.array_equal(t.populated(), o.populated())
    assert np.array_equal(t.cells(), o.table())


def test_table_build_speed1_full_size_cfg1(pair_cfg1):
    """The default speed-1 table (limit 1e8 cells) at 400x300, all stages bit-exact."""
    a, b = pair_cfg1
    p = speed1()
    t, o = _build_both(p, b)
    assert t.info()["nslices"] == o.nslices
    assert np.array_equal(t.cells(), o.table())
    gp, gd = t.kcoherence()
    op, od = o.kcoherence()
    assert np.array_equal(gp, op) and np.array_equal(gd.view(np.uint32), od.view(np.uint32))


@pytest.mark.parametrize("rounds", [0, 1, 4, 8])
def test_lookup_bit_exact_vs_parallel_model(pair_cfg1, rounds):
    a, b = pair_cfg1
    p = speed1(f"-limit 1 -jf_rounds {rounds}")
    t, o = _build_both(p, b)
    ann = t.lookup(a)
    ref = o.lookup_parallel(a)
    assert np.array_equal(ann, ref)


def test_lookup_quality_vs_raster_reference(pair_cfg1):
    """Mean SSD (and mean sqrt SSD) within 1% of the raster-order restatement at speed 1;
    every NNF_DIST equals the recomputed SSD of its match; every match in bounds."""
    a, b = pair_cfg1
    p = speed1()
    t, o = _build_both(p, b)
    ann = t.lookup(a)
    ref = o.lookup_raster(a)
    bw, bh = b.shape[1] - 7, b.shape[0] - 7
    assert ann.shape == ref.shape == (a.shape[0] - 7, a.shape[1] - 7, 3)
    assert (ann[..., 0] >= 0).all() and (ann[..., 0] < bw).all()
    assert (ann[..., 1] >= 0).all() and (ann[..., 1] < bh).all()
    m, d64 = O.recompute(a, b, ann)
    assert np.abs(d64 - ann[..., 2]).max() <= 1e-4
    _assert_within_1pct(ann, ref, "cfg1 400x300")


def _assert_within_1pct(ann, ref, what):
    """Mean SSD and mean sqrt SSD (the reference's mean_dist, patchtable_main.cpp:90-101) within
    1% of the raster-order restatement, both ways."""
    r_ssd = ann[..., 2].mean() / ref[..., 2].mean()
    r_dist = O.mean_dist(ann) / O.mean_dist(ref)
    print(f"{what}: mean SSD gpu {ann[..., 2].mean():.6f} raster {ref[..., 2].mean():.6f} (ratio {r_ssd:.5f}); "
          f"mean_dist gpu {O.mean_dist(ann):.6f} raster {O.mean_dist(ref):.6f} (ratio {r_dist:.5f})", flush=True)
    assert 0.99 <= r_ssd <= 1.01
    assert 0.99 <= r_dist <= 1.01


def test_lookup_bit_exact_default_limit_cfg1(pair_cfg1):
    """The default speed-1 table (limit 1e8 cells) at 400x300: lookup bit-exact vs the model."""
    a, b = pair_cfg1
    p = speed1()
    t, o = _build_both(p, b)
    assert np.array_equal(t.lookup(a), o.lookup_parallel(a))


def test_cfg2_1024x768_bit_exact_and_quality():
    """The 1024x768 generator pair (seed 2), PatchTableParams() + set_speed(1), default limit:
    descriptor, slices, cells and k-coherence lists bit-exact; lookup bit-exact vs the oracle's
    parallel model; mean SSD and mean sqrt SSD within 1% of the raster-order restatement."""
    from patchtable_amd import synth_pair
    print("cfg2 1024x768 seed 2: generating", flush=True)
    a, b = synth_pair(2, 768, 1024)
    p = speed1()
    print("cfg2: building GPU and oracle tables", flush=True)
    t, o = _build_both(p, b)
    assert t.info()["nslices"] == o.nslices
    assert t.info()["table_add_count"] == o.table_add_count
    assert np.array_equal(t.descriptor().view(np.uint32), o.descriptor().view(np.uint32))
    assert np.array_equal(t.cells(), o.table())
    gp, gd = t.kcoherence()
    op, od = o.kcoherence()
    assert np.array_equal(gp, op) and np.array_equal(gd.view(np.uint32), od.view(np.uint32))
    print("cfg2: lookups", flush=True)
    ann = t.lookup(a)
    assert np.array_equal(ann, o.lookup_parallel(a))
    m, d64 = O.recompute(a, b, ann)
    assert np.abs(d64 - ann[..., 2]).max() <= 1e-4
    _assert_within_1pct(ann, o.lookup_raster(a), "cfg2 1024x768")


# Shapes whose tile grids are 8 or 16 columns wide, so the kernels take the XCD column-band map
# (tile_coords): database 200x512 (505 patch columns = 8 tiles of 64), queries 200x512, 100x1031
# (1024 columns = 16 tiles) and 128x1031 with query_step 2 (512 lattice columns = 8 tiles; the
# gather runs over 16).
@pytest.mark.parametrize("qh,qw,extra", [(200, 512, ""), (100, 1031, ""), (128, 1031, "-query_step 2"),
                                         (200, 512, "-kcoherence_iter -1")])
def test_xcd_band_grids_bit_exact(qh, qw, extra):
    from patchtable_amd import synth_pair
    print(f"xcd bands: database 200x512, query {qh}x{qw} {extra}", flush=True)
    _, b = synth_pair(41, 200, 512)
    a, _ = synth_pair(42, qh, qw)
    p = speed1("-limit 1 " + extra)
    t, o = _build_both(p, b)
    assert np.array_equal(t.cells(), o.table())
    gp, gd = t.kcoherence()
    op, od = o.kcoherence()
    assert np.array_equal(gp, op) and np.array_equal(gd.view(np.uint32), od.view(np.uint32))
    assert np.array_equal(t.lookup(a), o.lookup_parallel(a))
