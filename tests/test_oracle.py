"""Pins the CPU oracle against the reference's golden vectors and against what the mathematics
fixes (closed forms, brute force on small inputs, invariants). CPU only."""
import json
import os

import numpy as np
import pytest

import oracle_ffi as O
from conftest import seeded, speed1

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
UNUSED = (1 << 30) - 1


def fnv1a(a):
    h = 1469598103934665603
    b = np.asarray(a, dtype="<u4").tobytes()
    for byte in b:
        h ^= byte
        h = (h * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return h


def fnv1a_fast(a):
    # vectorised over 4-byte words is not possible for FNV; use the C-compiled reference order in
    # chunks via numpy ufunc reduce on uint64 with wraparound.
    h = np.uint64(1469598103934665603)
    p = np.uint64(1099511628211)
    b = np.frombuffer(np.asarray(a, dtype="<u4").tobytes(), dtype=np.uint8)
    with np.errstate(over="ignore"):
        for byte in b:
            h = np.uint64((h ^ np.uint64(byte)) * p)
    return int(h)


# ---------------------------------------------------------------- glibc rand / random_shuffle
@pytest.fixture(scope="module")
def glibc():
    return json.load(open(os.path.join(GOLDEN, "glibc_rand.json")))


def test_glibc_rand_stream_matches_system_rand(glibc):
    n = 4200000
    r = np.empty(n, np.int32)
    O.lib().orc_glibc_rand(1, n, O.i32(r))
    assert r[:2000].tolist() == glibc["rand_first"]
    idx = np.array(glibc["rand_sparse_index"])
    assert r[idx].tolist() == glibc["rand_sparse_value"]


@pytest.mark.parametrize("k", [0, 1, 2, 3, 4])
def test_random_shuffle_matches_libstdcxx(glibc, k):
    s = glibc["shuffle"][k]
    n = s["n"]
    a = np.arange(n, dtype=np.int32)
    O.lib().orc_random_shuffle(n, O.i32(a))
    m = len(s["head"])
    assert a[:m].tolist() == s["head"]
    assert a[n - m:].tolist() == s["tail"]
    assert np.array_equal(np.sort(a), np.arange(n))
    if n <= 1000:
        assert str(fnv1a(a)) == s["fnv1a"]


# ---------------------------------------------------------------- colour conversion
def test_rgb2yuv_closed_form():
    img = seeded((31, 17, 3), 3)
    out = np.empty_like(img)
    O.lib().orc_rgb2yuv(O.f32(img), 31, 17, O.f32(out))
    R, G, B = (img[..., c].astype(np.float64) for c in range(3))
    Y = (0.299 * R + 0.587 * G) + 0.114 * B
    U = (-0.14713 * R + -0.28886 * G) + 0.436 * B
    V = (0.615 * R + -0.51499 * G) + -0.10001 * B
    ref = np.stack([Y, U, V], -1).astype(np.float32)
    assert np.array_equal(out, ref)


# ---------------------------------------------------------------- Walsh-Hadamard / GCK
def test_wh_basis_known_answer():
    """gck.py:70-77, the reference's only known-answer test."""
    vec = (O.ctypes.c_int * 64)()
    alp = (O.ctypes.c_int * 24)()
    O.lib().orc_wh_basis(8, vec, alp)
    v = np.array(vec[:]).reshape(8, 8).tolist()
    a = np.array(alp[:]).reshape(8, 3).tolist()
    assert v == [[1, 1, 1, 1, 1, 1, 1, 1], [1, 1, 1, 1, -1, -1, -1, -1],
                 [1, 1, -1, -1, -1, -1, 1, 1], [1, 1, -1, -1, 1, 1, -1, -1],
                 [1, -1, -1, 1, 1, -1, -1, 1], [1, -1, -1, 1, -1, 1, 1, -1],
                 [1, -1, 1, -1, -1, 1, -1, 1], [1, -1, 1, -1, 1, -1, 1, -1]]
    assert a == [[1, 1, 1], [1, 1, -1], [1, -1, -1], [1, -1, 1],
                 [-1, -1, 1], [-1, -1, -1], [-1, 1, -1], [-1, 1, 1]]
    vec4 = (O.ctypes.c_int * 16)()
    alp4 = (O.ctypes.c_int * 8)()
    O.lib().orc_wh_basis(4, vec4, alp4)
    assert np.array(vec4[:]).reshape(4, 4).tolist() == [[1, 1, 1, 1], [1, 1, -1, -1], [1, -1, -1, 1], [1, -1, 1, -1]]
    assert np.array(alp4[:]).reshape(4, 2).tolist() == [[1, 1], [1, -1], [-1, -1], [-1, 1]]


def test_gck_recurrences_match_generated_header():
    """Axis, offset and sign of kernels 2..64 against the table read from the reference's gck.h."""
    fx = json.load(open(os.path.join(GOLDEN, "gck_recurrences.json")))["kernels"]
    ax, de, sg = ((O.ctypes.c_int * 64)() for _ in range(3))
    n = O.lib().orc_gck_recurrences(8, ax, de, sg)
    assert n == 64
    for row in fx:
        k = row["kernel"] - 1
        assert (ax[k], de[k], sg[k]) == (row["axis"], row["delta"], row["sign"]), row


def _snake():
    r, c = (O.ctypes.c_int * 64)(), (O.ctypes.c_int * 64)()
    O.lib().orc_snake_order(8, r, c)
    vec = (O.ctypes.c_int * 64)()
    alp = (O.ctypes.c_int * 24)()
    O.lib().orc_wh_basis(8, vec, alp)
    return list(zip(r[:], c[:])), np.array(vec[:], np.float64).reshape(8, 8)


@pytest.mark.parametrize("n,nchroma", [(4, 1), (10, 1), (6, 0), (3, 3)])
def test_gck_equals_brute_force_correlation(n, nchroma):
    """Each output channel is the 64-tap correlation of the 8x8 block ending at (y, x) with a
    2-D Walsh-Hadamard kernel outer(wh[row], wh[col]) in snake order, up to one sign per kernel;
    checked in float64 on a small image (zero outside the image)."""
    h, w = 23, 29
    img = seeded((h, w, 3), 7) - 0.5
    out = np.empty((h + 7, w + 7, n + 2 * nchroma), np.float32)
    O.lib().orc_gck(O.f32(np.ascontiguousarray(img)), h, w, n, nchroma, O.f32(out))
    order, wh = _snake()
    pad = np.zeros((h + 14, w + 14, 3))
    pad[7:7 + h, 7:7 + w] = img
    for k in range(n):
        r, c = order[k]
        K = np.outer(wh[r], wh[c])
        chans = range(3) if k < nchroma else [0]
        base = 3 * k if k < nchroma else 3 * nchroma + (k - nchroma)
        for j, col in enumerate(chans):
            ref = np.zeros((h + 7, w + 7))
            for dy in range(8):
                for dx in range(8):
                    ref += K[dy, dx] * pad[dy:dy + h + 7, dx:dx + w + 7, col]
            got = out[..., base + j].astype(np.float64)
            sign = np.sign(np.sum(got * ref))
            assert sign != 0
            assert np.allclose(got, sign * ref, atol=2e-4), (k, col)


# ---------------------------------------------------------------- patch SSD
def test_patch_ssd_vs_float64():
    a = seeded((20, 30, 3), 1)
    b = seeded((25, 18, 3), 2)
    rng = np.random.default_rng(0)
    for _ in range(200):
        ax, ay = int(rng.integers(0, 23)), int(rng.integers(0, 13))
        bx, by = int(rng.integers(0, 11)), int(rng.integers(0, 18))
        d = O.lib().orc_patch_ssd(O.f32(a), 30, ax, ay, O.f32(b), 18, bx, by)
        ref = ((a[ay:ay + 8, ax:ax + 8].astype(np.float64) - b[by:by + 8, bx:bx + 8]) ** 2).sum()
        assert abs(d - ref) <= 1e-5 * ref


# ---------------------------------------------------------------- quantiser
@pytest.mark.parametrize("ns,limit", [([31, 26, 36, 17, 9, 22], 1e8), ([200, 150, 90, 60, 30, 10], 1e8),
                                      ([200, 1, 1, 0, 0, 5], 1e6), ([120, 80, 60], 1e5), ([5, 5], 1e8)])
def test_scale_slices_to_limit_invariants(ns, limit):
    arr = (O.ctypes.c_int * len(ns))(*ns)
    O.lib().orc_scale_slices_to_limit(len(ns), arr, limit)
    out = list(arr)
    prod = float(np.prod(np.array(out, dtype=np.float64)))
    assert prod <= limit and min(out) >= 1
    for i in range(len(out)):   # greedily maximal: no single +1 fits
        bumped = list(out)
        bumped[i] += 1
        assert float(np.prod(np.array(bumped, dtype=np.float64))) > limit
    if float(np.prod(np.maximum(np.array(ns, np.float64), 1))) <= limit:
        assert all(o >= max(v, 1) for o, v in zip(out, ns))


def test_partition_centres_and_bins_brute_force():
    rng = np.random.default_rng(4)
    vals = (rng.standard_normal(20001) ** 3).astype(np.float32)
    ns = 37
    mn, mx = float(vals.min()), float(vals.max())
    centers = np.empty(ns, np.float32)
    to_bin = np.empty(16384, np.int32)
    O.lib().orc_partition(O.f32(vals), len(vals), ns, mn, mx, O.f32(centers), O.i32(to_bin))
    s = np.sort(vals)
    n = len(vals)
    ref = np.array([(s[(n - 1) * i // ns] + s[(n - 1) * (i + 1) // ns]) / np.float32(2) for i in range(ns)], np.float32)
    assert np.array_equal(centers, ref)
    rng_ = np.float32(mx) - np.float32(mn)
    coord = mn + (np.arange(16384) + 0.5) * float(rng_) / 16384
    dist = np.abs(coord[:, None] - centers[None, :].astype(np.float64))
    nearest = dist.argmin(1)
    ok = to_bin == nearest
    # disagreement only where the coordinate sits on a midpoint (within float rounding)
    srt = np.sort(dist, 1)
    assert np.all(ok | (srt[:, 1] - srt[:, 0] < 1e-6))


# ---------------------------------------------------------------- distance transforms
def _brute_prop(ns, t, max_dist=None):
    """Nearest filled cell in L1 over the lattice, ties to the smallest linear index."""
    coords = np.array(np.unravel_index(np.arange(t.size), ns)).T
    src = np.nonzero(t != UNUSED)[0]
    out = t.copy()
    d = np.abs(coords[:, None, :] - coords[src][None, :, :]).sum(-1)
    best = np.argmin(d * (len(src) + 1) + np.arange(len(src))[None, :], axis=1)
    bd = d[np.arange(t.size), best]
    out = t[src[best]]
    if max_dist is not None:
        out = np.where(bd > max_dist, UNUSED, out)
    return out


def test_dt_prop_known_answers():
    U = UNUSED
    # 1-D: a tie at equal distance goes to the source of smaller index
    t = np.array([5, U, U, U, 9], np.int32)
    assert O.dt([5], t, 0).tolist() == [5, 5, 5, 9, 9]
    # 2-D 3x3, sources at (0,0)=1 and (2,2)=2: anti-diagonal ties go to (0,0)
    t = np.full(9, U, np.int32)
    t[0], t[8] = 1, 2
    assert O.dt([3, 3], t, 0).reshape(3, 3).tolist() == [[1, 1, 1], [1, 1, 2], [1, 2, 2]]
    # dt_iters 1: the BFS stops after 2 levels
    t = np.array([7, U, U, U, U, U], np.int32)
    assert O.dt([6], t, 0, dt_iters=1).tolist() == [7, 7, 7, U, U, U]


def test_dt_prop_bfs_equals_nearest_source_brute_force():
    rng = np.random.default_rng(1)
    for trial in range(400):
        G = int(rng.integers(1, 5))
        ns = [int(rng.integers(1, 7)) for _ in range(G)]
        n = int(np.prod(ns))
        t = np.full(n, UNUSED, np.int32)
        k = int(rng.integers(1, max(2, n // 3) + 1))
        idx = rng.choice(n, size=min(k, n), replace=False)
        t[idx] = rng.integers(0, 1 << 20, size=len(idx))
        it = int(rng.choice([-1, -1, 1, 2, 3]))
        bfs = O.dt(ns, t, 0, it)
        ref = _brute_prop(ns, t, max_dist=(it + 1) if it > 0 else None)
        assert np.array_equal(bfs, ref), (ns, it)
        assert np.array_equal(O.dt(ns, t, 1, it), bfs)


def test_dt_downsample_known_answers():
    U = UNUSED
    t = np.array([3, U, U, U, 9], np.int32)
    assert O.dt([5], t, 2).tolist() == [3, 3, 3, 3, 9]
    t = np.array([U, U, 7, U, U], np.int32)
    assert O.dt([5], t, 2).tolist() == [7] * 5
    # 2-D 4x4 with sources at (0,3)=4 and (3,0)=5: coarse cells take the first filled child
    t = np.full(16, U, np.int32)
    t[3], t[12] = 4, 5
    out = O.dt([4, 4], t, 2).reshape(4, 4)
    assert out.tolist() == [[4, 4, 4, 4], [4, 4, 4, 4], [5, 5, 4, 4], [5, 5, 4, 4]]


def _euclid_case(rng, G):
    ns = [int(rng.integers(2, 8)) for _ in range(G)]
    n = int(np.prod(ns))
    t = np.full(n, UNUSED, np.int32)
    k = int(rng.integers(1, max(2, n // 4) + 1))
    idx = rng.choice(n, size=min(k, n), replace=False)
    vals = np.arange(len(idx), dtype=np.int32) + 100
    t[idx] = vals
    out = O.dt(ns, t, 3)
    coords = np.array(np.unravel_index(np.arange(n), ns)).T
    pos_of = {int(v): coords[i] for v, i in zip(vals, idx)}
    d2 = ((coords[:, None, :] - coords[idx][None, :, :]) ** 2).sum(-1).min(1)
    got = np.array([((coords[c] - pos_of[int(out[c])]) ** 2).sum() for c in range(n)])
    return got, d2


def test_dt_euclidean_1d_is_nearest_source():
    """Along one dimension the lower envelope of Felzenszwalb-Huttenlocher gives a nearest
    source everywhere (the in-place index update of patchtable.h:4268 cannot read a stale entry:
    a source's own parabola is its envelope at its own position)."""
    rng = np.random.default_rng(2)
    for trial in range(300):
        got, d2 = _euclid_case(rng, 1)
        assert np.array_equal(got, d2), trial


def test_dt_euclidean_nd_mostly_nearest():
    """In several dimensions the reference updates index_array in place while still reading it
    (idx[q] = idx[v[k]], patchtable.h:4268), so a cell whose envelope parabola v < q was itself
    re-assigned inherits that re-assignment. The oracle keeps this behaviour; the chosen source is
    always populated and nearest for the large majority of cells."""
    rng = np.random.default_rng(3)
    tot = ok = 0
    for trial in range(200):
        got, d2 = _euclid_case(rng, int(rng.integers(2, 4)))
        assert (got >= d2).all()
        tot += got.size
        ok += int((got == d2).sum())
    assert ok / tot > 0.9


# ---------------------------------------------------------------- populate
def test_populate_is_min_rank_per_cell(pair_cfg1):
    a, b = pair_cfg1
    p = speed1("-limit 1")
    T = O.OracleTable(p.raw, b)
    pop = T.populated()
    wh = T.descriptor()
    h, w = b.shape[:2]
    n = (h - 7) * (w - 7)
    perm = np.arange(n, dtype=np.int32)
    O.lib().orc_random_shuffle(n, O.i32(perm))
    rank = np.empty(n, np.int64)
    rank[perm] = np.arange(n)
    cells = np.zeros(n, np.int64)
    stride = np.cumprod([1] + T.nslices[::-1])[:-1][::-1]
    ys, xs = np.divmod(np.arange(n), w - 7)
    for d in range(T.G):
        mn, mx, centers, to_bin = T.partition(d)
        v = wh[ys + 7, xs + 7, d]
        f = (v - np.float32(mn)) * np.float32(16383) / (np.float32(mx) - np.float32(mn))
        idx = np.clip(f.astype(np.int64), 0, 16383)
        cells += to_bin[idx].astype(np.int64) * stride[d]
    best = {}
    for k in np.argsort(rank):
        best.setdefault(int(cells[k]), k)
    ref = np.full(T.ncells, UNUSED, np.int64)
    for c, k in best.items():
        ref[c] = (xs[k] + 7) | ((ys[k] + 7) << 16)
    assert np.array_equal(pop, ref)
    assert T.table_add_count == len(best)
    # after the distance transform every cell holds a populated patch
    tab = T.table()
    assert not (tab == UNUSED).any()
    assert set(np.unique(tab).tolist()) <= set(np.unique(pop[pop != UNUSED]).tolist())


# ---------------------------------------------------------------- lookup invariants
def test_raster_lookup_invariants(pair_cfg1):
    """stress.py:134-154: matches in bounds, NNF_DIST equal to the recomputed SSD (1e-4),
    mean_dist equal to its float64 recompute."""
    a, b = pair_cfg1
    p = speed1("-limit 1")
    T = O.OracleTable(p.raw, b)
    ann = T.lookup_raster(a)
    assert ((ann[..., 0] >= 0) & (ann[..., 0] < b.shape[1] - 7)).all()
    assert ((ann[..., 1] >= 0) & (ann[..., 1] < b.shape[0] - 7)).all()
    m, d = O.recompute(a, b, ann)
    assert np.abs(d - ann[..., 2]).max() <= 1e-4
    assert abs(m - O.mean_dist(ann)) <= 1e-4
    par = T.lookup_parallel(a)
    m2, d2 = O.recompute(a, b, par)
    assert np.abs(d2 - par[..., 2]).max() <= 1e-4


def test_exact_nn_lower_bound_small():
    """Both lookups are bounded below by the exact nearest neighbour (brute force, float64) on a
    small pair, and stay close to it."""
    from patchtable_amd import synth_pair
    a, b = synth_pair(21, 40, 48)
    p = speed1("-limit 0.2")
    T = O.OracleTable(p.raw, b)
    ann = T.lookup_raster(a)
    par = T.lookup_parallel(a)
    ah, aw = a.shape[0] - 7, a.shape[1] - 7
    bh, bw = b.shape[0] - 7, b.shape[1] - 7
    A = np.lib.stride_tricks.sliding_window_view(a.astype(np.float64), (8, 8, 3))[:, :, 0].reshape(ah * aw, -1)
    B = np.lib.stride_tricks.sliding_window_view(b.astype(np.float64), (8, 8, 3))[:, :, 0].reshape(bh * bw, -1)
    d = (A ** 2).sum(1)[:, None] + (B ** 2).sum(1)[None, :] - 2 * A @ B.T
    exact = np.maximum(d.min(1), 0).reshape(ah, aw)
    assert (ann[..., 2] >= exact - 1e-4).all() and (par[..., 2] >= exact - 1e-4).all()
    assert par[..., 2].mean() <= 3.0 * exact.mean() + 1e-3
