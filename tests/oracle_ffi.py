"""ctypes binding of the CPU oracle (oracle/liboracle.so) for the tests.

The oracle is test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's CPU
baseline use it, to check the HIP library against.
"""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "liboracle.so")
# the OpenMP build (pt_params.threads > 1 runs the raster lookup's rows and the k-coherence
# patches on several threads): timing only, selected by use_openmp() before the first call
LIB_OMP = os.path.join(ROOT, "oracle", "liboracle_omp.so")

_P = ctypes.POINTER
_f32p = _P(ctypes.c_float)
_f64p = _P(ctypes.c_double)
_i32p = _P(ctypes.c_int32)
_int = ctypes.c_int

_lib = None
_path = LIB


def use_openmp():
    """Load the OpenMP build instead of the checker (before the first oracle call)."""
    global _path
    if _lib is not None and _path != LIB_OMP:
        raise RuntimeError("the oracle library is already loaded")
    _path = LIB_OMP


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_path):
            subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
        from patchtable_amd._ffi import PtParams
        L = ctypes.CDLL(_path)
        sig = {
            "orc_glibc_rand": (_int, [ctypes.c_uint32, _int, _i32p]),
            "orc_random_shuffle": (_int, [_int, _i32p]),
            "orc_rgb2yuv": (_int, [_f32p, _int, _int, _f32p]),
            "orc_wh_basis": (_int, [_int, _P(_int), _P(_int)]),
            "orc_snake_order": (_int, [_int, _P(_int), _P(_int)]),
            "orc_gck_recurrences": (_int, [_int, _P(_int), _P(_int), _P(_int)]),
            "orc_gck": (_int, [_f32p, _int, _int, _int, _int, _f32p]),
            "orc_descriptor": (_int, [_P(PtParams), _f32p, _int, _int, _f32p]),
            "orc_patch_ssd": (ctypes.c_float, [_f32p, _int, _int, _int, _f32p, _int, _int, _int]),
            "orc_scale_slices_to_limit": (_int, [_int, _P(_int), ctypes.c_double]),
            "orc_partition": (_int, [_f32p, _int, _int, ctypes.c_float, ctypes.c_float, _f32p, _i32p]),
            "orc_build": (ctypes.c_void_p, [_P(PtParams), _f32p, _int, _int, _int, _i32p]),
            "orc_shape": (_int, [ctypes.c_void_p, _P(_int), _P(_int), _P(_int)]),
            "orc_free": (None, [ctypes.c_void_p]),
            "orc_set_params": (_int, [ctypes.c_void_p, _P(PtParams)]),
            "orc_info": (_int, [ctypes.c_void_p, _P(_int), _P(_int), _P(ctypes.c_int64), _P(_int), _P(_int)]),
            "orc_get_descriptor": (_int, [ctypes.c_void_p, _f32p]),
            "orc_get_partition": (_int, [ctypes.c_void_p, _int, _f32p, _f32p, _f32p, _i32p]),
            "orc_get_table": (_int, [ctypes.c_void_p, _i32p]),
            "orc_get_populated": (_int, [ctypes.c_void_p, _i32p]),
            "orc_get_kcoherence": (_int, [ctypes.c_void_p, _i32p, _f32p]),
            "orc_extra_info": (_int, [ctypes.c_void_p, _int, _P(_int), _P(ctypes.c_int64), _P(_int)]),
            "orc_get_extra_table": (_int, [ctypes.c_void_p, _int, _i32p]),
            "orc_save": (_int, [ctypes.c_void_p, ctypes.c_char_p]),
            "orc_dt": (_int, [_int, _P(_int), _i32p, _int, _int]),
            "orc_lookup_raster": (_int, [ctypes.c_void_p, _f32p, _int, _int, _int, _f64p, _f64p, _f32p]),
            "orc_lookup_parallel": (_int, [ctypes.c_void_p, _f32p, _int, _int, _int, _f64p, _f64p, _f32p]),
            "orc_desc_dist": (ctypes.c_float, [_f32p, _f32p, _int]),
            "orc_recompute_mean_dist_desc": (ctypes.c_double, [_f32p, _int, _int, _f32p, _int, _int, _int, _f64p, _f64p]),
            "orc_mean_dist": (ctypes.c_double, [_f64p, ctypes.c_long]),
            "orc_recompute_mean_dist": (ctypes.c_double, [_f32p, _int, _int, _f32p, _int, _int, _f64p, _f64p]),
            "orc_build_keyed": (ctypes.c_void_p, [_P(PtParams), _f32p, _int, _int, _f32p]),
            "orc_lookup_parallel_keyed": (_int, [ctypes.c_void_p, _f32p, _int, _int, _f32p, _f64p]),
            "orc_lookup_raster_keyed": (_int, [ctypes.c_void_p, _f32p, _int, _int, _f32p, _f64p]),
            "orc_brute_f64": (_int, [_f32p, _int, _i32p, _int, _f32p, _int, _int, _i32p, _f64p]),
            "orc_key_round_weights": (_int, [_int]),
            "orc_mfma_items": (_int, [_P(ctypes.c_long), _int]),
            "orc_set_mfma_seen_skip": (None, [_int]),
            "orc_key_loss": (ctypes.c_double, [_f64p, _f64p, _f64p, _f64p, _f32p, _f32p, _int, _i32p, _int,
                                               ctypes.c_double]),
            "orc_key_forward": (_int, [_f32p, _f32p, _f32p, _f32p, _f32p, _int, _int, _f32p]),
            "orc_key_gradients": (_int, [_f32p, _f32p, _f32p, _f32p, _f32p, _f32p, _int, _i32p, _int, ctypes.c_float,
                                         _f32p, _f32p, _f32p, _f32p, _f32p]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def ptr(a, ctype):
    return a.ctypes.data_as(_P(ctype))


def f32(a):
    return ptr(a, ctypes.c_float)


def f64(a):
    return ptr(a, ctypes.c_double)


def i32(a):
    return ptr(a, ctypes.c_int32)


class OracleTable:
    """CPU restatement of PatchTable construction (oracle/pt_oracle.cpp `build`)."""

    def __init__(self, params, b, allowed=None, key=None):
        """key: for a learned-key table, the key of b ((h+7) x (w+7) x 8 float32)."""
        b = np.ascontiguousarray(b, dtype=np.float32)
        self.b = b
        self.params = params
        self.allowed = None if allowed is None else np.ascontiguousarray(allowed, dtype=np.int32)
        if key is not None:
            self._key = np.ascontiguousarray(key, dtype=np.float32)
            self.handle = lib().orc_build_keyed(ctypes.byref(params), f32(b), b.shape[0], b.shape[1], f32(self._key))
        else:
            self.handle = lib().orc_build(ctypes.byref(params), f32(b), b.shape[0], b.shape[1], b.shape[2],
                                          None if self.allowed is None else i32(self.allowed))
        h, w, nd = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        lib().orc_shape(self.handle, ctypes.byref(h), ctypes.byref(w), ctypes.byref(nd))
        self.h, self.w, self.nd = h.value, w.value, nd.value
        G = ctypes.c_int()
        ns = (ctypes.c_int * 64)()
        nc = ctypes.c_int64()
        add = ctypes.c_int()
        K = ctypes.c_int()
        lib().orc_info(self.handle, ctypes.byref(G), ns, ctypes.byref(nc), ctypes.byref(add), ctypes.byref(K))
        self.G = G.value
        self.nslices = [ns[i] for i in range(self.G)]
        self.ncells = nc.value
        self.table_add_count = add.value
        self.K = K.value

    def __del__(self):
        # at interpreter exit the module's globals may already be gone (lib is None); the
        # process then frees the memory anyway
        if getattr(self, "handle", None) and lib is not None:
            lib().orc_free(self.handle)
            self.handle = None

    def set_params(self, params):
        self.params = params
        lib().orc_set_params(self.handle, ctypes.byref(params))

    def extra(self, i):
        """(nslices, table_add_count, cells) of further table i."""
        ns = (ctypes.c_int * 64)()
        nc = ctypes.c_int64()
        add = ctypes.c_int()
        assert lib().orc_extra_info(self.handle, i, ns, ctypes.byref(nc), ctypes.byref(add)) == 0
        cells = np.empty(nc.value, np.int32)
        lib().orc_get_extra_table(self.handle, i, i32(cells))
        return [ns[k] for k in range(self.G)], add.value, cells

    def save(self, path):
        assert lib().orc_save(self.handle, str(path).encode()) == 0

    def descriptor(self):
        nd = self.nd
        out = np.empty((self.h + 7, self.w + 7, nd), np.float32)
        lib().orc_get_descriptor(self.handle, f32(out))
        return out

    def partition(self, dim):
        mn, mx = ctypes.c_float(), ctypes.c_float()
        centers = np.empty(self.nslices[dim], np.float32)
        to_bin = np.empty(16384, np.int32)
        lib().orc_get_partition(self.handle, dim, ctypes.byref(mn), ctypes.byref(mx), f32(centers), i32(to_bin))
        return mn.value, mx.value, centers, to_bin

    def table(self):
        out = np.empty(self.ncells, np.int32)
        lib().orc_get_table(self.handle, i32(out))
        return out

    def populated(self):
        out = np.empty(self.ncells, np.int32)
        lib().orc_get_populated(self.handle, i32(out))
        return out

    def kcoherence(self):
        n = (self.h - 7) * (self.w - 7) * self.K
        pos = np.empty(n, np.int32)
        dist = np.empty(n, np.float32)
        lib().orc_get_kcoherence(self.handle, i32(pos), f32(dist))
        return pos.reshape(self.h - 7, self.w - 7, self.K), dist.reshape(self.h - 7, self.w - 7, self.K)

    def _ann_shape(self, a):
        if self.params.is_descriptor:
            pad = 14 if self.params.descriptor_padded else 0
            return (a.shape[0] - pad, a.shape[1] - pad, 3)
        return (a.shape[0] - 7, a.shape[1] - 7, 3)

    def _lookup(self, fn, a, prev, sv):
        a = np.ascontiguousarray(a, dtype=np.float32)
        ann = np.empty(self._ann_shape(a), np.float64)
        pv = None if prev is None else np.ascontiguousarray(prev, dtype=np.float64)
        s = None if sv is None else np.ascontiguousarray(sv, dtype=np.float32)
        fn(self.handle, f32(a), a.shape[0], a.shape[1], a.shape[2], f64(ann),
           None if pv is None else f64(pv), None if s is None else f32(s))
        return ann

    def lookup_raster(self, a, prev=None, sv=None):
        return self._lookup(lib().orc_lookup_raster, a, prev, sv)

    def lookup_parallel(self, a, prev=None, sv=None):
        return self._lookup(lib().orc_lookup_parallel, a, prev, sv)

    def lookup_keyed(self, a, qkey, raster=False):
        """Learned-key table: lookup of RGB a with its key qkey ((h+7) x (w+7) x 8)."""
        a = np.ascontiguousarray(a, dtype=np.float32)
        q = np.ascontiguousarray(qkey, dtype=np.float32)
        ann = np.empty((a.shape[0] - 7, a.shape[1] - 7, 3), np.float64)
        fn = lib().orc_lookup_raster_keyed if raster else lib().orc_lookup_parallel_keyed
        fn(self.handle, f32(a), a.shape[0], a.shape[1], f32(q), f64(ann))
        return ann


def key_forward(w1, b1, w2, b2, img):
    """Key planes 8 x (h+7) x (w+7) of an RGB image (double accumulation over BF16-rounded
    weights and inputs)."""
    img = np.ascontiguousarray(img, dtype=np.float32)
    h, w = img.shape[:2]
    out = np.zeros((8, h + 7, w + 7), np.float32)
    ws = [np.ascontiguousarray(x, dtype=np.float32) for x in (w1, b1, w2, b2)]
    lib().orc_key_forward(*[f32(x) for x in ws], f32(img), h, w, f32(out))
    return out


def brute_f64(a, b, qpos):
    """(positions XY_TO_INT, FP64 distances) of the exact nearest neighbours of the query patches
    at qpos (XY_TO_INT, patch coordinates)."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    b = np.ascontiguousarray(b, dtype=np.float32)
    q = np.ascontiguousarray(qpos, dtype=np.int32)
    pos = np.empty(len(q), np.int32)
    dist = np.empty(len(q), np.float64)
    lib().orc_brute_f64(f32(a), a.shape[1], i32(q), len(q), f32(b), b.shape[0], b.shape[1], i32(pos), f64(dist))
    return pos, dist


def key_loss(w1, b1, w2, b2, a, b, pos, margin):
    """Mean triplet loss in double for double weights (no BF16 rounding)."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    b = np.ascontiguousarray(b, dtype=np.float32)
    pos = np.ascontiguousarray(pos, dtype=np.int32)
    ws = [np.ascontiguousarray(x, dtype=np.float64) for x in (w1, b1, w2, b2)]
    return lib().orc_key_loss(*[f64(x) for x in ws], f32(a), f32(b), a.shape[1], i32(pos), pos.shape[0],
                              float(margin))


def key_round_weights(on):
    lib().orc_key_round_weights(1 if on else 0)


def key_gradients(w1, b1, w2, b2, a, b, pos, margin):
    """(gW1, gb1, gW2, gb2), loss of the mean triplet loss; pos: [nt, 3] XY_TO_INT int32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    b = np.ascontiguousarray(b, dtype=np.float32)
    pos = np.ascontiguousarray(pos, dtype=np.int32)
    ws = [np.ascontiguousarray(x, dtype=np.float32) for x in (w1, b1, w2, b2)]
    g = [np.empty_like(x) for x in ws]
    loss = np.zeros(1, np.float32)
    lib().orc_key_gradients(*[f32(x) for x in ws], f32(a), f32(b), a.shape[1], i32(pos), pos.shape[0], float(margin),
                            *[f32(x) for x in g], f32(loss))
    return g, float(loss[0])


def mean_dist(ann):
    ann = np.ascontiguousarray(ann, dtype=np.float64)
    return lib().orc_mean_dist(f64(ann), ann.shape[0] * ann.shape[1])


def recompute(a, b, ann):
    """(mean sqrt SSD recomputed in float64, per-patch float64 SSD)."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    b = np.ascontiguousarray(b, dtype=np.float32)
    ann = np.ascontiguousarray(ann, dtype=np.float64)
    d = np.empty(ann.shape[:2], np.float64)
    m = lib().orc_recompute_mean_dist(f32(a), a.shape[0], a.shape[1], f32(b), b.shape[0], b.shape[1], f64(ann), f64(d))
    return m, d


def recompute_desc(a, b, ann):
    """Descriptor mode: (mean sqrt of the float64 squared descriptor distance, per-patch values)."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    b = np.ascontiguousarray(b, dtype=np.float32)
    ann = np.ascontiguousarray(ann, dtype=np.float64)
    d = np.empty(ann.shape[:2], np.float64)
    m = lib().orc_recompute_mean_dist_desc(f32(a), a.shape[0], a.shape[1], f32(b), b.shape[0], b.shape[1], a.shape[2],
                                           f64(ann), f64(d))
    return m, d


def dt(nslices, table, algo, dt_iters=-1):
    t = np.ascontiguousarray(table, dtype=np.int32).copy()
    ns = (ctypes.c_int * len(nslices))(*nslices)
    rc = lib().orc_dt(len(nslices), ns, i32(t), algo, dt_iters)
    assert rc == 0
    return t


def set_mfma_seen_skip(on):
    """1 (default): the MFMA model skips anchors a group searched in its last round with anchors;
    0: it searches every anchor (the full search the skip is checked against)."""
    lib().orc_set_mfma_seen_skip(int(on))


def mfma_items():
    """(group, anchor) items of each round of the oracle's last MFMA lookup."""
    buf = (ctypes.c_long * 64)()
    n = lib().orc_mfma_items(buf, 64)
    return [int(buf[i]) for i in range(min(n, 64))]
