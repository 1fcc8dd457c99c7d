"""MI355X-native PatchTable: Python host mirror of the reference's C++ API.

    from patchtable_amd import PatchTableParams, PatchTable
    p = PatchTableParams(); p.set_speed(1)
    table = PatchTable(p, b)              # b: H x W x 3 float32 in [0,1]
    ann = table.lookup(a)                 # (Ha-7) x (Wa-7) x 3 float64: NNF_X, NNF_Y, NNF_DIST

Mirrors PatchTableParams (patchtable/patchtable_params.h:23-161, patchtable.cpp:26-387) and
PatchTable<> (patchtable/patchtable.h:4359-4491). All work runs in the HIP library
libpatchtable_hip.so through its C ABI (include/patchtable_amd.h); there is no CPU fallback.
"""
import ctypes

import numpy as np

from . import _ffi
from ._ffi import PatchTableError, PtParams, check, load

NNF_X = 0
NNF_Y = 1
NNF_DIST = 2

__all__ = ["PatchTableParams", "PatchTable", "PatchTableBatch", "PatchKey", "PatchTableError", "NNF_X",
           "NNF_Y", "NNF_DIST", "synth_pair", "load_color_image", "save_color_image", "brute_lookup",
           "brute_exact_f64"]


class PatchTableParams:
    """Parameters with the reference's defaults, speed presets and switches.

    PatchTableParams()            -> set_defaults() (patchtable.cpp:375-377)
    PatchTableParams(argv)        -> defaults then set_from_switches(argc, argv); argv[0] is the
                                     program name and is skipped as in parse_switches
    p.set_speed(i)                -> patchtable.cpp:138-150
    p.set_from_switches("...")    -> patchtable.cpp:152-161 (first token dropped, like the
                                     reference)
    Attributes read and write the pt_params fields by their reference names.
    """

    def __init__(self, argv=None):
        object.__setattr__(self, "_p", PtParams())
        check(load().pt_params_default(ctypes.byref(self._p)))
        if argv is not None:
            self.set_from_switches(" ".join(str(a) for a in argv))

    def set_speed(self, i):
        check(load().pt_params_set_speed(ctypes.byref(self._p), int(i)))

    def set_from_switches(self, s):
        check(load().pt_params_set_switches(ctypes.byref(self._p), s.encode()))

    def __getattr__(self, name):
        p = object.__getattribute__(self, "_p")
        v = getattr(p, name)
        if name in ("nslices", "filter_dims"):
            n = p.nslices_count if name == "nslices" else p.filter_dims_count
            return [v[i] for i in range(n)]
        if isinstance(v, bytes):
            return v.decode()
        return v

    def __setattr__(self, name, value):
        p = self._p
        if name in ("nslices", "filter_dims"):
            arr = getattr(p, name)
            for i, x in enumerate(value):
                arr[i] = int(x)
            setattr(p, name + "_count", len(value))
            if name == "nslices":
                p.ndims = len(value)
            return
        if name in ("load_filename", "save_filename"):
            value = value.encode()
        setattr(p, name, value)

    @property
    def raw(self):
        return self._p


class PatchTable:
    """PatchTable<float, float, int32_t> (patchtable.h:4359-4491) backed by HIP kernels.

    b may be a numpy array (host) or a torch CUDA tensor (device, HWC float32 contiguous).
    """

    def __init__(self, p, b, allowed_patches=None):
        self.p = p
        self._t = ctypes.c_void_p()
        allowed = None
        if allowed_patches is not None:
            self._allowed = np.ascontiguousarray(allowed_patches, dtype=np.int32)
            allowed = self._allowed.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
        if _is_device_tensor(b):
            import torch
            if b.dim() != 3:
                raise PatchTableError("b must be h x w x c")
            _check_dev(b, "b", torch.float32)
            h, w, c = (int(s) for s in b.shape)
            _after_torch(b.device)
            check(load().pt_table_create_dev(ctypes.byref(p.raw), ctypes.c_void_p(b.data_ptr()),
                                             h, w, c, allowed, ctypes.byref(self._t)))
        else:
            b = np.ascontiguousarray(b, dtype=np.float32)
            h, w, c = b.shape
            check(load().pt_table_create(ctypes.byref(p.raw), b.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                         h, w, c, allowed, ctypes.byref(self._t)))
        self._shape()

    def _shape(self):
        """Table "image" size (b, or the descriptor size + 7 in descriptor mode) and dims."""
        h, w, nd = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        check(load().pt_table_shape(self._t, ctypes.byref(h), ctypes.byref(w), ctypes.byref(nd)))
        self.b_shape = (h.value, w.value, 3)
        self.ndims = nd.value

    def ann_shape(self, qh, qw):
        """NNF shape for a query of qh x qw: (qh-7, qw-7) images; (qh, qw) descriptors
        ((qh-14, qw-14) when descriptor_padded)."""
        if self.p.is_descriptor:
            pad = 14 if self.p.descriptor_padded else 0
            return (qh - pad, qw - pad, 3)
        return (qh - 7, qw - 7, 3)

    @classmethod
    def keyed(cls, p, key, b):
        """Learned-key table of the device image b (torch CUDA tensor, h x w x 3): the 8-dim key
        of `key` (a PatchKey) indexes the table; the GCK table of b with p is its fallback."""
        if not _is_device_tensor(b):
            raise PatchTableError("PatchTable.keyed expects a torch CUDA tensor")
        import torch
        if b.dim() != 3 or int(b.shape[2]) != 3:
            raise PatchTableError("b must be h x w x 3")
        _check_dev(b, "b", torch.float32)
        self = cls.__new__(cls)
        self.p = p
        self._t = ctypes.c_void_p()
        h, w, _ = (int(s) for s in b.shape)
        _after_torch(b.device)
        check(load().pt_table_create_keyed(ctypes.byref(p.raw), key._k, ctypes.c_void_p(b.data_ptr()), h, w,
                                           ctypes.byref(self._t)))
        self._shape()
        return self

    @classmethod
    def load(cls, p, filename):
        self = cls.__new__(cls)
        self.p = p
        self._t = ctypes.c_void_p()
        check(load().pt_table_load(ctypes.byref(p.raw), filename.encode(), ctypes.byref(self._t)))
        self._shape()
        return self

    def close(self):
        if getattr(self, "_t", None) and self._t.value:
            load().pt_table_destroy(self._t)
            self._t = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def lookup(self, a, ann=None, ann_prev=None, coherence_temporal_sv=None):
        """Returns the NNF; also stores the lookup's wall time in self.last_seconds.

        Host arrays in, host array out; with a torch CUDA tensor `a`, the field is returned as a
        torch CUDA float64 tensor (and `ann`, if given, must be one of the right shape)."""
        secs = ctypes.c_double()
        if _is_device_tensor(a):
            import torch
            if a.dim() != 3:
                raise PatchTableError("a must be h x w x c")
            _check_dev(a, "a", torch.float32)
            h, w, c = (int(s) for s in a.shape)
            shp = self.ann_shape(h, w)
            if ann is None:
                ann = torch.empty(shp, dtype=torch.float64, device=a.device)
            _check_dev(ann, "ann", torch.float64, shp, a.device)
            if ann_prev is not None:
                _check_dev(ann_prev, "ann_prev", torch.float64, shp, a.device)
            if coherence_temporal_sv is not None:
                _check_dev(coherence_temporal_sv, "coherence_temporal_sv", torch.float32, shp[:2], a.device)
            _after_torch(a.device)
            prev = ctypes.c_void_p(ann_prev.data_ptr()) if ann_prev is not None else None
            sv = ctypes.c_void_p(coherence_temporal_sv.data_ptr()) if coherence_temporal_sv is not None else None
            check(load().pt_table_lookup_dev(self._t, ctypes.c_void_p(a.data_ptr()), h, w, c,
                                             ctypes.c_void_p(ann.data_ptr()), prev, sv, ctypes.byref(secs)))
        else:
            a = np.ascontiguousarray(a, dtype=np.float32)
            h, w, c = a.shape
            if ann is None:
                ann = np.empty(self.ann_shape(h, w), np.float64)
            prev = None
            if ann_prev is not None:
                self._prev = np.ascontiguousarray(ann_prev, dtype=np.float64)
                prev = self._prev.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
            sv = None
            if coherence_temporal_sv is not None:
                self._sv = np.ascontiguousarray(coherence_temporal_sv, dtype=np.float32)
                sv = self._sv.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
            check(load().pt_table_lookup(self._t, a.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), h, w, c,
                                         ann.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), prev, sv,
                                         ctypes.byref(secs)))
        self.last_seconds = secs.value
        return ann

    def save(self, filename):
        check(load().pt_table_save(self._t, filename.encode()))

    # ---- introspection (tests) ----
    def info(self):
        G = ctypes.c_int()
        ns = (ctypes.c_int * _ffi.PT_MAX_DIMS)()
        nc = ctypes.c_int64()
        add = ctypes.c_int()
        K = ctypes.c_int()
        check(load().pt_table_info(self._t, ctypes.byref(G), ns, ctypes.byref(nc), ctypes.byref(add), ctypes.byref(K)))
        return dict(grid_ndims=G.value, nslices=[ns[i] for i in range(G.value)], ncells=nc.value,
                    table_add_count=add.value, kcoherence=K.value)

    def descriptor(self):
        h, w, _ = self.b_shape
        out = np.empty((h + 7, w + 7, self.ndims), np.float32)
        check(load().pt_table_get_descriptor(self._t, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return out

    def query_descriptor(self, h, w):
        out = np.empty((h + 7, w + 7, self.ndims), np.float32)
        check(load().pt_table_get_query_descriptor(self._t, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return out

    def partition(self, dim):
        info = self.info()
        mn, mx = ctypes.c_float(), ctypes.c_float()
        centers = np.empty(info["nslices"][dim], np.float32)
        to_bin = np.empty(16384, np.int32)
        check(load().pt_table_get_partition(self._t, dim, ctypes.byref(mn), ctypes.byref(mx),
                                            centers.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                            to_bin.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))
        return mn.value, mx.value, centers, to_bin

    def cells(self):
        out = np.empty(self.info()["ncells"], np.int32)
        check(load().pt_table_get_cells(self._t, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))
        return out

    def populated(self):
        out = np.empty(self.info()["ncells"], np.int32)
        check(load().pt_table_get_populated(self._t, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))
        return out

    def extra_info(self, i):
        """nslices, ncells, table_add_count of further table i (ntables > 1)."""
        ns = (ctypes.c_int * _ffi.PT_MAX_DIMS)()
        nc = ctypes.c_int64()
        add = ctypes.c_int()
        check(load().pt_table_extra_info(self._t, i, ns, ctypes.byref(nc), ctypes.byref(add)))
        G = self.info()["grid_ndims"]
        return dict(nslices=[ns[k] for k in range(G)], ncells=nc.value, table_add_count=add.value)

    def extra_cells(self, i):
        out = np.empty(self.extra_info(i)["ncells"], np.int32)
        check(load().pt_table_get_extra_cells(self._t, i, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))
        return out

    def kcoherence(self):
        h, w, _ = self.b_shape
        K = self.info()["kcoherence"]
        pos = np.empty((h - 7, w - 7, K), np.int32)
        dist = np.empty((h - 7, w - 7, K), np.float32)
        check(load().pt_table_get_kcoherence(self._t, pos.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                             dist.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return pos, dist

    def set_tile(self, tile_x):
        """Lattice tile of the plain lookup kernels: 64 x 4 (default), 32 x 8 or 16 x 16 patches
        (tile_x 64, 32, 16); results do not depend on it."""
        check(load().pt_table_set_tile(self._t, int(tile_x)))

    def autotune(self, a, ann, reps=5):
        """Times lookups of the device query a (into the device field ann) with each lattice tile
        and keeps the fastest; returns {"tile_x": chosen, "median_ms": {64: .., 32: .., 16: ..}}."""
        import torch
        _check_dev(a, "a", torch.float32)
        h, w = int(a.shape[0]), int(a.shape[1])
        _check_dev(ann, "ann", torch.float64, (h - 7, w - 7, 3), a.device)
        _after_torch(a.device)
        tx = ctypes.c_int()
        ms = (ctypes.c_double * 3)()
        check(load().pt_table_autotune(self._t, ctypes.c_void_p(a.data_ptr()), h, w, int(a.shape[2]),
                                       ctypes.c_void_p(ann.data_ptr()), int(reps), ctypes.byref(tx), ms))
        return {"tile_x": tx.value, "median_ms": {64: ms[0], 32: ms[1], 16: ms[2]}}

    def set_mfma_stats(self, on=True):
        """Count the MFMA rounds' items in later lookups (instrumentation; see mfma_items)."""
        check(load().pt_table_set_mfma_stats(self._t, 1 if on else 0))

    def mfma_items(self):
        """Per round of the last lookup (with set_mfma_stats on): the (16-patch group, anchor)
        items scored, each 16 query patches x 96 candidates. Empty when nothing was counted."""
        buf = (ctypes.c_int64 * 64)()
        n = ctypes.c_int()
        check(load().pt_table_mfma_items(self._t, buf, 64, ctypes.byref(n)))
        return [int(buf[i]) for i in range(min(n.value, 64))]

    def stage_ms(self):
        out = (ctypes.c_double * 12)()
        check(load().pt_table_stage_ms(self._t, out))
        names = ["gck", "quantiser", "populate", "dt", "kcoherence", "build_total",
                 "query_gck", "query_candidates", "query_propagation", "query_total",
                 "propagation_launches"]
        return dict(zip(names, list(out)))


class PatchTableBatch:
    """One query against many databases (torch CUDA tensors, HWC float32): one table per
    database, built concurrently; a lookup computes the query descriptor once and runs each phase
    as one launch over every pair, replayed from one captured HIP graph."""

    def __init__(self, p, databases, threads=4):
        import torch
        self.p = p
        if not databases:
            raise PatchTableError("PatchTableBatch needs at least one database")
        dev = databases[0].device if _is_device_tensor(databases[0]) else None
        for i, d in enumerate(databases):
            if not _is_device_tensor(d) or d.dim() != 3 or int(d.shape[2]) != 3:
                raise PatchTableError(f"database {i} must be an h x w x 3 torch CUDA tensor")
            _check_dev(d, f"database {i}", torch.float32, device=dev)
        self._dbs = list(databases)
        n = len(self._dbs)
        _after_torch(dev)
        ptrs = (ctypes.c_void_p * n)(*[d.data_ptr() for d in self._dbs])
        hs = (ctypes.c_int * n)(*[int(d.shape[0]) for d in self._dbs])
        ws = (ctypes.c_int * n)(*[int(d.shape[1]) for d in self._dbs])
        self._bt = ctypes.c_void_p()
        check(load().pt_batch_create(ctypes.byref(p.raw), n, ptrs, hs, ws, int(threads), ctypes.byref(self._bt)))
        self.n = n

    def lookup(self, a, anns=None):
        """Returns (list of n NNFs as torch CUDA float64 tensors, per-pair device ms, wall s)."""
        import torch
        if not _is_device_tensor(a) or a.dim() != 3 or int(a.shape[2]) != 3:
            raise PatchTableError("a must be an h x w x 3 torch CUDA tensor")
        _check_dev(a, "a", torch.float32)
        h, w = int(a.shape[0]), int(a.shape[1])
        if anns is None:
            anns = [torch.empty((h - 7, w - 7, 3), dtype=torch.float64, device=a.device) for _ in range(self.n)]
        if len(anns) != self.n:
            raise PatchTableError(f"expected {self.n} output fields, got {len(anns)}")
        for i, x in enumerate(anns):
            _check_dev(x, f"anns[{i}]", torch.float64, (h - 7, w - 7, 3), a.device)
        _after_torch(a.device)
        ptrs = (ctypes.c_void_p * self.n)(*[x.data_ptr() for x in anns])
        ms = (ctypes.c_double * self.n)()
        secs = ctypes.c_double()
        check(load().pt_batch_lookup(self._bt, ctypes.c_void_p(a.data_ptr()), h, w, ptrs, ms, ctypes.byref(secs)))
        return anns, list(ms), secs.value

    @property
    def tile_x(self):
        """Lattice tile width of the batch's lookup kernels (64 x 4, 32 x 8 or 16 x 16 patches)."""
        return int(load().pt_batch_get_tile(self._bt))

    @tile_x.setter
    def tile_x(self, tx):
        check(load().pt_batch_set_tile(self._bt, int(tx)))

    def autotune(self, a, anns, reps=5):
        """Times the batched lookup of query a into anns with each tile width and keeps the
        fastest; returns {"tile_x": chosen, "median_ms": {64: .., 32: .., 16: ..}}."""
        import torch
        _check_dev(a, "a", torch.float32)
        h, w = int(a.shape[0]), int(a.shape[1])
        if len(anns) != self.n:
            raise PatchTableError(f"expected {self.n} output fields, got {len(anns)}")
        for i, x in enumerate(anns):
            _check_dev(x, f"anns[{i}]", torch.float64, (h - 7, w - 7, 3), a.device)
        _after_torch(a.device)
        ptrs = (ctypes.c_void_p * self.n)(*[x.data_ptr() for x in anns])
        tx = ctypes.c_int()
        ms = (ctypes.c_double * 3)()
        check(load().pt_batch_autotune(self._bt, ctypes.c_void_p(a.data_ptr()), h, w, ptrs, int(reps),
                                       ctypes.byref(tx), ms))
        return {"tile_x": tx.value, "median_ms": {64: ms[0], 32: ms[1], 16: ms[2]}}

    def close(self):
        if getattr(self, "_bt", None) and self._bt.value:
            load().pt_batch_destroy(self._bt)
            self._bt = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PatchKey:
    """Learned patch key: a 192 -> 64 -> 8 MLP (ReLU) from an 8x8 RGB patch
    (x[dy*24 + dx*3 + c] = img(y+dy, x+dx, c) - 0.5) to an 8-dim key, BF16 weights with FP32
    accumulation, trained on the GPU from triplets mined by the library's own lookups.
    Images are torch CUDA tensors (h x w x 3 float32)."""

    KI, KH, KO = 192, 64, 8

    def __init__(self, seed=1):
        self._k = ctypes.c_void_p()
        check(load().pt_key_create(int(seed), ctypes.byref(self._k)))

    def close(self):
        if getattr(self, "_k", None) and self._k.value:
            load().pt_key_destroy(self._k)
            self._k = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_weights(self, w1, b1, w2, b2):
        arrs = [np.ascontiguousarray(x, dtype=np.float32) for x in (w1, b1, w2, b2)]
        shapes = [(self.KH, self.KI), (self.KH,), (self.KO, self.KH), (self.KO,)]
        for a, sh in zip(arrs, shapes):
            if a.shape != sh:
                raise PatchTableError(f"expected weights of shape {sh}, got {a.shape}")
        check(load().pt_key_set_weights(self._k, *[_ffi_f32(a) for a in arrs]))

    def weights(self):
        """(w1, b1, w2, b2) FP32 masters and (w1, w2) BF16 bit patterns (uint16)."""
        w1 = np.empty((self.KH, self.KI), np.float32)
        b1 = np.empty(self.KH, np.float32)
        w2 = np.empty((self.KO, self.KH), np.float32)
        b2 = np.empty(self.KO, np.float32)
        w1b = np.empty((self.KH, self.KI), np.uint16)
        w2b = np.empty((self.KO, self.KH), np.uint16)
        u16 = ctypes.POINTER(ctypes.c_uint16)
        check(load().pt_key_get_weights(self._k, _ffi_f32(w1), _ffi_f32(b1), _ffi_f32(w2), _ffi_f32(b2),
                                        w1b.ctypes.data_as(u16), w2b.ctypes.data_as(u16)))
        return (w1, b1, w2, b2), (w1b, w2b)

    def forward(self, img):
        """Key planes of a device image: torch tensor 8 x (h+7) x (w+7), patch (x, y) at
        (y+7, x+7), zeros elsewhere."""
        import torch
        if not _is_device_tensor(img) or img.dim() != 3 or int(img.shape[2]) != 3:
            raise PatchTableError("img must be an h x w x 3 torch CUDA tensor")
        _check_dev(img, "img", torch.float32)
        h, w, _ = (int(s) for s in img.shape)
        out = torch.zeros((self.KO, h + 7, w + 7), dtype=torch.float32, device=img.device)
        _after_torch(img.device)
        check(load().pt_key_forward_dev(self._k, ctypes.c_void_p(img.data_ptr()), h, w,
                                        ctypes.c_void_p(out.data_ptr())))
        return out

    def gradients(self, a, b, pos, margin):
        """Triplet-loss gradients (gW1, gb1, gW2, gb2) and loss for triplet positions pos
        (device int32 tensor [nt, 3] of XY_TO_INT: anchor in a, positive and negative in b)."""
        nt = int(pos.shape[0])
        g = [np.empty((self.KH, self.KI), np.float32), np.empty(self.KH, np.float32),
             np.empty((self.KO, self.KH), np.float32), np.empty(self.KO, np.float32)]
        loss = np.zeros(1, np.float32)
        import torch
        if tuple(a.shape) != tuple(b.shape):
            raise PatchTableError("anchor and positive/negative images must have equal sizes")
        if a.dim() != 3 or int(a.shape[2]) != 3:
            raise PatchTableError("images must be h x w x 3")
        _check_dev(a, "a", torch.float32)
        _check_dev(b, "b", torch.float32, device=a.device)
        _check_dev(pos, "pos", torch.int32, (nt, 3), a.device)
        _after_torch(a.device)
        check(load().pt_key_gradients(self._k, ctypes.c_void_p(a.data_ptr()), ctypes.c_void_p(b.data_ptr()),
                                      int(a.shape[0]), int(a.shape[1]), ctypes.c_void_p(pos.data_ptr()), nt,
                                      float(margin),
                                      *[_ffi_f32(x) for x in g], _ffi_f32(loss)))
        return g, float(loss[0])

    def train(self, p, pairs, tpp=2048, steps=300, batch=2048, lr=1e-3, margin=1.0, seed=1, positives="lookup"):
        """Mines triplets from the (a, b) device image pairs -- positives from the library's
        lookups with parameters p ("lookup") or the exhaustive search's exact nearest neighbours
        ("exact", p unused) -- then runs `steps` Adam steps; returns the per-step losses (numpy)."""
        if positives not in ("lookup", "exact"):
            raise PatchTableError("positives must be 'lookup' or 'exact'")
        n = len(pairs)
        A = (ctypes.c_void_p * n)(*[a.data_ptr() for a, _ in pairs])
        B = (ctypes.c_void_p * n)(*[b.data_ptr() for _, b in pairs])
        H = (ctypes.c_int * n)(*[int(a.shape[0]) for a, _ in pairs])
        W = (ctypes.c_int * n)(*[int(a.shape[1]) for a, _ in pairs])
        import torch
        for i, (a, b) in enumerate(pairs):
            if tuple(a.shape) != tuple(b.shape):
                raise PatchTableError("training pairs must have equal image sizes")
            if a.dim() != 3 or int(a.shape[2]) != 3:
                raise PatchTableError("training images must be h x w x 3")
            _check_dev(a, f"pairs[{i}][0]", torch.float32)
            _check_dev(b, f"pairs[{i}][1]", torch.float32, device=a.device)
        if pairs:
            _after_torch(pairs[0][0].device)
        losses = np.zeros(max(steps, 1), np.float32)
        if positives == "exact":
            check(load().pt_key_train_exact(self._k, n, A, B, H, W, int(tpp), int(steps), int(batch), float(lr),
                                            float(margin), int(seed), _ffi_f32(losses)))
        else:
            check(load().pt_key_train(self._k, ctypes.byref(p.raw), n, A, B, H, W, int(tpp), int(steps), int(batch),
                                      float(lr), float(margin), int(seed), _ffi_f32(losses)))
        return losses[:steps]


def _queries(a, b, qpos):
    import torch
    for name, x in (("a", a), ("b", b)):
        if not _is_device_tensor(x) or x.dim() != 3 or int(x.shape[2]) != 3:
            raise PatchTableError(f"{name} must be an h x w x 3 torch CUDA tensor")
        _check_dev(x, name, torch.float32, device=a.device)
    if qpos is None:
        return None, (int(a.shape[0]) - 7) * (int(a.shape[1]) - 7)
    if not (_is_device_tensor(qpos) and qpos.dtype == torch.int32):
        raise PatchTableError("query positions must be a CUDA int32 tensor of XY_TO_INT values")
    _check_dev(qpos, "qpos", torch.int32, device=a.device)
    return ctypes.c_void_p(qpos.data_ptr()), int(qpos.numel())


def brute_lookup(a, b, qpos=None, precision="bf16"):
    """Nearest neighbour of each query patch of a among all patches of b (torch CUDA tensors,
    HWC float32): MFMA candidate scoring in BF16 or FP8, exact FP32 rescoring of the 8 best.
    qpos: CUDA int32 XY_TO_INT query positions, or None for every query patch (raster order).
    Returns dict(pos, dist, cand, score, stage_ms) with pos/dist/cand/score CUDA tensors."""
    import torch
    prec = {"bf16": _ffi.PT_BRUTE_BF16, "fp8": _ffi.PT_BRUTE_FP8}[precision]
    qp, nq = _queries(a, b, qpos)
    dev = a.device
    pos = torch.empty(nq, dtype=torch.int32, device=dev)
    dist = torch.empty(nq, dtype=torch.float32, device=dev)
    cand = torch.empty((nq, 8), dtype=torch.int32, device=dev)
    score = torch.empty((nq, 8), dtype=torch.float32, device=dev)
    ms = (ctypes.c_double * 3)()
    torch.cuda.synchronize()
    check(load().pt_brute_lookup_dev(ctypes.c_void_p(a.data_ptr()), int(a.shape[0]), int(a.shape[1]),
                                     ctypes.c_void_p(b.data_ptr()), int(b.shape[0]), int(b.shape[1]), qp, nq, prec,
                                     ctypes.c_void_p(pos.data_ptr()), ctypes.c_void_p(dist.data_ptr()),
                                     ctypes.c_void_p(cand.data_ptr()), ctypes.c_void_p(score.data_ptr()), ms))
    return dict(pos=pos, dist=dist, cand=cand, score=score,
                stage_ms=dict(vectors=ms[0], search=ms[1], rescore=ms[2]))


def brute_exact_f64(a, b, qpos=None):
    """Validation: exact FP64 nearest neighbours (pos XY_TO_INT, dist float64), CUDA tensors."""
    import torch
    qp, nq = _queries(a, b, qpos)
    pos = torch.empty(nq, dtype=torch.int32, device=a.device)
    dist = torch.empty(nq, dtype=torch.float64, device=a.device)
    torch.cuda.synchronize()
    check(load().pt_brute_exact_f64_dev(ctypes.c_void_p(a.data_ptr()), int(a.shape[0]), int(a.shape[1]),
                                        ctypes.c_void_p(b.data_ptr()), int(b.shape[0]), int(b.shape[1]), qp, nq,
                                        ctypes.c_void_p(pos.data_ptr()), ctypes.c_void_p(dist.data_ptr())))
    return pos, dist


def _ffi_f32(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _is_device_tensor(x):
    return hasattr(x, "is_cuda") and bool(getattr(x, "is_cuda"))


def _check_dev(x, name, dtype, shape=None, device=None):
    """A torch CUDA tensor argument handed to the library as a raw pointer: it must have the
    dtype and shape the C entry point assumes, be contiguous (the library reads dense HWC
    rows) and live on the device of the call. Anything else raises instead of reading or
    writing out of bounds."""
    if not _is_device_tensor(x):
        raise PatchTableError(f"{name} must be a torch CUDA tensor")
    if x.dtype != dtype:
        raise PatchTableError(f"{name} must be {dtype}, got {x.dtype}")
    if not x.is_contiguous():
        raise PatchTableError(f"{name} must be contiguous")
    if shape is not None and tuple(x.shape) != tuple(shape):
        raise PatchTableError(f"{name} must have shape {tuple(shape)}, got {tuple(x.shape)}")
    if device is not None and x.device != device:
        raise PatchTableError(f"{name} is on {x.device}, expected {device}")
    return x


def _after_torch(device):
    """The library works on its own non-blocking HIP streams and returns only once its work is
    done. Torch's current stream may still be writing an input (or reading a buffer the call
    will overwrite), so wait for it before handing the pointers over."""
    import torch
    torch.cuda.current_stream(device).synchronize()


def load_color_image(filename):
    """load_color_image<float> (solver/array.h:1637-1684): a .png as 8-bit RGB * (1/255.0), a
    .pfm as stored; returns h x w x 3 (h x w x 1 for a one-channel .pfm) float32."""
    h, w, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    check(load().pt_image_size(filename.encode(), ctypes.byref(h), ctypes.byref(w), ctypes.byref(c)))
    out = np.empty((h.value, w.value, c.value), np.float32)
    check(load().pt_image_load(filename.encode(), out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                               h.value, w.value, c.value))
    return out


def save_color_image(a, filename):
    """save_color_image (solver/array.h:1745-1827) of a 1 or 3 channel array (an NNF) to .pfm."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    if a.ndim == 2:
        a = a[..., None]
    check(load().pt_pfm_save(filename.encode(), a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                             a.shape[0], a.shape[1], a.shape[2]))


def synth_pair(seed, h, w):
    """Seeded synthetic (a, b) pair, each h x w x 3 float32 (see include/patchtable_amd.h)."""
    a = np.empty((h, w, 3), np.float32)
    b = np.empty((h, w, 3), np.float32)
    check(load().pt_synth_pair(int(seed), int(h), int(w), a.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                               b.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
    return a, b
