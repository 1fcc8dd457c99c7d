"""GPU: pt_table_lookup_dev_noalloc (the lookup without device allocation) through the C ABI
against pt_table_lookup_dev, bit for bit, on consecutive calls whose second query is smaller
(the shared device array keeps the first call's contents), with query_step 2, with a previous
field, in descriptor mode, in the MFMA form, and a query too large for the array rejected before
any launch."""
import ctypes

import numpy as np
import pytest

from conftest import speed1

pytestmark = pytest.mark.gpu


def _vp(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _both(lib, table, a, prev=None):
    """(lookup_dev result, lookup_dev_noalloc result) for device query a."""
    import torch
    h, w, c = (int(s) for s in a.shape)
    shp = table.ann_shape(h, w)
    ref = torch.full(shp, -1.0, dtype=torch.float64, device="cuda")
    got = torch.full(shp, -2.0, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    s1, s2 = ctypes.c_double(), ctypes.c_double()
    assert lib.pt_table_lookup_dev(table._t, _vp(a), h, w, c, _vp(ref), _vp(prev), None, ctypes.byref(s1)) == 0
    rc = lib.pt_table_lookup_dev_noalloc(table._t, _vp(a), h, w, c, _vp(got), _vp(prev), None, ctypes.byref(s2))
    assert rc == 0, lib.pt_last_error().decode()
    assert s2.value > 0
    return ref.cpu().numpy(), got.cpu().numpy()


def test_noalloc_matches_lookup_dev_two_calls(lib):
    import torch
    from patchtable_amd import PatchTable, synth_pair
    print("noalloc: 300x400 then 150x220 queries against a 300x400 table", flush=True)
    a1, b = synth_pair(71, 300, 400)
    a2, _ = synth_pair(72, 150, 220)
    p = speed1()
    table = PatchTable(p, torch.from_numpy(b).cuda())
    for a in (a1, a2):
        ref, got = _both(lib, table, torch.from_numpy(a).cuda())
        assert np.array_equal(ref, got)
    # with a previous field (the initial field also lives in the shared array)
    prev = torch.from_numpy(np.ascontiguousarray(ref)).cuda()
    ref, got = _both(lib, table, torch.from_numpy(a2).cuda(), prev)
    assert np.array_equal(ref, got)


def test_noalloc_query_step_and_descriptor_mode(lib):
    import torch
    from patchtable_amd import PatchTable, synth_pair
    print("noalloc: query_step 2, then descriptor mode", flush=True)
    a, b = synth_pair(73, 200, 260)
    p = speed1("-query_step 2 -limit 1")
    table = PatchTable(p, torch.from_numpy(b).cuda())
    ref, got = _both(lib, table, torch.from_numpy(a).cuda())
    assert np.array_equal(ref, got)
    rng = np.random.default_rng(5)
    bd = rng.random((90, 120, 8), dtype=np.float32)
    ad = rng.random((70, 100, 8), dtype=np.float32)
    pd = speed1("-limit 0.5")
    pd.is_descriptor = 1
    dt = PatchTable(pd, torch.from_numpy(bd).cuda())
    ref, got = _both(lib, dt, torch.from_numpy(ad).cuda())
    assert np.array_equal(ref, got)


def test_noalloc_mfma_form_two_calls(lib):
    """The MFMA lookup form in the shared array (its int8 query copy lives there too)."""
    import torch
    from patchtable_amd import PatchTable, synth_pair
    print("noalloc: MFMA form, 260x300 then 120x170 queries", flush=True)
    a1, b = synth_pair(75, 260, 300)
    a2, _ = synth_pair(76, 120, 170)
    table = PatchTable(speed1("-limit 1 -mfma_rounds 4 -mfma_anchors 5"), torch.from_numpy(b).cuda())
    for a in (a1, a2):
        ref, got = _both(lib, table, torch.from_numpy(a).cuda())
        assert np.array_equal(ref, got)


def test_noalloc_rejects_queries_beyond_the_array(lib):
    import torch
    from patchtable_amd import PatchTable, synth_pair, _ffi
    _, b = synth_pair(74, 64, 64)
    table = PatchTable(speed1("-limit 0.1"), torch.from_numpy(b).cuda())
    h, w = 1600, 1200   # about 250 MB of working set
    a = torch.zeros((h, w, 3), dtype=torch.float32, device="cuda")
    ann = torch.zeros((h - 7, w - 7, 3), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    rc = lib.pt_table_lookup_dev_noalloc(table._t, _vp(a), h, w, 3, _vp(ann), None, None, None)
    assert rc == _ffi.PT_E_INVALID
    assert "scratch" in lib.pt_last_error().decode()
    assert float(ann.abs().max()) == 0.0   # nothing was launched
