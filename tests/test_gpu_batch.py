"""GPU: one query against many databases (pt_batch_*): each phase one launch over every pair,
captured in one graph. Every pair's NNF equals the single-table lookup of the same pair bit for
bit, on the first call (eager, then captured) and on graph replays, with every lattice tile, and
matches the oracle's model of the parallel lookup."""
import numpy as np
import pytest

import oracle_ffi as O

pytestmark = pytest.mark.gpu


def test_batch_matches_single_tables():
    import torch
    from patchtable_amd import PatchTable, PatchTableBatch, PatchTableParams, synth_pair
    sizes = [(120, 160), (200, 150), (96, 260), (170, 170)]
    dbs = [synth_pair(200 + i, h, w)[1] for i, (h, w) in enumerate(sizes)]
    a_np = synth_pair(199, 150, 180)[0]
    a = torch.from_numpy(a_np).cuda()
    p = PatchTableParams()
    p.set_speed(1)
    p.set_from_switches("-x 0 -limit 1")
    batch = PatchTableBatch(p, [torch.from_numpy(d).cuda() for d in dbs], threads=3)
    singles = []
    for d in dbs:
        t = PatchTable(p, d)
        singles.append(t.lookup(a_np))
    for call in range(3):
        anns, ms, secs = batch.lookup(a)
        print(f"call {call}: per-pair ms {[round(m, 3) for m in ms]}, wall {secs * 1e3:.2f} ms", flush=True)
        for i, ann in enumerate(anns):
            assert np.array_equal(ann.cpu().numpy(), singles[i]), (call, i)
        assert all(m > 0 for m in ms)
    o = O.OracleTable(p.raw, dbs[1])
    assert np.array_equal(singles[1], o.lookup_parallel(a_np))
    batch.close()


@pytest.mark.parametrize("switches", ["-x 0 -limit 1", "-x 0 -limit 1 -query_step 2",
                                      "-x 0 -limit 1 -mfma_rounds 4 -mfma_anchors 5"])
def test_batch_one_launch_per_phase_every_tile(switches):
    """The batched kernels (one launch per phase for all pairs) with each lattice tile give every
    pair the single-table lookup's NNF bit for bit, also with query_step 2 (the gather phase) and
    in the MFMA form; a single table gives the same NNF with every tile. Databases of uneven
    sizes, among them 8 and 16 tile columns wide (the XCD column bands)."""
    import torch
    from patchtable_amd import PatchTable, PatchTableBatch, PatchTableParams, synth_pair
    sizes = [(120, 160), (200, 150), (96, 260), (170, 170), (90, 1031)]
    dbs = [synth_pair(240 + i, h, w)[1] for i, (h, w) in enumerate(sizes)]
    a_np = synth_pair(239, 140, 519)[0]          # 512 patches wide: 8 tile columns of 64
    a = torch.from_numpy(a_np).cuda()
    p = PatchTableParams()
    p.set_speed(1)
    p.set_from_switches(switches)
    print(f"batch tiles: {switches}", flush=True)
    batch = PatchTableBatch(p, [torch.from_numpy(d).cuda() for d in dbs], threads=3)
    singles = [PatchTable(p, d) for d in dbs]
    refs = [t.lookup(a_np) for t in singles]
    for tx in (64, 32, 16):
        singles[0].set_tile(tx)
        assert np.array_equal(singles[0].lookup(a_np), refs[0]), tx
        batch.tile_x = tx
        for call in range(2):                    # eager + capture, then a replay
            anns, ms, _ = batch.lookup(a)
            for i, ann in enumerate(anns):
                assert np.array_equal(ann.cpu().numpy(), refs[i]), (tx, call, i)
            assert len(set(ms)) == 1 and ms[0] > 0
    anns = [torch.empty((133, 512, 3), dtype=torch.float64, device="cuda") for _ in dbs]
    res = batch.autotune(a, anns, reps=2)
    print(f"autotune: {res}", flush=True)
    assert res["tile_x"] in (64, 32, 16) and batch.tile_x == res["tile_x"]
    assert all(v > 0 for v in res["median_ms"].values())
    out, _, _ = batch.lookup(a, anns)
    for i in range(len(dbs)):
        assert np.array_equal(out[i].cpu().numpy(), refs[i]), i
    batch.close()


def test_table_autotune_keeps_results():
    """The single-table autotuner times each lattice tile and keeps the fastest; the lookup after
    it gives the same NNF as with the default tile."""
    import torch
    from patchtable_amd import PatchTable, synth_pair
    from conftest import speed1
    a_np, b_np = synth_pair(260, 150, 530)
    for switches in ("-limit 1", "-limit 1 -mfma_rounds 3 -mfma_anchors 4"):
        p = speed1(switches)
        t = PatchTable(p, torch.from_numpy(b_np).cuda())
        ref = t.lookup(a_np)
        a = torch.from_numpy(a_np).cuda()
        ann = torch.empty((143, 523, 3), dtype=torch.float64, device="cuda")
        res = t.autotune(a, ann, reps=2)
        print(f"autotune {switches}: {res}", flush=True)
        assert res["tile_x"] in (64, 32, 16)
        assert np.array_equal(t.lookup(a, ann).cpu().numpy(), ref)


def test_batch_replay_follows_parameter_changes():
    """The tables read the caller's parameters at every lookup (patchtable.h:2602). Changing
    lookup-time fields between calls with the same buffers must not replay a graph captured with
    the old values: each call equals the single-table lookup under the current parameters."""
    import torch
    from patchtable_amd import PatchTable, PatchTableBatch, PatchTableParams, synth_pair
    sizes = [(110, 140), (130, 120)]
    dbs = [synth_pair(220 + i, h, w)[1] for i, (h, w) in enumerate(sizes)]
    a_np = synth_pair(219, 120, 150)[0]
    a = torch.from_numpy(a_np).cuda()
    p = PatchTableParams()
    p.set_speed(1)
    p.set_from_switches("-x 0 -limit 1")
    batch = PatchTableBatch(p, [torch.from_numpy(d).cuda() for d in dbs], threads=2)
    singles = [PatchTable(p, d) for d in dbs]
    anns = [torch.empty((113, 143, 3), dtype=torch.float64, device="cuda") for _ in dbs]
    settings = [(6, 1, 10), (6, 1, 10), (2, 1, 10), (2, 0, 4), (6, 1, 10)]   # jf_rounds, spatial, kcoherence
    for call, (jr, sp, kc) in enumerate(settings):
        p.jf_rounds, p.spatial, p.kcoherence = jr, sp, kc
        out, _, _ = batch.lookup(a, anns)
        for i in range(len(dbs)):
            ref = singles[i].lookup(a_np)
            assert np.array_equal(out[i].cpu().numpy(), ref), (call, i)
    batch.close()


def test_device_arguments_are_checked():
    """Device tensors handed to the library must have the dtype, shape and layout the C entry
    points assume; anything else raises before a pointer is passed."""
    import torch
    from patchtable_amd import PatchTable, PatchTableError, PatchTableParams, synth_pair
    a_np, b_np = synth_pair(230, 64, 72)
    p = PatchTableParams()
    p.set_speed(1)
    p.set_from_switches("-x 0 -limit 0.1")
    b = torch.from_numpy(b_np).cuda()
    with pytest.raises(PatchTableError, match="float32"):
        PatchTable(p, b.double())
    with pytest.raises(PatchTableError, match="contiguous"):
        PatchTable(p, b.transpose(0, 1))
    t = PatchTable(p, b)
    a = torch.from_numpy(a_np).cuda()
    with pytest.raises(PatchTableError, match="float32"):
        t.lookup(a.half())
    with pytest.raises(PatchTableError, match="shape"):
        t.lookup(a, ann=torch.empty((10, 10, 3), dtype=torch.float64, device="cuda"))
    with pytest.raises(PatchTableError, match="float64"):
        t.lookup(a, ann=torch.empty((57, 65, 3), dtype=torch.float32, device="cuda"))
    # a field written by a torch kernel just before the call is read complete
    a2 = torch.zeros_like(a)
    a2.copy_(a)
    assert np.array_equal(t.lookup(a2).cpu().numpy(), t.lookup(a_np))
