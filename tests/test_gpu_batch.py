"""GPU: one query against many databases (pt_batch_*). Every pair's NNF equals the single-table
lookup of the same pair bit for bit, on the first call (eager, then captured) and on graph
replays, and matches the oracle's model of the parallel lookup."""
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
