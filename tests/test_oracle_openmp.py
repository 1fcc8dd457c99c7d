"""The OpenMP build of the oracle (oracle/liboracle_omp.so), which the benchmark's CPU baseline
times on all host threads: with pt_params.threads > 1 its k-coherence lists and table cells are
those of one thread (patches are independent), its 1-thread raster lookup is the checker's, and
its multi-thread raster lookup (rows in parallel, as under the reference's OpenMP,
patchtable.h:3617-3618) stays within 2% of the 1-thread mean SSD. Runs in a child process, so
the checker library this process loads is untouched."""
import json
import os
import subprocess
import sys

from conftest import ROOT

CHILD = r'''
import json, sys
sys.path.insert(0, ROOT); sys.path.insert(0, ROOT + "/tests")
import numpy as np
import oracle_ffi as O
O.use_openmp()
from patchtable_amd import PatchTableParams, synth_pair
a, b = synth_pair(1, 150, 200)
def params(threads):
    p = PatchTableParams(); p.set_speed(1); p.threads = threads
    return p
p1, p4 = params(1), params(4)
t1, t4 = O.OracleTable(p1.raw, b), O.OracleTable(p4.raw, b)
k1, k4 = t1.kcoherence(), t4.kcoherence()
out = {"cells": bool(np.array_equal(t1.table(), t4.table())),
       "kpos": bool(np.array_equal(k1[0], k4[0])), "kdist": bool(np.array_equal(k1[1], k4[1]))}
r1 = t1.lookup_raster(a)
t4.set_params(p4.raw)
r4 = t4.lookup_raster(a)
out["ssd1"] = float(r1[..., 2].mean())
out["ssd4"] = float(r4[..., 2].mean())
np.save(sys.argv[1], r1)
print(json.dumps(out))
'''


def test_openmp_oracle_matches_single_thread(tmp_path):
    import numpy as np
    import oracle_ffi as O
    from patchtable_amd import PatchTableParams, synth_pair
    subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    out_npy = str(tmp_path / "r1.npy")
    r = subprocess.run([sys.executable, "-c", "ROOT = %r\n" % ROOT + CHILD, out_npy], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    res = json.loads(r.stdout.strip().splitlines()[-1])
    assert res["cells"] and res["kpos"] and res["kdist"]
    assert abs(res["ssd4"] / res["ssd1"] - 1) < 0.02, res
    # the 1-thread raster lookup of the OpenMP build is the checker's, bit for bit
    a, b = synth_pair(1, 150, 200)
    p = PatchTableParams()
    p.set_speed(1)
    ref = O.OracleTable(p.raw, b).lookup_raster(a)
    assert np.array_equal(np.load(out_npy), ref)
