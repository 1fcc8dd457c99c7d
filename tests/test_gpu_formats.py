"""GPU: table save/load in the reference's file format, PNG inputs and the match front end.

* A table saved by the library is byte-identical to the oracle's file for the same build, parses
  with the independent reader of test_io, and loads back (directly or through -load) into a
  table whose lookup is bit-identical to the original's.
* vidpair0 (the reference's own pair, match_ours_full.sh / match_ours_full.txt): the mean patch
  distance at each tested speed is within 1% of the reference's published value (a tolerance
  check: the reference's k-coherence lists come from FLANN and its lookup is raster-ordered).
"""
import os
import subprocess

import numpy as np
import pytest

import oracle_ffi as O
from conftest import speed1
from test_io import parse_table_file

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden", "vidpair0")


def reference_mean_dist():
    """match_ours_full.txt: line i+1 is `-speed i` (match_ours_full.sh)."""
    out = {}
    for i, line in enumerate(open(os.path.join(GOLD, "match_ours_full.txt"))):
        if "mean_dist:" in line:
            out[i] = float(line.split("mean_dist:")[1].split(",")[0])
    return out


def test_save_matches_oracle_file_and_loads_back(pair_cfg1, tmp_path):
    from patchtable_amd import PatchTable
    a, b = pair_cfg1
    p = speed1("-limit 1")
    t = PatchTable(p, b)
    gpath, opath = str(tmp_path / "gpu.table"), str(tmp_path / "oracle.table")
    t.save(gpath)
    o = O.OracleTable(p.raw, b)
    o.save(opath)
    assert open(gpath, "rb").read() == open(opath, "rb").read()
    f = parse_table_file(gpath)
    assert f["nslices"] == t.info()["nslices"]
    assert np.array_equal(f["table"].reshape(-1), t.cells())
    assert np.array_equal(f["b0"], b)
    assert np.array_equal(f["wh0"].view(np.uint32), t.descriptor().view(np.uint32))
    for d, P in enumerate(f["part"]):
        mn, mx, centers, to_bin = t.partition(d)
        assert (P["min_val"], P["max_val"], P["nslices"]) == (mn, mx, len(centers))
        assert np.array_equal(P["centers"], centers) and np.array_equal(P["to_bin"], to_bin)
    ann0 = t.lookup(a)
    t2 = PatchTable.load(p, gpath)
    assert np.array_equal(t2.cells(), t.cells())
    gp, gd = t.kcoherence()
    lp, ld = t2.kcoherence()
    assert np.array_equal(gp, lp) and np.array_equal(gd.view(np.uint32), ld.view(np.uint32))
    assert np.array_equal(t2.lookup(a), ann0)
    # the -load switch of the constructor reads the oracle's file
    p2 = speed1(f"-limit 1 -load {opath}")
    t3 = PatchTable(p2, b)
    assert np.array_equal(t3.lookup(a), ann0)


def test_save_switch_writes_file(pair_cfg1, tmp_path):
    from patchtable_amd import PatchTable
    a, b = pair_cfg1
    path = str(tmp_path / "s.table")
    p = speed1(f"-limit 0.5 -save {path}")
    t = PatchTable(p, b)
    assert os.path.exists(path)
    assert np.array_equal(parse_table_file(path)["table"].reshape(-1), t.cells())


@pytest.mark.parametrize("speed", range(10))
def test_vidpair0_mean_dist_vs_published(speed):
    from patchtable_amd import PatchTable, PatchTableParams, load_color_image
    a = load_color_image(os.path.join(GOLD, "a.png"))
    b = load_color_image(os.path.join(GOLD, "b.png"))
    # PatchTableParams(argc, argv) of `patchtable match a b out -speed i -verbose 0`
    p = PatchTableParams(["match", "a.png", "b.png", "out.pfm", "-speed", str(speed), "-verbose", "0"])
    t = PatchTable(p, b)
    ann = t.lookup(a)
    m, _ = O.recompute(a, b, ann)
    ref = reference_mean_dist()[speed]
    print(f"vidpair0 speed {speed}: mean_dist {O.mean_dist(ann):.6f} (recomputed {m:.6f}) reference {ref:.6f}")
    assert abs(m - O.mean_dist(ann)) < 1e-5
    assert O.mean_dist(ann) <= 1.01 * ref


def test_match_cli_on_vidpair0(tmp_path):
    from patchtable_amd import load_color_image
    exe = os.path.join(ROOT, "patchtable_amd", "patchtable")
    out = str(tmp_path / "out.pfm")
    r = subprocess.run([exe, "match", os.path.join(GOLD, "a.png"), os.path.join(GOLD, "b.png"), out,
                        "-speed", "1", "-verbose", "0"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    line = r.stdout.strip().splitlines()[-1]
    assert line.startswith("total_time: ")
    fields = dict(kv.split(": ") for kv in line.split(", "))
    ann = load_color_image(out)
    assert ann.shape == (555 - 7, 1332 - 7, 3)
    assert abs(float(fields["mean_dist"]) - O.mean_dist(ann.astype(np.float64))) < 1e-5
    assert float(fields["mean_dist"]) <= 1.01 * reference_mean_dist()[1]
    r2 = subprocess.run([exe, "check_dist", os.path.join(GOLD, "a.png"), os.path.join(GOLD, "b.png"), out],
                        capture_output=True, text=True, timeout=120)
    assert r2.returncode == 0, r2.stderr
    assert abs(float(r2.stdout.strip()) - float(fields["mean_dist_recomputed"])) < 1e-5


def test_cli_test_descriptor():
    """`patchtable test_descriptor` (patchtable_main.cpp:174-240, 365-366) passes."""
    exe = os.path.join(ROOT, "patchtable_amd", "patchtable")
    r = subprocess.run([exe, "test_descriptor"], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stderr
    assert "test_descriptor: OK" in r.stdout
