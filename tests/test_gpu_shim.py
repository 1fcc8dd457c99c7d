"""GPU: the C++ drop-in header (include/patchtable/patchtable.h) used the way the reference's
callers use PatchTable<> (patchtable_main.cpp:206-209, 284-299): construct from an Array, look up
twice, save. The program is host code only (g++), linked against the in-tree library; its NNF
must equal the Python binding's lookup and the oracle's parallel model bit for bit."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

import oracle_ffi as O
from conftest import ROOT, speed1

pytestmark = pytest.mark.gpu

PROGRAM = r'''
#include "patchtable/patchtable.h"
int main(int argc, char **argv) {
    /* argv: out.bin saved_table seed h w [switches...]; argv[0] is skipped by the parser */
    PatchTableParams p;
    p.set_speed(1);
    std::string sw = "prog";
    for (int i = 6; i < argc; i++) { sw += " "; sw += argv[i]; }
    p.set_from_switches(sw);
    const int seed = atoi(argv[3]), h = atoi(argv[4]), w = atoi(argv[5]);
    Array<float> a(h, w, 3), b(h, w, 3);
    if (pt_synth_pair(seed, h, w, a.data, b.data) != 0) return 2;
    PatchTable<> table(&p, b);
    Array<double> ann;
    table.lookup(a, ann);
    double secs = table.lookup(a, ann);   /* the warm run, as test_match times it */
    table.save(argv[2]);
    FILE *f = fopen(argv[1], "wb");
    fwrite(ann.data, sizeof(double), ann.nelems, f);
    fclose(f);
    printf("%d %d %d %.6f\n", ann.height(), ann.width(), ann.channels(), secs);
    return 0;
}
'''


def test_cpp_shim_constructs_and_looks_up():
    from patchtable_amd import PatchTable, synth_pair
    h, w, seed = 150, 200, 61
    with tempfile.TemporaryDirectory() as d:
        src, exe = os.path.join(d, "shim.cpp"), os.path.join(d, "shim")
        with open(src, "w") as f:
            f.write(PROGRAM)
        libdir = os.path.join(ROOT, "patchtable_amd")
        subprocess.check_call(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"), src, "-o", exe,
                               "-L", libdir, "-lpatchtable_hip", f"-Wl,-rpath,{libdir}"])
        out_bin, tab = os.path.join(d, "ann.bin"), os.path.join(d, "t.bin")
        print("running the C++ shim program", flush=True)
        r = subprocess.run([exe, out_bin, tab, str(seed), str(h), str(w), "-limit", "1"], capture_output=True,
                           text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        hh, ww, cc, secs = r.stdout.split()
        assert (int(hh), int(ww), int(cc)) == (h - 7, w - 7, 3) and float(secs) > 0
        ann = np.fromfile(out_bin, np.float64).reshape(h - 7, w - 7, 3)
        a, b = synth_pair(seed, h, w)
        p = speed1("-limit 1")
        assert np.array_equal(ann, PatchTable(p, b).lookup(a))
        o = O.OracleTable(p.raw, b)
        assert np.array_equal(ann, o.lookup_parallel(a))
        # the saved table reloads to the same cells
        ref = os.path.join(d, "oracle.bin")
        o.save(ref)
        assert open(tab, "rb").read() == open(ref, "rb").read()
