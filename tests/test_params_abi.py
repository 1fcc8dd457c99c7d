"""Host logic without a GPU: parameter defaults/presets/switches (patchtable.cpp:26-387), the
C ABI's exported symbols and error paths, and the C++ drop-in header."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT


def P(argv=None):
    from patchtable_amd import PatchTableParams
    return PatchTableParams(argv)


def test_defaults_are_set_defaults_then_speed3():
    p = P()
    # set_defaults (patchtable.cpp:26-134) then set_speed(3) (:142); the preset's leading
    # "-ndims 6" is the dropped first token.
    assert (p.ndims, p.nchroma, p.ntables) == (6, 1, 1)
    assert p.limit == 100e6
    assert (p.do_rs, p.run_dt, p.partition_step, p.dt_iters, p.do_prop, p.query_step) == (0, 1, 1, -1, 1, 1)
    assert (p.kcoherence, p.kcoherence_iter, p.prop_iters, p.spatial) == (2, -1, 2, 0)
    assert (p.threads, p.dt_algo, p.populate_random, p.colorspace, p.dim_algo) == (1, 1, 1, 0, 0)
    assert (p.triangle_factor, p.kcoherence_min_dist, p.patch_w, p.grid_ndims) == (2.0, 2, 8, -1)
    assert (p.flann_checks, p.flann_eps, p.flann_reset_step, p.query_iters, p.verbose) == (1, 500000, 10, 2, 1)


def test_speed_presets_effective_values():
    expect = {  # (limit, do_rs, kcoherence, kcoherence_iter, prop_iters, spatial, query_step); ndims stays 6
        0: (10e6, 1, 20, -1, 2, 1, 1), 1: (100e6, 1, 10, 0, 2, 1, 1), 2: (100e6, 1, 10, 0, 2, 0, 1),
        3: (100e6, 0, 2, -1, 2, 0, 1), 4: (10e6, 1, 20, 0, 1, 1, 2), 5: (100e6, 1, 4, 0, 1, 1, 2),
        6: (100e6, 0, 2, -1, 1, 0, 2), 7: (1e6, 1, 3, 0, 1, 1, 4), 8: (100e6, 1, 5, 0, 1, 0, 6),
        9: (100e6, 0, 5, 0, 1, 0, 10)}
    for s, e in expect.items():
        p = P()
        p.set_speed(s)
        assert (p.limit, p.do_rs, p.kcoherence, p.kcoherence_iter, p.prop_iters, p.spatial, p.query_step) == e, s
        assert p.ndims == 6 and p.nchroma == 1


def test_cli_argv_and_switch_errors():
    from patchtable_amd import PatchTableError
    p = P(["patchtable", "match", "a.png", "b.png", "out.pfm", "-speed", "1", "-limit", "2.5", "-dt_algo", "euclidean"])
    assert (p.kcoherence, p.spatial, p.limit, p.dt_algo) == (10, 1, 2.5e6, 6)
    p = P()
    p.set_from_switches("-x 0 -nslices 10,20,30")
    assert p.ndims == 3 and p.nslices == [10, 20, 30]
    p = P()
    p.set_from_switches("-x 0 -mfma_rounds 4 -mfma_jumps 16,4,1,1")
    assert p.mfma_rounds == 4 and list(p.raw.mfma_jumps)[:5] == [16, 4, 1, 1, 0]
    assert list(P().raw.mfma_jumps) == [0] * 16
    for bad in ("16,,1", "16,x", "0,1", "2000", ",".join(["1"] * 17)):
        with pytest.raises(PatchTableError, match="mfma_jumps"):
            P().set_from_switches("-x 0 -mfma_jumps " + bad)
    p = P()
    p.set_from_switches("-x 0 -auto_filter 1")
    assert p.filter_dims == [2, 1, 5, 0, 3, 10]
    p = P()
    p.set_from_switches("-limit 3")          # first token dropped: unchanged
    assert p.limit == 100e6
    with pytest.raises(PatchTableError, match="invalid dt_algo"):
        P().set_from_switches("-x 0 -dt_algo nope")
    with pytest.raises(PatchTableError, match="set_speed"):
        P().set_speed(10)


def declared_functions():
    text = open(os.path.join(ROOT, "include", "patchtable_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pt_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(lib):
    names = declared_functions()
    assert len(names) >= 20
    out = subprocess.check_output(["nm", "-D", "--defined-only", os.path.join(ROOT, "patchtable_amd", "libpatchtable_hip.so")]).decode()
    exported = set(l.split()[-1] for l in out.splitlines() if " T " in l)
    missing = [n for n in names if n not in exported]
    assert not missing, missing


def test_params_struct_layout_matches_c():
    from patchtable_amd._ffi import PtParams
    src = os.path.join(ROOT, "_build_sizeof_test.c")
    exe = os.path.join(ROOT, "_build_sizeof_test")
    try:
        with open(src, "w") as f:
            f.write('#include <stdio.h>\n#include <stddef.h>\n#include "patchtable_amd.h"\n'
                    'int main(){printf("%zu %zu %zu\\n", sizeof(pt_params), offsetof(pt_params, kcoh_window),'
                    ' offsetof(pt_params, save_filename));return 0;}\n')
        subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), "-o", exe, src])
        size, off, off2 = (int(v) for v in subprocess.check_output([exe]).split())
    finally:
        for pth in (src, exe):
            if os.path.exists(pth):
                os.remove(pth)
    assert ctypes.sizeof(PtParams) == size
    assert PtParams.kcoh_window.offset == off and PtParams.save_filename.offset == off2


def test_create_rejects_bad_input_before_gpu_work(lib):
    from patchtable_amd import PatchTable, PatchTableError
    b = np.zeros((20, 20, 3), np.float32)
    p = P()
    p.is_descriptor = 1   # descriptor mode needs at least 4 channels (patch_dist_approx reads 4)
    with pytest.raises(PatchTableError, match="3 channels unsupported"):
        PatchTable(p, b)
    p = P()
    p.is_descriptor = 1
    with pytest.raises(PatchTableError, match="allows no patch"):
        PatchTable(p, np.zeros((20, 20, 8), np.float32), allowed_patches=np.zeros((20, 20), np.int32))
    p = P()
    p.ntables = 40
    with pytest.raises(PatchTableError, match="ntables"):
        PatchTable(p, b)
    p = P()
    with pytest.raises(PatchTableError, match="at least 8x8"):
        PatchTable(p, np.zeros((5, 20, 3), np.float32))
    with pytest.raises(PatchTableError, match="3 channel"):
        PatchTable(P(), np.zeros((20, 20, 4), np.float32))
    p = P()
    p.set_from_switches("-x 0 -lookup_algo pm")
    with pytest.raises(PatchTableError, match="lookup_algo"):
        PatchTable(p, b)


def test_cpp_header_shim_compiles_and_parses_params():
    """The C++ drop-in header builds against the library and reproduces PatchTableParams."""
    src = os.path.join(ROOT, "_build_shim_test.cpp")
    exe = os.path.join(ROOT, "_build_shim_test")
    code = r'''
#include "patchtable/patchtable.h"
int main(int argc, char **argv) {
    PatchTableParams p(argc, argv);
    printf("%d %d %g %d %d\n", p.ndims, p.kcoherence, p.limit, p.spatial, p.prop_iters);
    Array<float> img(10, 12, 3);
    img.clear(0.5f);
    printf("%d %d %d %d\n", img.height(), img.width(), img.channels(), img.nelems);
    /* construct and lookup run on the GPU in tests/test_gpu_shim.py; here the code only has to
       compile and link against the library */
    if (argc < 0) { PatchTable<> t(&p, img); Array<double> ann; t.lookup(img, ann); }
    return 0;
}
'''
    try:
        with open(src, "w") as f:
            f.write(code)
        libdir = os.path.join(ROOT, "patchtable_amd")
        subprocess.check_call(["g++", "-std=c++17", "-I", os.path.join(ROOT, "include"), src, "-o", exe,
                               "-L", libdir, "-lpatchtable_hip", f"-Wl,-rpath,{libdir}"])
        out = subprocess.check_output([exe, "-speed", "1", "-limit", "3"]).decode().split("\n")
    finally:
        for pth in (src, exe):
            if os.path.exists(pth):
                os.remove(pth)
    assert out[0].split() == ["6", "10", "3e+06", "1", "2"]
    assert out[1].split() == ["10", "12", "3", "360"]


def test_synth_pair_deterministic():
    from patchtable_amd import synth_pair
    a1, b1 = synth_pair(1, 30, 40)
    a2, b2 = synth_pair(1, 30, 40)
    a3, _ = synth_pair(2, 30, 40)
    assert np.array_equal(a1, a2) and np.array_equal(b1, b2) and not np.array_equal(a1, a3)
    assert a1.min() >= 0 and a1.max() <= 1 and b1.min() >= 0 and b1.max() <= 1


def test_packing_size_limits_rejected_before_launch():
    """XY_TO_INT keeps y in the bits above 16 of an int32, so a padded y of 32768 or more would
    decode as negative: heights above 32768 (descriptor mode: descriptor height + 7) are
    rejected on the host, before any device work (this runs without a GPU)."""
    from patchtable_amd import PatchTable, PatchTableError, _ffi
    with pytest.raises(PatchTableError, match="too large"):
        PatchTable(P(), np.zeros((32769, 8, 3), np.float32))
    p = P()
    p.is_descriptor = 1
    with pytest.raises(PatchTableError, match="too large"):
        PatchTable(p, np.zeros((32762, 1, 4), np.float32))
    lib = _ffi.load()
    fake = ctypes.c_void_p(256)   # never dereferenced: the size check comes first
    pos, dist = ctypes.c_void_p(512), ctypes.c_void_p(768)
    assert lib.pt_brute_exact_f64_dev(fake, 16, 16, fake, 32769, 8, None, 0, pos, dist) == _ffi.PT_E_INVALID
    assert "too large" in lib.pt_last_error().decode()
    assert lib.pt_brute_exact_f64_dev(fake, 32769, 16, fake, 16, 16, None, 0, pos, dist) == _ffi.PT_E_INVALID
