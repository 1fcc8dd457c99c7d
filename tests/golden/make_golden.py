"""Regenerates the committed golden fixtures under tests/golden/.

glibc_rand.json
    Runs tests/golden/glibc_golden.cpp, built with the system g++ against the system glibc and
    libstdc++: the rand() stream of a fresh process and std::random_shuffle permutations. These
    pin the oracle's restatement of the populate order (patchtable.h:3121-3159).

gck_recurrences.json
    Reads the reference's generated patchtable/gck.h as text and records, for each kernel index
    2..64, the recurrence axis, offset and sign of its "Kernel k" block (gck.h:222-1946). These
    pin the oracle's restatement of gck.py's snake order and Walsh-Hadamard alpha signs.

Run from the repository root:  python tests/golden/make_golden.py  (needs /root/reference).
"""
import json
import os
import re
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
REFERENCE_GCK = "/root/reference/patchtable/gck.h"


def make_glibc():
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "glibc_golden")
        subprocess.check_call(["g++", "-O2", "-o", exe, os.path.join(HERE, "glibc_golden.cpp")])
        out = subprocess.check_output([exe]).decode()
    data = json.loads(out)
    with open(os.path.join(HERE, "glibc_rand.json"), "w") as f:
        json.dump(data, f)


def make_gck():
    text = open(REFERENCE_GCK).read()
    # First (generic, get_zero) block of each kernel in the y<4 loop.
    pat = re.compile(
        r"/\* Kernel (\d+) \*/\s*if \(n > \d+\) \{\s*for \(int c = 0; c < [^;]*; c\+\+\) \{\s*"
        r"out\(y, x, [^=]*\) = (-?1) \* out\.get_zero\(([^,]*), ([^,]*), ")
    rows = {}
    for m in pat.finditer(text):
        k = int(m.group(1))
        if k in rows:
            continue
        sign = int(m.group(2))
        yexpr, xexpr = m.group(3).strip(), m.group(4).strip()
        if yexpr == "y":
            axis, delta = 0, int(xexpr.split("-")[1])
        else:
            axis, delta = 1, int(yexpr.split("-")[1])
        rows[k] = {"axis": axis, "delta": delta, "sign": sign}
    table = [dict(kernel=k, **rows[k]) for k in sorted(rows)]
    with open(os.path.join(HERE, "gck_recurrences.json"), "w") as f:
        json.dump({"source": "patchtable/gck.h (kernel k block, k=2..64)", "kernels": table}, f, indent=1)


if __name__ == "__main__":
    make_glibc()
    if os.path.exists(REFERENCE_GCK):
        make_gck()
