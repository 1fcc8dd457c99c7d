"""Image and NNF file formats (CPU): PNG loading, PFM read/write, the table file layout.

PNG decoding is pinned against PIL on the reference's own vidpair0 inputs and on synthetic
images of every colour type; PFM against an independent numpy reader/writer of the format of
solver/pfm.cpp:15-116 (little-endian, bottom row first)."""
import os
import struct

import numpy as np
import pytest

from patchtable_amd import load_color_image, save_color_image

GOLD = os.path.join(os.path.dirname(__file__), "golden", "vidpair0")
PIL = pytest.importorskip("PIL.Image")


def pil_rgb(path):
    return np.asarray(PIL.open(path).convert("RGB"))


@pytest.mark.parametrize("name", ["a.png", "b.png"])
def test_png_matches_pil_on_vidpair0(name):
    path = os.path.join(GOLD, name)
    got = load_color_image(path)
    ref8 = pil_rgb(path)
    assert got.shape == (555, 1332, 3)
    # load_color_image<float>: (real)(v * (1/255.0)) with the product in double
    ref = (ref8.astype(np.float64) * (1.0 / 255.0)).astype(np.float32)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


@pytest.mark.parametrize("mode", ["RGB", "RGBA", "L", "LA", "P", "I;16", "1"])
def test_png_colour_types(tmp_path, mode):
    rng = np.random.default_rng(7)
    rgb = rng.integers(0, 256, (23, 37, 3), dtype=np.uint8)
    img = PIL.fromarray(rgb, "RGB")
    if mode == "P":
        img = img.quantize(colors=50)
    elif mode == "I;16":
        img = PIL.fromarray(rng.integers(0, 65536, (23, 37), dtype=np.uint16))
    elif mode == "1":
        img = img.convert("1")
    elif mode != "RGB":
        img = img.convert(mode)
    path = str(tmp_path / f"t_{mode.replace(';', '_')}.png")
    img.save(path)
    got = load_color_image(path)
    ref8 = pil_rgb(path)
    if mode == "I;16":   # 16-bit grey: the high byte of each sample
        raw = np.asarray(PIL.open(path)).astype(np.uint16)
        ref8 = np.repeat((raw >> 8).astype(np.uint8)[..., None], 3, axis=2)
    ref = (ref8.astype(np.float64) * (1.0 / 255.0)).astype(np.float32)
    assert np.array_equal(got, ref)


def write_pfm_np(path, arr, little=True):
    h, w, c = arr.shape
    with open(path, "wb") as f:
        f.write(b"PF\n" if c == 3 else b"Pf\n")
        f.write(f"{w} {h}\n".encode())
        f.write(b"-1.000000\n" if little else b"1.000000\n")
        data = arr[::-1].astype("<f4" if little else ">f4")
        f.write(data.tobytes())


def read_pfm_np(path):
    with open(path, "rb") as f:
        tag = f.readline().strip()
        w, h = (int(v) for v in f.readline().split())
        scale = float(f.readline())
        c = 3 if tag == b"PF" else 1
        data = np.frombuffer(f.read(), dtype="<f4" if scale < 0 else ">f4")
    return data.reshape(h, w, c)[::-1]


def test_pfm_write_layout(tmp_path):
    rng = np.random.default_rng(3)
    ann = rng.normal(size=(11, 17, 3)) * 100
    path = str(tmp_path / "nnf.pfm")
    save_color_image(ann, path)
    raw = open(path, "rb").read()
    assert raw.startswith(b"PF\n17 11\n-1.000000\n")
    assert np.array_equal(read_pfm_np(path), ann.astype(np.float32))
    assert np.array_equal(load_color_image(path), ann.astype(np.float32))


@pytest.mark.parametrize("little", [True, False])
def test_pfm_read_both_byte_orders(tmp_path, little):
    rng = np.random.default_rng(4)
    arr = rng.normal(size=(9, 5, 3)).astype(np.float32)
    path = str(tmp_path / "x.pfm")
    write_pfm_np(path, arr, little)
    assert np.array_equal(load_color_image(path), arr)


def test_pfm_one_channel(tmp_path):
    arr = np.arange(12, dtype=np.float64).reshape(3, 4)
    path = str(tmp_path / "m.pfm")
    save_color_image(arr, path)
    assert open(path, "rb").read().startswith(b"Pf\n4 3\n")
    got = load_color_image(path)
    assert got.shape == (3, 4, 1) and np.array_equal(got[..., 0], arr.astype(np.float32))


def test_bad_files_raise(tmp_path):
    from patchtable_amd import PatchTableError
    p = tmp_path / "junk.png"
    p.write_bytes(b"not a png at all")
    with pytest.raises(PatchTableError):
        load_color_image(str(p))
    with pytest.raises(PatchTableError):
        load_color_image(str(tmp_path / "missing.pfm"))


def parse_table_file(path):
    """Independent reader of the reference's PatchTable file (patchtable.h:1574-1593,
    1145-1170, 1334-1353; patchtable.cpp:499-573)."""
    buf = open(path, "rb").read()
    off = 0

    def i32():
        nonlocal off
        v = struct.unpack_from("<i", buf, off)[0]
        off += 4
        return v

    def f32():
        nonlocal off
        v = struct.unpack_from("<f", buf, off)[0]
        off += 4
        return v

    def string():
        nonlocal off
        n = i32()
        s = buf[off:off + n].decode()
        off += n
        return s

    def vec(dtype):
        nonlocal off
        n = i32()
        a = np.frombuffer(buf, dtype=dtype, count=n, offset=off)
        off += n * 4
        return a

    def array(dtype):
        nonlocal off
        sizes = vec("<i4").tolist()
        n = int(np.prod(sizes))
        a = np.frombuffer(buf, dtype=dtype, count=n, offset=off).reshape(sizes)
        off += n * 4
        return a

    assert string() == "PatchTable"
    out = {"table": array("<i4"), "b0": array("<f4"), "wh0": array("<f4"), "nslices": vec("<i4").tolist()}
    assert string() == "PatchPartition"
    parts = []
    for _ in range(i32()):
        assert string() == "PatchPartition1D"
        P = {"to_bin": vec("<i4"), "centers": vec("<f4"), "lsizes": vec("<f4"), "rsizes": vec("<f4"),
             "min_val": f32(), "max_val": f32(), "range": f32(), "nslices": i32()}
        assert string() == "EndPatchPartition1D"
        parts.append(P)
    assert string() == "EndPatchPartition"
    assert string() == "EndPatchTable"
    assert off == len(buf)
    out["part"] = parts
    return out
