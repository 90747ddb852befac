"""PNG codec of the command-line evaluator (include/pdflow_png.h) against an independent decoder
and encoder (PIL), and the reference's frame conversion (scene_flow_impair.cpp:218-277)."""
import os
import subprocess

import numpy as np
import pytest
from PIL import Image

from pdflow import png, synth_pair

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "pd-flow_amd", "pdflow", "pdflow-impair")


@pytest.mark.parametrize("mode,dtype,shape", [("L", np.uint8, (37, 53)), ("I;16", np.uint16, (40, 31)),
                                              ("RGB", np.uint8, (21, 17, 3)), ("RGBA", np.uint8, (9, 11, 4)),
                                              ("LA", np.uint8, (8, 5, 2))])
def test_reader_matches_pil(tmp_path, mode, dtype, shape):
    rng = np.random.default_rng(len(shape) * 7 + shape[0])
    a = rng.integers(0, np.iinfo(dtype).max + 1, size=shape, dtype=np.int64).astype(dtype)
    # smooth rows too, so that PIL's encoder picks the Sub/Up/Average/Paeth filters
    a[: shape[0] // 2] = np.sort(a[: shape[0] // 2], axis=1)
    p = tmp_path / "x.png"
    Image.fromarray(a).save(p, optimize=True)
    assert Image.open(p).mode == mode
    got = png.read(p)
    assert got.dtype == dtype and np.array_equal(got, a)


def test_palette_and_pil_reads_our_writer(tmp_path):
    rng = np.random.default_rng(3)
    idx = rng.integers(0, 16, size=(12, 20), dtype=np.uint8)
    im = Image.fromarray(idx, mode="P")
    pal = rng.integers(0, 256, size=48, dtype=np.uint8)
    im.putpalette(list(pal))
    p = tmp_path / "p.png"
    im.save(p)
    got = png.read(p)
    assert np.array_equal(got, pal.reshape(16, 3)[idx])
    for arr in (rng.integers(0, 256, (13, 7), dtype=np.uint8), rng.integers(0, 65536, (5, 9), dtype=np.uint16).astype(np.uint16),
                rng.integers(0, 256, (6, 4, 3), dtype=np.uint8)):
        q = tmp_path / "w.png"
        png.write(q, arr)
        back = np.array(Image.open(q))
        assert np.array_equal(back.astype(arr.dtype), arr)
        assert np.array_equal(png.read(q), arr)


def test_corrupt_files_are_rejected(tmp_path):
    p = tmp_path / "c.png"
    png.write(p, np.zeros((4, 4), np.uint8))
    data = bytearray(p.read_bytes())
    data[30] ^= 0xFF                     # inside IHDR/IDAT: CRC mismatch
    p.write_bytes(bytes(data))
    with pytest.raises(OSError):
        png.read(p)
    q = tmp_path / "n.png"
    q.write_bytes(b"not a png at all")
    with pytest.raises(OSError):
        png.read(q)


def test_load_frame_matches_the_reference_conversion(tmp_path):
    i1, z1, _, _ = synth_pair(5, 240, 320)
    I8 = i1.reshape(320, 240).T          # column-major planes -> (rows, cols) images
    Z16 = z1.reshape(320, 240).T
    png.write(tmp_path / "i.png", np.ascontiguousarray(I8))
    png.write(tmp_path / "z.png", np.ascontiguousarray(Z16))
    I, Z = png.load_frame(tmp_path / "i.png", tmp_path / "z.png", 320, 240)
    assert np.array_equal(I, i1.astype(np.float32))
    # convertTo(CV_32FC1, 1/5000): the double product rounded to float
    assert np.array_equal(Z, (z1.astype(np.float64) * (1.0 / 5000.0)).astype(np.float32))


def test_cli_rejects_bad_arguments_without_a_gpu():
    r = subprocess.run([CLI, "--rows", "100"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "--rows r" in r.stdout
    r = subprocess.run([CLI, "--i1", "/nonexistent.png", "--no-show"], capture_output=True, text=True, timeout=60)
    assert "cannot be found" in r.stdout
