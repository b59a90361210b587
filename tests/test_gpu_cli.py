"""The command-line driver (mms-gpu_amd/mmsgpu) against the reference's own
logged runs: the 18 positional arguments of main.cpp:49-70, the built-in
and PNG masks, the single-domain, vortex and srand(seed) random initial
conditions (main.cpp:113-227), M.dat and material.dat (main.cpp:231-245),
and the step records time_marching writes."""
import json
import os
import struct
import subprocess
import zlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
CLI = os.path.join(ROOT, "mms-gpu_amd", "mmsgpu")


def _say(*a):
    print("[gpu-test]", *a, flush=True)


def _kat(key):
    return json.load(open(os.path.join(GOLD, "log_kats.json")))[key]


def _run(tmp_path, args):
    assert os.path.exists(CLI), "mms-gpu_amd/mmsgpu is not built"
    _say("mmsgpu", " ".join(args))
    r = subprocess.run([CLI] + args, cwd=str(tmp_path), capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return r.stdout


def _rows(path):
    return [list(map(float, l.split(","))) for l in open(path) if l.strip()]


def _printed(x, ref, rel=4e-6):
    e = np.floor(np.log10(abs(ref)))
    return abs(x - ref) <= 0.51 * 10.0 ** (e - 5) + rel * abs(ref)


def _check_rows(rows, recs, n):
    assert len(rows) >= n
    for r, rec in zip(rows[:n], recs[:n]):
        _, tindex, t, dt, E2, mx, my, mz, mag, T = rec
        assert int(r[0]) == int(tindex)
        assert _printed(r[1], t) and _printed(r[2], dt), (r[1:3], t, dt)
        assert _printed(r[3], E2), (r[3], E2)
        assert _printed(r[8], T), (r[8], T)
        assert abs(r[4] - mx) <= 2e-5 and abs(r[5] - my) <= 1e-5 * abs(my) + 1e-9


@pytest.mark.parametrize("key", ["square32_random", "square32_vortex", "square16_random"])
def test_cli_builtin_mask_initial_conditions(tmp_path, key):
    """No mask image (the reference's build without an image library):
    every cell but the border; the random and vortex initial states give the
    reference's logged first steps, and M.dat / material.dat hold them."""
    k = _kat(key)
    dims = (k["xdim"], k["ydim"], k["zdim"])
    args = ["dummy.png", "3", "1e-14", "%g" % k["timestep"], "%g" % k["width"], "%g" % k["height"],
            "%g" % k["depth"], *map(str, dims), "1", "1", "0", "0", "1", "sim", str(k["seed"]), str(k["ic"])]
    out = _run(tmp_path, args)
    assert "Initial condition: " + {1: "Vortex", 2: "Random"}[k["ic"]] in out
    _check_rows(_rows(tmp_path / "sim" / "dynamics.dat"), k["records"], 2)


def _png_rgba(mask2):
    """RGBA PNG of a material mask: black where magnetic, white elsewhere; the
    mask's row y = 0 is the bottom image row (load_mask's orientation)."""
    h, w = mask2.shape
    img = np.where(mask2[::-1, :, None] == 1, 0, 255).astype(np.uint8).repeat(4, axis=2)
    img[:, :, 3] = 255
    raw = b"".join(b"\x00" + img[r].tobytes() for r in range(h))

    def chunk(t, b):
        return struct.pack(">I", len(b)) + t + b + struct.pack(">I", zlib.crc32(t + b) & 0xFFFFFFFF)
    return (b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, 6, 0, 0, 0)) +
            chunk(b"IDAT", zlib.compress(raw)) + chunk(b"IEND", b""))


def test_cli_png_mask_single_domain(tmp_path):
    """The nanowire4 mask as a PNG (its size replaces xdim, ydim), single
    domain: the first logged step of the reference's nanowire4 run."""
    k = _kat("nanowire4_fixed_1e-13")
    mask = np.loadtxt(os.path.join(GOLD, "nanowire4_mask.txt"), dtype=np.uint8)
    (tmp_path / "nanowire4.png").write_bytes(_png_rgba(mask))
    args = ["nanowire4.png", "3", "1e-13", "1e-13", "2.1e-07", "1.3e-07", "1.5e-08", "16", "16", "3", "1", "1",
            "0", "0", "1", "nw", "2010", "0"]
    out = _run(tmp_path, args)
    assert "(xdim, ydim, zdim) = (42, 26, 3)" in out
    rows = _rows(tmp_path / "nw" / "dynamics.dat")
    rec = k["records"][0]
    assert int(rows[0][0]) == 1 and _printed(rows[0][3], rec[4]) and _printed(rows[0][8], rec[9])
