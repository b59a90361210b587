"""Host-side pieces that need no GPU: the C ABI exports, the ctypes binding,
seeded workloads and the PNG mask reader."""
import os
import re
import struct
import zlib

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    src = open(os.path.join(ROOT, "include", "mms_gpu.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?(?:int|long|double|char|void)\s*\*?\s*(\w+)\s*\(", src, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol(mg):
    import ctypes
    from mms_gpu import _lib
    declared = _declared_functions()
    assert len(declared) >= 25
    L = ctypes.CDLL(_lib.LIB_PATH)
    for name in declared:
        assert hasattr(L, name), name
    assert set(declared) == set(_lib.SIGNATURES), set(declared) ^ set(_lib.SIGNATURES)
    assert "gfx950" in mg.version()


def test_enum_values_match_python_constants(mg):
    """Every solver, M2L precision and external-field value declared in the
    header has the same value in the Python package."""
    src = open(os.path.join(ROOT, "include", "mms_gpu.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    header = {k: int(v) for k, v in re.findall(r"\bMMS_(\w+)\s*=\s*(\d+)", src)}
    checked = 0
    for k, v in header.items():
        if k.startswith(("SOLVER_", "M2L_", "EXT_")):
            assert getattr(mg, k) == v, k
            checked += 1
    assert checked >= 10 and header["M2L_FP16X2"] == 5


def test_reference_entry_points_present():
    declared = _declared_functions()
    for name in ("calc_potential_exact", "calc_potential_exact_gpu", "fmm_calc", "Hfield", "rk4_step",
                 "time_marching"):
        assert name in declared


def test_splitmix64_known_answers():
    from mms_gpu import workloads as wl
    # published SplitMix64 output for seed 0
    assert [int(v) for v in wl.splitmix64(0, 0, 3)] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4,
                                                        0x06C45D188009454F]
    u = wl.uniform(2010, 1, 100000)
    assert 0 <= u.min() and u.max() < 1 and abs(u.mean() - 0.5) < 0.005
    v = wl.random_unit(2010, 1, 20000)
    assert np.allclose(np.linalg.norm(v, axis=1), 1)
    assert np.abs(v.mean(0)).max() < 0.02


def test_workloads_shapes():
    from mms_gpu import workloads as wl
    c1 = wl.c1_square()
    assert c1.n == 3072 and c1.n_magnetic == 3072
    assert np.allclose(np.linalg.norm(c1.M.reshape(-1, 3), axis=1), float(wl.MS), rtol=1e-6)
    c2 = wl.c2_dots()
    assert c2.n == 256 * 256 * 3 and 60000 <= c2.n_magnetic <= 63000
    Mn = np.linalg.norm(c2.M.reshape(-1, 3), axis=1)
    assert np.allclose(Mn[c2.material == 1], float(wl.MS), rtol=1e-6) and np.all(Mn[c2.material == 0] == 0)
    assert abs(c2.hext[0] - 7957.7) < 0.1
    c3 = wl.c3_domains(side=128, zd=4)
    m = c3.M.reshape(4, 128, 128, 3)
    assert np.all(m[0] == m[3]) and np.all(m[..., 2] == 0)
    p = wl.probes(1, c3.n, 500)
    assert len(p) == 500 and np.all(np.diff(p) > 0) and p.max() < c3.n


def test_reference_default_geometry():
    from mms_gpu import workloads as wl
    mat, M = wl.single_domain_square(16, 16, 3)
    mat = mat.reshape(3, 16, 16)
    assert mat[0].sum() == 0 and mat[2].sum() == 0 and mat[1].sum() == 14 * 14
    assert M.reshape(-1, 3)[mat.ravel() == 1][:, 0].min() == np.float32(8.6e5)


def _png(rows, ctype, filt):
    h, w = len(rows), len(rows[0])
    ch = {0: 1, 2: 3, 6: 4}[ctype]
    raw = b""
    prev = np.zeros(w * ch, np.int32)
    for r in rows:
        cur = np.array(r, np.int32).ravel()
        if filt == 1:
            enc = cur - np.concatenate([np.zeros(ch, np.int32), cur[:-ch]])
        elif filt == 2:
            enc = cur - prev
        else:
            enc = cur
        raw += bytes([filt]) + bytes((enc & 0xFF).astype(np.uint8))
        prev = cur
    def chunk(t, b):
        return struct.pack(">I", len(b)) + t + b + struct.pack(">I", zlib.crc32(t + b) & 0xFFFFFFFF)
    return (b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, ctype, 0, 0, 0)) +
            chunk(b"IDAT", zlib.compress(raw)) + chunk(b"IEND", b""))


@pytest.mark.parametrize("filt", [0, 1, 2])
def test_png_mask_reader(tmp_path, filt):
    from mms_gpu import geometry
    rows = [[(255, 255, 255, 255), (0, 0, 0, 255), (0, 0, 0, 255)],
            [(0, 0, 0, 255), (200, 10, 10, 255), (255, 255, 255, 255)]]
    p = tmp_path / "m.png"
    p.write_bytes(_png(rows, 6, filt))
    mat = geometry.material_from_png(str(p))
    # y = 0 is the bottom image row (FreeImage scanline order)
    assert mat.tolist() == [[1, 0, 0], [0, 1, 1]]
    w, h, grey = geometry.load_mask(str(p))
    assert (w, h) == (3, 2) and grey[0, 1] == int(.3 * 200 + .59 * 10 + .11 * 10)


def test_nanowire4_fixture_mask():
    mask = np.loadtxt(os.path.join(ROOT, "tests", "golden", "nanowire4_mask.txt"), dtype=np.uint8)
    assert mask.shape == (26, 42) and mask.sum() == 360


# ---------------------------------------------------------- command line --
CLI = os.path.join(ROOT, "mms-gpu_amd", "mmsgpu")


def _cli_files(tmp_path, args):
    """Runs the command-line driver with a negative final time: M.dat and
    material.dat are written before the time loop (which needs a GPU; without
    one the driver exits with an error after writing them)."""
    import subprocess
    assert os.path.exists(CLI), "mms-gpu_amd/mmsgpu is not built"
    subprocess.run([CLI] + args, cwd=str(tmp_path), capture_output=True, text=True, timeout=120)
    sim = tmp_path / args[15]
    return open(sim / "M.dat").read().split("\n")[:-1], open(sim / "material.dat").read().split("\n")[:-1]


def _vec_lines(M, dims):
    xd, yd, zd = dims
    v = M.reshape(zd, yd, xd, 3)
    return ["%s %s %s " % tuple("%g" % float(c) for c in v[z, y, x])
            for z in range(zd) for x in range(xd) for y in range(yd)]


def _mat_lines(mat, dims):
    xd, yd, zd = dims
    m = mat.reshape(zd, yd, xd)[1]
    return ["".join("%g " % float(m[y, x]) for x in range(xd)) for y in range(yd)]



@pytest.mark.parametrize("ic", [0, 1, 2])
def test_cli_front_end_built_in_mask(tmp_path, ic):
    """main.cpp:113-245 without a mask image: border mask, layers 1..zdim-2,
    the single-domain / vortex / srand(seed) random states (the last two pin
    against the reference's logged runs through the oracle tests), written
    as M.dat (save_vector3d order) and material.dat (layer 1)."""
    from mms_gpu import workloads as wl
    dims = (20, 12, 4)
    args = ["dummy.png", "3", "-1", "1e-14", "1e-7", "6e-8", "2e-8", *map(str, dims), "1", "1", "0", "0", "1",
            "sim", "2010", str(ic)]
    Mlines, matlines = _cli_files(tmp_path, args)
    mask = np.zeros(dims[1::-1], np.uint8)
    mask[1:-1, 1:-1] = 1
    mat, M = wl.reference_initial_condition(mask, dims[2], ic, 2010)
    assert Mlines == _vec_lines(M, dims)
    assert matlines == _mat_lines(mat, dims)


def test_cli_front_end_png_mask(tmp_path):
    """A PNG mask replaces xdim, ydim; grey == 0 is magnetic, the image's
    bottom row is y = 0."""
    import sys
    sys.path.insert(0, os.path.dirname(__file__))
    from mms_gpu import workloads as wl
    from test_gpu_cli import _png_rgba
    mask = np.loadtxt(os.path.join(ROOT, "tests", "golden", "nanowire4_mask.txt"), dtype=np.uint8)
    (tmp_path / "nw.png").write_bytes(_png_rgba(mask))
    Mlines, matlines = _cli_files(tmp_path, ["nw.png", "3", "-1", "1e-13", "2.1e-07", "1.3e-07", "1.5e-08", "8",
                                             "8", "3", "1", "1", "0", "0", "1", "nw", "1", "0"])
    mat, M = wl.layered_single_domain(mask, 3)
    assert Mlines == _vec_lines(M, (42, 26, 3))
    assert matlines == _mat_lines(mat, (42, 26, 3))


def test_cli_rejects_bad_arguments(tmp_path):
    import subprocess
    assert os.path.exists(CLI)
    r = subprocess.run([CLI, "none", "3", "1e-14", "1e-14", "1e-7", "1e-7", "1e-8", "8", "8", "3", "1", "1", "0",
                        "0", "1", "x", "1", "7"], cwd=str(tmp_path), capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "Unknown Initial Condition" in r.stderr
    (tmp_path / "bad.png").write_bytes(b"not a png")
    r = subprocess.run([CLI, "bad.png"], cwd=str(tmp_path), capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "not a PNG" in r.stderr


def test_committed_traffic_summary_serves_the_default_bench_line():
    """bench.py's roofline.traffic for the default form reads the committed
    rocprofv3 summary: every kernel class it prices has a line keyed by the
    workload name the bench asks for."""
    import importlib.util
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    workload = "c3_leaf8_fp16x2_f16x2near"
    for kernel in ("stage", "div", "m2l", "p2p", "upward", "downward", "m2m"):
        t = bench.pmc_traffic(kernel, workload)
        assert t is not None and t > 0, kernel
