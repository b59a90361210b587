"""Batched many-film stepping against the same films stepped one by one."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
F = np.float32


def _say(*a):
    print("[gpu-test]", *a, flush=True)


def _films():
    from mms_gpu import workloads as wl
    out = []
    for k, (side, kind) in enumerate([(20, "disc"), (36, "U"), (40, "square"), (50, "U"), (70, "disc"),
                                       (90, "square")]):
        mask = wl.film_mask(kind, side)
        mat = np.broadcast_to(mask, (3, side, side)).ravel().copy()
        M = wl._finish(mat, wl.random_unit(77, 100 + k, side * side * 3))
        hx = float(F((-50 + 20 * k) * 1e-3) / wl.MU_0)
        out.append(wl.Workload(f"f{k}", side, side, 3, mat, M, dt=1e-13, steps=10, hext=(hx, 0.0, 0.0),
                               solver="direct" if side * side * 3 < 4096 else "fmm", P=4))
    return out


@pytest.mark.parametrize("form", ["fp32_leaf4", "fp16x2_leaf8"])
@pytest.mark.parametrize("use_graph", [True, False])
def test_batch_matches_single_films(mg, use_graph, form):
    from mms_gpu import workloads as wl
    prec, leaf = {"fp32_leaf4": (mg.M2L_FP32, 4), "fp16x2_leaf8": (mg.M2L_FP16X2, 8)}[form]
    films = _films()
    mat = np.concatenate([f.material for f in films])
    M0 = np.concatenate([f.M for f in films])
    kw = dict(mu_0=wl.MU_0, Ms=wl.MS, Aexch=wl.AEXCH, alfa=wl.ALFA, gamma=wl.GAMMA)
    _say("batch of", len(films), "films, graph", use_graph, form)
    with mg.FilmBatch(films, 3, *films[0].d, solver=mg.SOLVER_FMM, P=4, m2l_precision=prec, leaf_width=leaf,
                      material=mat, **kw) as b:
        info = b.info()
        _say(info)
        assert info["direct_films"] == 2 and info["fmm_films"] == 4 and info["fmm_size_classes"] == 2
        b.set_m(M0)
        b.steps(10, F(1e-13), use_graph=use_graph)
        Mb = b.get_m()
        offsets = b.offsets
    worst = 0.0
    for i, f in enumerate(films):
        solver = mg.SOLVER_DIRECT if f.solver == "direct" else mg.SOLVER_FMM
        with mg.Film(*f.dims, *f.d, solver=solver, P=4, m2l_precision=prec, leaf_width=leaf,
                     external=mg.EXT_CONST, hext=f.hext, material=f.material, **kw) as s:
            s.set_m(f.M)
            s.steps(10, F(1e-13), use_graph=False)
            Ms = s.get_m()
        d = np.abs(Mb[3 * offsets[i]:3 * offsets[i + 1]] - Ms).max() / float(wl.MS)
        worst = max(worst, d)
        _say(f"  film {f.name} {f.dims} {f.solver}: max |dm| {d:.2e}")
    assert worst <= 1e-5


def test_batch_autotune_and_sweep_shape(mg):
    """The 512-film sweep builds, autotunes and steps; magnetisation stays unit."""
    from mms_gpu import workloads as wl
    films = wl.c5_films(nfilm=64)
    mat = np.concatenate([f.material for f in films])
    M0 = np.concatenate([f.M for f in films])
    with mg.FilmBatch(films, 3, *films[0].d, solver=mg.SOLVER_FMM, P=4, material=mat, mu_0=wl.MU_0, Ms=wl.MS,
                      Aexch=wl.AEXCH, alfa=wl.ALFA, gamma=wl.GAMMA) as b:
        b.set_m(M0)
        rep = b.autotune(reps=2)
        _say(rep)
        b.steps(5, F(1e-13))
        M = b.get_m().reshape(-1, 3)
    mag = np.linalg.norm(M[mat == 1].astype(float), axis=1) / float(wl.MS)
    assert np.abs(mag - 1).max() < 1e-6
    assert np.all(M[mat == 0] == 0)


def test_batch_direct_table_ragged_films(mg, orc):
    """Small films of every shape take the direct sum from per-film 1/r
    tables (widths 1, 3, 5, 20, 37 and 64: rows padded to a multiple of 4 with
    zero charges; a 1-cell-wide film reads past its table rows); anisotropic
    cells (dx, dy, dz all different) enter through the table.  Each film's
    potential is within 1e-6 of the FP64 oracle sum."""
    dims = [(1, 40), (37, 11), (5, 50), (64, 20), (3, 3), (20, 1)]
    zd = 3
    d = (F(5e-9), F(4e-9), F(3e-9))
    rng = np.random.default_rng(8)
    sizes = [x * y * zd for x, y in dims]
    q = (rng.standard_normal(sum(sizes)) * 1e14).astype(F)
    with mg.FilmBatch([(x, y, (0.0, 0.0, 0.0)) for x, y in dims], zd, *d, solver=mg.SOLVER_FMM, P=4) as b:
        assert b.info()["direct_films"] == len(dims)
        pot = b.potential(q)
    off = 0
    for (x, y), m in zip(dims, sizes):
        ref = orc.potential_exact_f64(q[off:off + m], (x, y, zd), d)
        err = np.linalg.norm(pot[off:off + m] - ref) / np.linalg.norm(ref)
        _say(f"  film {x}x{y}x{zd}: rel L2 vs FP64 {err:.2e}")
        assert np.all(np.isfinite(pot[off:off + m]))
        assert err < 1e-6
        off += m


def test_batch_direct_sum_without_tables(mg, orc):
    """A direct-sum batch whose largest film's 1/r table exceeds the LDS
    budget (100 x 100 x 3 cells: a 120 KB table) takes the per-pair direct
    sum for every film; each potential is within 1e-5 of the FP64 oracle sum."""
    dims = [(100, 100), (7, 9), (33, 2)]
    zd = 3
    d = (F(5e-9), F(5e-9), F(5e-9))
    rng = np.random.default_rng(12)
    sizes = [x * y * zd for x, y in dims]
    q = (rng.standard_normal(sum(sizes)) * 1e14).astype(F)
    with mg.FilmBatch([(x, y, (0.0, 0.0, 0.0)) for x, y in dims], zd, *d, solver=mg.SOLVER_DIRECT, P=4) as b:
        assert b.info()["direct_films"] == len(dims)
        pot = b.potential(q)
    off = 0
    for (x, y), m in zip(dims, sizes):
        ref = orc.potential_exact_f64(q[off:off + m], (x, y, zd), d)
        err = np.linalg.norm(pot[off:off + m] - ref) / np.linalg.norm(ref)
        _say(f"  film {x}x{y}x{zd} (no table): rel L2 vs FP64 {err:.2e}")
        assert np.all(np.isfinite(pot[off:off + m]))
        assert err < 1e-5
        off += m
