"""The measured configurations at their real size on the GPU.

* 1024 x 1024 x 4 domain film (4,194,304 cells, FMM P = 6, leaf 8): every
  operand form inside the workload's precision split, stepped 250 RK4 steps
  of 5e-14 s, demag field within 1e-3 (relative L2) of an FP64 direct sum on
  8,192 seeded probes.
* 512-film sweep (16^2..96^2 x 3; direct sum below 4,096 cells, FMM P = 4
  above): the batch's potential on every 32nd film against the FP64 direct
  sum, and every 32nd direct-sum film's trajectory against the oracle's RK4.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
F = np.float32


def _say(*a):
    print("[gpu-test]", *a, flush=True)


@pytest.mark.parametrize("form", ["fp16x2_f16x2near", "fp16x2_bf16x3near", "bf16x3", "fp32"])
def test_c3_full_size_field_error_along_run(mg, form):
    from mms_gpu import workloads as wl
    prec = {"fp32": mg.M2L_FP32, "bf16x3": mg.M2L_BF16X3, "fp16x2_bf16x3near": mg.M2L_FP16X2_BF16X3NEAR,
            "fp16x2_f16x2near": mg.M2L_FP16X2_FP16X2NEAR}[form]
    w = wl.c3_domains()
    assert w.dims == (1024, 1024, 4)
    _say("C3 1024x1024x4", form)
    with mg.Film(*w.dims, *w.d, solver=mg.SOLVER_FMM, P=6, m2l_precision=prec, leaf_width=8, mu_0=wl.MU_0,
                 Ms=wl.MS, Aexch=wl.AEXCH, alfa=wl.ALFA, gamma=wl.GAMMA, material=w.material) as f:
        f.set_m(w.M)
        errs = []
        for target in (0, 250):
            f.steps(target - (0 if not errs else errs[-1][0]), w.dt, use_graph=True)
            f.sync()
            l2, mx, npr = wl.demag_field_error(f, w, 8192)
            _say(f"  step {target}: field rel L2 {l2:.2e}, rel max {mx:.2e} ({npr} probes)")
            errs.append((target, l2, mx))
            assert npr == 8192
        m = f.get_m().reshape(-1, 3).astype(np.float64)
    assert max(e[1] for e in errs) <= 1e-3
    assert np.abs(np.linalg.norm(m, axis=1) / float(wl.MS) - 1).max() < 1e-6


def _c5_batch(mg, wl, films, prec=None):
    mat = np.concatenate([f.material for f in films])
    return mg.FilmBatch(films, 3, *films[0].d, solver=mg.SOLVER_FMM, P=4,
                        m2l_precision=mg.M2L_FP32 if prec is None else prec, material=mat, mu_0=wl.MU_0,
                        Ms=wl.MS, Aexch=wl.AEXCH, alfa=wl.ALFA, gamma=wl.GAMMA)


@pytest.mark.parametrize("form", ["fp32", "fp16x2_bf16x3near"])
def test_c5_sweep_potential_against_fp64(mg, orc, form):
    """All 512 films; the potential of every 32nd film through the batch's
    solvers against the FP64 direct sum: 1e-5 for direct-sum films, the C2
    dots film's FMM tolerance (order-4 expansion, 2e-3) for FMM films."""
    from mms_gpu import workloads as wl
    prec = {"fp32": mg.M2L_FP32, "fp16x2_bf16x3near": mg.M2L_FP16X2_BF16X3NEAR}[form]
    films = wl.c5_films()
    assert len(films) == 512
    ch = np.concatenate([orc.divergence_3d(f.M, f.dims, f.d) for f in films])
    _say("C5 potential, 512 films", form)
    with _c5_batch(mg, wl, films, prec) as b:
        info = b.info()
        _say(info)
        pot = b.potential(ch)
        offs = b.offsets
    kinds = set()
    for i in range(0, 512, 32):
        f = films[i]
        sl = slice(int(offs[i]), int(offs[i + 1]))
        ex = orc.potential_exact_f64(ch[sl], f.dims, f.d)
        rel = np.linalg.norm(pot[sl] - ex) / np.linalg.norm(ex)
        mx = np.abs(pot[sl] - ex).max() / np.abs(ex).max()
        _say(f"  film {i} {f.dims} {f.solver}: rel L2 {rel:.2e}, rel max {mx:.2e}")
        kinds.add(f.solver)
        if f.solver == "direct":
            assert rel <= 1e-5 and mx <= 1e-5
        else:
            assert rel <= 2e-3
    assert kinds == {"direct", "fmm"}


def test_c5_sweep_trajectories_against_oracle(mg, orc):
    """All 512 films stepped 50 RK4 steps of 1e-13 s together; every 32nd
    direct-sum film against the oracle's RK4 (FP64 combine, per-film field
    ramp value) run on that film alone: within 1e-5 of Ms."""
    from mms_gpu import workloads as wl
    films = wl.c5_films()
    M0 = np.concatenate([f.M for f in films])
    nsteps = 50
    _say("C5 trajectories, 512 films x", nsteps, "steps")
    with _c5_batch(mg, wl, films) as b:
        b.set_m(M0)
        b.steps(nsteps, F(1e-13))
        b.sync()
        M = b.get_m()
        offs = b.offsets
    direct = [i for i, f in enumerate(films) if f.solver == "direct"]
    sample = direct[::32] if len(direct) >= 64 else direct[::max(1, len(direct) // 2)]
    assert sample
    worst = 0.0
    for i in sample:
        f = films[i]
        m = f.M.copy()
        kw = dict(P=0, mu_0=wl.MU_0, Ms=wl.MS, A=wl.AEXCH, external=1, hext=f.hext)
        H, _, _ = orc.hfield(m, f.dims, f.d, **kw)
        for _ in range(nsteps):
            m, _ = orc.rk4_step(m, H, f.material, f.dims, f.d, F(1e-13), 0, wl.MU_0, wl.MS, wl.AEXCH, wl.ALFA,
                                wl.GAMMA, external=1, hext=f.hext, combine64=1)
            H, _, _ = orc.hfield(m, f.dims, f.d, **kw)
        d = np.abs(M[3 * int(offs[i]):3 * int(offs[i + 1])] - m).max() / float(wl.MS)
        worst = max(worst, d)
        _say(f"  film {i} {f.dims}: max |dm| {d:.2e}")
    assert worst <= 1e-5
