"""GPU field assembly, RK4 stepping and time loop against the oracle, the
golden trajectory and the reference's own logged trajectory."""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
F = np.float32
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _say(*a):
    print("[gpu-test]", *a, flush=True)


def _rel(a, b):
    return np.linalg.norm(np.asarray(a, float) - b) / np.linalg.norm(b)


def _nanowire():
    from mms_gpu import workloads as wl
    mask = np.loadtxt(os.path.join(GOLD, "nanowire4_mask.txt"), dtype=np.uint8)
    mat, M = wl.layered_single_domain(mask, 3)
    d = (F(F(2.1e-7) / 42), F(F(1.3e-7) / 26), F(F(1.5e-8) / 3))
    return mat, M, (42, 26, 3), d


def test_hfield_matches_oracle(mg, orc):
    from mms_gpu import workloads as wl
    mat, M, dims, d = _nanowire()
    n = int(np.prod(dims))
    for demag, exch, ext, t in [(1, 1, 0, 0.0), (0, 1, 1, 3e-10), (1, 1, 1, 2e-9), (1, 0, 0, 0.0)]:
        _say("Hfield", demag, exch, ext, t)
        H = np.zeros(3 * n, F)
        ch = np.zeros(n, F)
        pot = np.zeros(n, F)
        st = mg.Hfield(M, H, ch, pot, F(t), *dims, *d, 3, wl.MU_0, wl.MS, wl.AEXCH, demag, exch, ext, 0, 1, 0)
        assert st == 0, mg.last_error()
        Ho, cho, poto = orc.hfield(M, dims, d, 3, wl.MU_0, wl.MS, wl.AEXCH, demag=demag, exchange=exch,
                                   external=ext, t=F(t))
        if demag:
            assert np.array_equal(ch, cho)          # div: same float expression
            assert _rel(pot, poto) < 3e-6
        tol = 3e-6 if demag else 1e-6
        assert np.abs(H - Ho).max() <= tol * np.abs(Ho).max(), (demag, exch, ext)


@pytest.mark.parametrize("dims", [(36, 20, 3), (32, 32, 4), (64, 8, 2), (4, 4, 1)])
def test_charge_four_wide_matches_oracle(mg, orc, dims):
    """Widths divisible by 4 take the four-cells-per-thread charge kernel:
    bitwise equal to divergence_3d (vector_functions.cpp:8-28), boundary
    cells (x = 0, x = xd - 1, y and z faces) included."""
    from mms_gpu import workloads as wl
    _say("charge x4", dims)
    n = int(np.prod(dims))
    rng = np.random.default_rng(sum(dims))
    M = (rng.standard_normal(3 * n) * wl.MS).astype(F)
    M[rng.random(3 * n) < 0.2] = 0.0
    d = (F(5e-9), F(4e-9), F(3e-9))
    H = np.zeros(3 * n, F)
    ch = np.zeros(n, F)
    pot = np.zeros(n, F)
    st = mg.Hfield(M, H, ch, pot, F(0.0), *dims, *d, 3, wl.MU_0, wl.MS, wl.AEXCH, 1, 0, 0, 0, 1, 0)
    assert st == 0, mg.last_error()
    assert np.array_equal(ch, orc.divergence_3d(M, dims, d))


@pytest.mark.parametrize("case", ["c2_dots", "ragged_direct"])
def test_stage_four_wide_matches_one_cell(mg, case, monkeypatch):
    """Widths divisible by 4 run the resident step's stage kernel four cells
    per thread; MMS_GPU_STAGE_ONE_CELL=1 forces the one-cell form.  Both give
    bitwise identical states, energies and torques: on the C2 dots film
    (non-magnetic cells between the dots, constant external field, FMM) and
    on a 36 x 20 x 3 film with a ragged mask and the ramp field (direct sum)."""
    from mms_gpu import workloads as wl
    if case == "c2_dots":
        w = wl.c2_dots()
        kw = dict(solver=mg.SOLVER_FMM, P=4, m2l_precision=mg.M2L_FP32, leaf_width=4,
                  external=mg.EXT_CONST, hext=(7957.7, 0.0, 0.0))
        dims, d, mat, M, dt = w.dims, w.d, w.material, w.M, 1e-13
    else:
        dims = (36, 20, 3)
        rng = np.random.default_rng(9)
        mat = (rng.random(int(np.prod(dims))) < 0.8).astype(np.uint8)
        M = wl.random_unit(2010, 3, int(np.prod(dims))).astype(F).ravel() * F(wl.MS)
        M = (M.reshape(-1, 3) * mat[:, None]).astype(F).ravel()
        d = (F(5e-9), F(5e-9), F(5e-9))
        kw = dict(solver=mg.SOLVER_DIRECT, external=mg.EXT_RAMP)
        dt = 1e-13
    out = []
    for one in ("0", "1"):
        _say("stage four-wide vs one-cell", case, one)
        monkeypatch.setenv("MMS_GPU_STAGE_ONE_CELL", one)
        with mg.Film(*dims, *d, Ms=wl.MS, Aexch=wl.AEXCH, mu_0=wl.MU_0, material=mat, **kw) as f:
            f.set_m(M)
            f.steps(3, dt, use_graph=False)
            f.sync()
            out.append((f.get_m(), f.energy_torque()))
    assert np.array_equal(out[0][0], out[1][0])
    assert out[0][1] == out[1][1]


@pytest.mark.parametrize("combine64", [1, 0])
def test_rk4_step_matches_oracle(mg, orc, combine64):
    from mms_gpu import workloads as wl
    w = wl.c1_square()
    _say("rk4_step C1 combine64 =", combine64)
    H0, _, _ = orc.hfield(w.M, w.dims, w.d, 0, wl.MU_0, wl.MS, wl.AEXCH)
    M2o, Ho = orc.rk4_step(w.M, H0, w.material, w.dims, w.d, w.dt, 0, wl.MU_0, wl.MS, wl.AEXCH, wl.ALFA,
                           wl.GAMMA, combine64=combine64)
    M2 = np.zeros_like(w.M)
    H = H0.copy()
    slope = np.zeros_like(w.M)
    k4 = np.zeros_like(w.M)
    st, t2 = mg.rk4_step(F(1e-12), w.dt, w.material, w.M, M2, H, slope, k4, None, None, *w.dims, *w.d, 0,
                         wl.MU_0, wl.MS, wl.AEXCH, wl.ALFA, wl.GAMMA, 1, 1, 0, 0, 1, 1, 0)
    assert st == 0, mg.last_error()
    assert t2 == F(F(1e-12) + w.dt)
    # the device always combines in FP64; against the FP32 restatement the gap
    # is the FP32 rounding of the reference's combine
    tol = 2e-7 if combine64 else 2e-6
    assert np.abs(M2 - M2o).max() <= tol * float(wl.MS)
    assert np.abs(H - Ho).max() <= 3e-6 * np.abs(Ho).max()
    Mm = M2.reshape(-1, 3)
    assert np.allclose(np.linalg.norm(Mm.astype(float), axis=1), float(wl.MS), rtol=1e-6)


def test_resident_steps_match_dropin_and_graph(mg, orc):
    from mms_gpu import workloads as wl
    w = wl.c1_square()
    kw = dict(solver=mg.SOLVER_DIRECT, mu_0=wl.MU_0, Ms=wl.MS, Aexch=wl.AEXCH, alfa=wl.ALFA, gamma=wl.GAMMA,
              material=w.material)
    # drop-in sequence: Hfield once, then rk4_step; each later step's k1 is the
    # field of the new M (the reference time loop's post-step Hfield)
    M = w.M.copy()
    H = np.zeros_like(M)
    assert mg.Hfield(M, H, None, None, F(0), *w.dims, *w.d, 0, wl.MU_0, wl.MS, wl.AEXCH, 1, 1, 0, 0, 1, 0) == 0
    for _ in range(5):
        M2 = np.zeros_like(M)
        st, _ = mg.rk4_step(F(0), w.dt, w.material, M, M2, H, None, None, None, None, *w.dims, *w.d, 0,
                            wl.MU_0, wl.MS, wl.AEXCH, wl.ALFA, wl.GAMMA, 1, 1, 0, 0, 1, 1, 0)
        assert st == 0
        M = M2
        assert mg.Hfield(M, H, None, None, F(0), *w.dims, *w.d, 0, wl.MU_0, wl.MS, wl.AEXCH, 1, 1, 0, 0, 1,
                         0) == 0
    with mg.Film(*w.dims, *w.d, **kw) as fa, mg.Film(*w.dims, *w.d, **kw) as fb:
        fa.set_m(w.M)
        fb.set_m(w.M)
        fa.steps(5, w.dt, use_graph=False)
        ea = fa.get_m()
        # the resident FP64 base carries extra bits the drop-in re-rounds each step
        assert np.abs(ea - M).max() <= 1e-6 * float(wl.MS)
        # graph replays, with eager work and synchronisation between replays,
        # must equal eager stepping bit for bit
        for rep in range(3):
            fb.steps(5 if rep == 0 else 2, w.dt, use_graph=True)
            fb.sync()
            if rep > 0:
                fa.steps(2, w.dt, use_graph=False)
            assert np.array_equal(fb.get_m(), fa.get_m()), rep
            _ = fb.energy_torque()


def test_c1_trajectory_1000_steps(mg):
    """32x32x3 film, random m, direct sum, 1000 RK4 steps of 1e-13 s: within
    1e-4 (|dm|, m = M/Ms) of the reference path (FP32 restatement)."""
    from mms_gpu import workloads as wl
    g = np.load(os.path.join(GOLD, "c1_traj.npz"))
    w = wl.c1_square()
    assert np.array_equal(g["M0"], w.M)
    _say("C1 trajectory, 1000 steps")
    with mg.Film(*w.dims, *w.d, solver=mg.SOLVER_DIRECT, mu_0=wl.MU_0, Ms=wl.MS, Aexch=wl.AEXCH, alfa=wl.ALFA,
                 gamma=wl.GAMMA, material=w.material) as f:
        f.set_m(w.M)
        f.steps(int(g["nsteps"][0]), w.dt, use_graph=True)
        M = f.get_m()
        E, T = f.energy_torque()
    dm32 = np.abs(M - g["M_f32"]).max() / float(wl.MS)
    dm64 = np.abs(M - g["M_f64c"]).max() / float(wl.MS)
    _say(f"  max|dm| vs FP32 reference path {dm32:.2e}, vs FP64-combine restatement {dm64:.2e}")
    assert dm32 <= 1e-4
    assert dm64 <= 1e-5
    Eo, To = g["Etorque_f32"][-1]
    assert abs(E - Eo) <= 1e-4 * abs(Eo) and abs(T - To) <= 1e-3 * To


def test_nanowire4_matches_logged_reference_trajectory(mg):
    """Fixed-step (1e-13 s) run of the reference's nanowire4 geometry: energy,
    <M> and torque after every step follow the reference's own log over 1000
    steps (printed to 6 significant digits)."""
    from mms_gpu import workloads as wl
    k = json.load(open(os.path.join(GOLD, "log_kats.json")))["nanowire4_fixed_1e-13"]
    mat, M0, dims, d = _nanowire()
    recs = k["records"]
    _say("nanowire4 logged trajectory", len(recs), "steps")
    worst = {"E2": 0.0, "my": 0.0, "mz": 0.0, "T": 0.0}
    with mg.Film(*dims, *d, solver=mg.SOLVER_DIRECT, mu_0=wl.MU_0, Ms=wl.MS, Aexch=wl.AEXCH, alfa=wl.ALFA,
                 gamma=F(k["gamma"]), material=mat) as f:
        f.set_m(M0)
        done = 0
        for rec in recs:
            tindex = int(rec[1])
            f.steps(tindex - done, F(k["timestep"]), use_graph=True)
            done = tindex
            if tindex not in (1, 2, 3, 10, 50, 100, 250, 500, 750, 1000):
                continue
            E2, T = f.energy_torque()
            m = f.get_m().reshape(-1, 3)[mat == 1].astype(float) / float(wl.MS)
            mx, my, mz = m.mean(0)
            _, _, _, _, E2r, mxr, myr, mzr, _, Tr = rec
            worst["E2"] = max(worst["E2"], abs(E2 - E2r))
            worst["my"] = max(worst["my"], abs(my - myr) / abs(myr))
            worst["mz"] = max(worst["mz"], abs(mz - mzr) / abs(mzr))
            worst["T"] = max(worst["T"], abs(T - Tr) / Tr)
    _say("  worst deviations", worst)
    assert worst["E2"] <= 2e-3          # eV, log prints 373.xxx
    assert worst["my"] <= 5e-5 and worst["mz"] <= 5e-5
    assert worst["T"] <= 5e-5


def test_time_marching_writes_reference_logs(mg, tmp_path):
    from mms_gpu import workloads as wl
    k = json.load(open(os.path.join(GOLD, "log_kats.json")))["nanowire4_fixed_1e-13"]
    mat, M0, dims, d = _nanowire()
    M = M0.copy()
    sim = str(tmp_path / "sim")
    _say("time_marching nanowire4")
    st = mg.time_marching(mat, M, F(3e-12), F(1e-13), *dims, *d, 3, wl.MU_0, wl.MS, wl.AEXCH, wl.ALFA,
                          F(k["gamma"]), 1, 1, 0, 0, 1, sim, 0)
    assert st == 0, mg.last_error()
    rows = [list(map(float, l.split(","))) for l in open(os.path.join(sim, "dynamics.dat")) if l.strip()]
    assert len(rows) >= 2
    r1 = rows[0]
    ref1 = k["records"][0]
    assert int(r1[0]) == 1 and abs(r1[1] - 1e-13) < 1e-20
    assert abs(r1[3] - ref1[4]) <= 1e-3 and abs(r1[8] - ref1[9]) <= 1e-5 * ref1[9]
    # accepted-step rule of the current code: dt = 1e-14 / torque / 2
    assert abs(r1[2] - 1e-14 / r1[8] / 2) <= 1e-5 * r1[2]
    nlines = sum(1 for _ in open(os.path.join(sim, "Mdynamics.dat")))
    assert nlines == len(rows) * int(np.prod(dims))
    assert np.abs(np.linalg.norm(M.reshape(-1, 3)[mat == 1], axis=1) - float(wl.MS)).max() < 1.0


def test_energy_non_increasing_c1(mg):
    from mms_gpu import workloads as wl
    w = wl.c1_square()
    with mg.Film(*w.dims, *w.d, mu_0=wl.MU_0, Ms=wl.MS, Aexch=wl.AEXCH, alfa=wl.ALFA, gamma=wl.GAMMA,
                 material=w.material) as f:
        f.set_m(w.M)
        Es = []
        for _ in range(20):
            Es.append(f.energy_torque()[0])
            f.steps(10, w.dt)
    assert all(Es[i + 1] <= Es[i] + 1e-6 * abs(Es[i]) for i in range(len(Es) - 1)), Es
