"""rk4_step on a ragged film (non-material cells) and the time loop's
policy branches against the oracle: the field carried from the previous
step into k1 under a time-dependent field, energy-increase rejection with the
50-rejection abort, and the low-torque stop (ode_functions.cpp:76-148,
226-341)."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
F = np.float32
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _say(*a):
    print("[gpu-test]", *a, flush=True)


def _nanowire_random():
    """The reference's nanowire4 mask (layer 1 magnetic, layers 0 and 2 not)
    with a seeded random state on the magnetic cells."""
    from mms_gpu import workloads as wl
    mask = np.loadtxt(os.path.join(GOLD, "nanowire4_mask.txt"), dtype=np.uint8)
    mat, _ = wl.layered_single_domain(mask, 3)
    n = mat.size
    M = (wl.random_unit(2010, 7, n) * float(wl.MS)).astype(F)
    M = (M * mat[:, None]).astype(F).ravel()
    d = (F(F(2.1e-7) / 42), F(F(1.3e-7) / 26), F(F(1.5e-8) / 3))
    return mat, M, (42, 26, 3), d


@pytest.mark.parametrize("external", [0, 1])
def test_rk4_step_ragged_every_output(mg, orc, external):
    """M2, H, slope and this_slope on every cell of a ragged film.  The
    caller's M2 and this_slope carry garbage on the non-material cells: the
    reference leaves both untouched there (its stage fields then see that M2)
    and zeroes the slope."""
    from mms_gpu import workloads as wl
    mat, M, dims, d = _nanowire_random()
    n = mat.size
    rng = np.random.default_rng(11)
    off = np.repeat(mat == 0, 3)
    M2_in = np.zeros(3 * n, F)
    M2_in[off] = (rng.standard_normal(off.sum()) * 0.3 * float(wl.MS)).astype(F)
    ts_in = np.zeros(3 * n, F)
    ts_in[off] = rng.standard_normal(off.sum()).astype(F) * F(1e9)
    t = F(2e-10)
    _say("rk4_step ragged nanowire4, external", external)
    H0, _, _ = orc.hfield(M, dims, d, 0, wl.MU_0, wl.MS, wl.AEXCH, external=external, t=t)
    M2o, Ho, so, tso = orc.rk4_step_full(M, H0, mat, dims, d, F(1e-13), 0, wl.MU_0, wl.MS, wl.AEXCH, wl.ALFA,
                                         wl.GAMMA, external=external, t=t, combine64=1, M2_in=M2_in,
                                         this_slope_in=ts_in)
    M2 = M2_in.copy()
    H = H0.copy()
    slope = np.full(3 * n, F(7.0))
    ts = ts_in.copy()
    st, t2 = mg.rk4_step(t, F(1e-13), mat, M, M2, H, slope, ts, None, None, *dims, *d, 0, wl.MU_0, wl.MS,
                         wl.AEXCH, wl.ALFA, wl.GAMMA, 1, 1, external, 0, 1, 1, 0)
    assert st == 0, mg.last_error()
    on = ~off
    # non-material cells: bitwise the caller's M2 and this_slope, slope 0
    assert np.array_equal(M2[off], M2_in[off])
    assert np.array_equal(ts[off], ts_in[off])
    assert not slope[off].any()
    # material cells against the oracle (FP64 combine in both)
    assert np.abs(M2[on] - M2o[on]).max() <= 2e-7 * float(wl.MS)
    assert np.abs(H - Ho).max() <= 3e-6 * np.abs(Ho).max()
    assert np.abs(slope - so).max() <= 3e-6 * np.abs(so).max()
    assert np.abs(ts - tso).max() <= 3e-6 * np.abs(tso).max()


def _oracle_time_loop(orc, mat, M, finaltime, timestep, dims, d, external, gamma, max_steps=10 ** 6, demag=1,
                      exchange=1):
    """The reference time loop (ode_functions.cpp:228-341) over the oracle's
    Hfield / rk4_step (FP64 combine, as the device) / energy: returns the
    accepted rows (tindex, t, dt, E2, torque) and the panic rows."""
    from mms_gpu import workloads as wl
    kw = dict(P=0, mu_0=wl.MU_0, Ms=wl.MS, A=wl.AEXCH, demag=demag, exchange=exchange)
    t = F(0)
    dt = F(timestep)
    H, _, _ = orc.hfield(M, dims, d, external=external, t=t, **kw)
    M2 = np.zeros_like(M)
    rows, panic = [], []
    tindex, errs = 0, 0
    while t <= finaltime and len(rows) < max_steps:
        E, _ = orc.energy_torque(mat, M, H, wl.MU_0, wl.MS, d)
        M2, _, _, _ = orc.rk4_step_full(M, H, mat, dims, d, dt, 0, wl.MU_0, wl.MS, wl.AEXCH, wl.ALFA, gamma,
                                        demag=demag, exchange=exchange, external=external, t=t, combine64=1,
                                        M2_in=M2)
        H2, _, _ = orc.hfield(M2, dims, d, external=external, t=t, **kw)
        E2, torque = orc.energy_torque(mat, M2, H2, wl.MU_0, wl.MS, d)
        if E2 - E > 0.1:
            dt = F(dt / F(2.5))
            panic.append((tindex, float(t), float(dt)))
            errs += 1
            if errs > 50:
                break
        else:
            errs = 0
            t = F(t + dt)
            dt = F(F(1e-14) / torque / F(2))
            tindex += 1
            M, H = M2.copy(), H2
            rows.append((tindex, float(t), float(dt), float(E2), float(torque)))
            if torque < 1e-3:
                break
    return rows, panic


def _read_rows(path):
    return [list(map(float, l.split(","))) for l in open(path) if l.strip()]


def test_time_marching_ramp_field_follows_oracle_loop(mg, orc, tmp_path):
    """With the reference's time-dependent field (0.3 T ramp), k1 of each step
    uses the post-step field of the previous one, evaluated at that step's
    (unadvanced) time (ode_functions.cpp:246-247,267); a field recomputed at
    the new time would shift k1 by one dt of the ramp."""
    from mms_gpu import workloads as wl
    mat, M, dims, d = _nanowire_random()
    sim = str(tmp_path / "ramp")
    _say("time_marching ramp field vs oracle loop")
    Mg = M.copy()
    st = mg.time_marching(mat, Mg, F(1.1e-13), F(1e-13), *dims, *d, 3, wl.MU_0, wl.MS, wl.AEXCH, wl.ALFA,
                          wl.GAMMA, 1, 1, 1, 0, 1, sim, 0)
    assert st == 0, mg.last_error()
    rows = _read_rows(os.path.join(sim, "dynamics.dat"))
    ref, panic = _oracle_time_loop(orc, mat, M, F(1.1e-13), F(1e-13), dims, d, 1, wl.GAMMA)
    assert not panic
    assert len(rows) == len(ref) >= 5
    for r, o in zip(rows, ref):
        assert int(r[0]) == o[0]
        assert abs(r[1] - o[1]) <= 1e-5 * o[1]
        assert abs(r[3] - o[3]) <= 3e-5 * abs(o[3]) + 1e-3
        assert abs(r[8] - o[4]) <= 3e-5 * o[4]


def test_time_marching_energy_rejection_and_abort(mg, orc, tmp_path):
    """Energy-increase rejection (ode_functions.cpp:282-294): a uniform state
    at twice Ms along the ramp field, with the ramp as the only field, loses
    half its (negative) Zeeman energy on renormalisation whatever the step,
    so every step is rejected, dt shrinks by 2.5 each time, and the loop stops
    after the 51st rejection with the abort line in panic.dat."""
    from mms_gpu import workloads as wl
    mat, _, dims, d = _nanowire_random()
    n = mat.size
    M = np.zeros((n, 3), F)
    M[mat == 1, 0] = F(2.0) * wl.MS
    M = M.ravel()
    # the oracle agrees that the first attempt is rejected
    _, panic_o = _oracle_time_loop(orc, mat, M, F(1e-12), F(1e-13), dims, d, 1, wl.GAMMA, max_steps=1,
                                   demag=0, exchange=0)
    assert len(panic_o) == 51
    sim = str(tmp_path / "reject")
    _say("time_marching forced rejection")
    Mg = M.copy()
    st = mg.time_marching(mat, Mg, F(1e-12), F(1e-13), *dims, *d, 3, wl.MU_0, wl.MS, wl.AEXCH, wl.ALFA,
                          wl.GAMMA, 0, 0, 1, 0, 1, sim, 0)
    assert st == 0, mg.last_error()
    assert _read_rows(os.path.join(sim, "dynamics.dat")) == []
    lines = [l for l in open(os.path.join(sim, "panic.dat")) if l.strip()]
    assert len(lines) == 52
    head = "PANIC!! energy is increasing, so halving the step. "
    dt = F(1e-13)
    for k, l in enumerate(lines[:51]):
        assert l.startswith(head)
        f = [float(v) for v in l[len(head):].split(",")]
        dt = F(dt / F(2.5))
        assert int(f[0]) == 0 and f[1] == 0.0
        assert abs(f[2] - float(dt)) <= 1e-5 * float(dt), k
    assert lines[51].strip() == "PANIC!! Stuck in energy violation errors. Terminating the simulation."
    # no step accepted: M is returned as it was passed in
    assert np.array_equal(Mg, M)


def test_time_marching_torque_stop(mg, tmp_path):
    """Low-torque stop (ode_functions.cpp:337-340): a uniform state with only
    exchange (which stays parallel to M) has zero torque, so the loop accepts
    one step and stops."""
    from mms_gpu import workloads as wl
    mat, _, dims, d = _nanowire_random()
    n = mat.size
    M = np.zeros((n, 3), F)
    M[mat == 1, 1] = wl.MS
    M = M.ravel()
    sim = str(tmp_path / "torque")
    _say("time_marching torque stop")
    st = mg.time_marching(mat, M, F(1e-9), F(1e-13), *dims, *d, 3, wl.MU_0, wl.MS, wl.AEXCH, wl.ALFA,
                          wl.GAMMA, 0, 1, 0, 0, 1, sim, 0)
    assert st == 0, mg.last_error()
    rows = _read_rows(os.path.join(sim, "dynamics.dat"))
    assert len(rows) == 1
    assert int(rows[0][0]) == 1 and rows[0][8] < 1e-3
    assert abs(rows[0][1] - 1e-13) <= 1e-19
    assert [l for l in open(os.path.join(sim, "panic.dat")) if l.strip()] == []


def test_mdynamics_dump_order(mg, tmp_path):
    """Mdynamics.dat holds the full M after every accepted step, z outermost,
    then x, then y (helper_functions.cpp:107-123), '%g %g %g \\n' per cell."""
    from mms_gpu import workloads as wl
    mat, M, dims, d = _nanowire_random()
    xd, yd, zd = dims
    sim = str(tmp_path / "order")
    Mg = M.copy()
    st = mg.time_marching(mat, Mg, F(1.5e-13), F(1e-13), *dims, *d, 3, wl.MU_0, wl.MS, wl.AEXCH, wl.ALFA,
                          wl.GAMMA, 1, 1, 0, 0, 1, sim, 0)
    assert st == 0, mg.last_error()
    rows = _read_rows(os.path.join(sim, "dynamics.dat"))
    lines = open(os.path.join(sim, "Mdynamics.dat")).read().split("\n")
    n = xd * yd * zd
    assert len(lines) == len(rows) * n + 1 and lines[-1] == ""
    last = Mg.reshape(zd, yd, xd, 3)
    expect = []
    for z in range(zd):
        for x in range(xd):
            for y in range(yd):
                v = last[z, y, x]
                expect.append("%s %s %s " % tuple(_g(float(c)) for c in v))
    assert lines[(len(rows) - 1) * n:len(rows) * n] == expect


def _g(v):
    """C's %g"""
    return "%g" % v
