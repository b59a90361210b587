"""Closed forms and invariants that pin the oracle where no reference output
exists (direct sum, tree-code truncation, LL dynamics, RK4 order)."""
import math

import numpy as np
import pytest

F = np.float32
DX = (F(5e-9), F(5e-9), F(5e-9))


def test_legendre_textbook(orc):
    L = orc.lib()
    x = 0.3
    s = math.sqrt(1 - x * x)
    table = {(0, 0): 1, (1, 0): x, (1, 1): -s, (2, 0): 0.5 * (3 * x * x - 1), (2, 1): -3 * x * s,
             (2, 2): 3 * (1 - x * x), (3, 0): 0.5 * (5 * x ** 3 - 3 * x), (3, 1): -1.5 * (5 * x * x - 1) * s,
             (3, 2): 15 * x * (1 - x * x), (3, 3): -15 * s ** 3, (4, 0): (35 * x ** 4 - 30 * x * x + 3) / 8,
             (4, 1): -2.5 * (7 * x ** 3 - 3 * x) * s, (4, 2): 7.5 * (7 * x * x - 1) * (1 - x * x),
             (4, 3): -105 * x * s ** 3, (4, 4): 105 * (1 - x * x) ** 2}
    for (l, m), v in table.items():
        assert abs(L.or_associated_legendre(l, m, x) - v) < 1e-5 * max(1, abs(v)), (l, m)
    assert math.isnan(L.or_associated_legendre(5, 0, x))   # numerics.cu:100-102


def test_two_charges_closed_form(orc):
    dims = (5, 1, 1)
    q = np.zeros(5, F)
    q[0], q[3] = 2e14, -1e14
    phi = orc.potential_exact_f64(q, dims, DX)
    dV, dx = float(DX[0]) ** 3, float(DX[0])
    expect = [(-1e14) / (3 * dx), 2e14 / dx + (-1e14) / (2 * dx), 2e14 / (2 * dx) + (-1e14) / dx,
              2e14 / (3 * dx), 2e14 / (4 * dx) + (-1e14) / dx]
    assert np.allclose(phi, np.array(expect) * dV, rtol=1e-6)
    f32 = orc.calc_potential_exact(q, dims, DX)
    assert np.allclose(f32, phi, rtol=2e-6)


def test_self_excluded_and_linear(orc):
    dims = (4, 3, 2)
    n = 24
    one = np.zeros(n, F)
    one[7] = 1e14
    assert orc.potential_exact_f64(one, dims, DX)[7] == 0.0
    rng = np.random.default_rng(3)
    a, b = rng.normal(size=n).astype(F) * 1e14, rng.normal(size=n).astype(F) * 1e14
    pa, pb = orc.potential_exact_f64(a, dims, DX), orc.potential_exact_f64(b, dims, DX)
    pab = orc.potential_exact_f64((a + b).astype(F), dims, DX)
    assert np.allclose(pab, pa + pb, rtol=1e-5, atol=1e-9 * np.abs(pa).max())
    assert np.allclose(orc.potential_exact_f64(-a, dims, DX), -pa)


def test_fp64_direct_vs_fft_and_fp32(orc):
    from oracle import fmm_oracle
    rng = np.random.default_rng(5)
    dims, d = (12, 9, 3), (F(5e-9), F(4e-9), F(3e-9))
    q = rng.normal(size=int(np.prod(dims))).astype(F) * 1e14
    brute = orc.potential_exact_f64(q, dims, d)
    fft = fmm_oracle.direct_fft(q.reshape(3, 9, 12).astype(float), ay=float(d[1]) / float(d[0]),
                                az=float(d[2]) / float(d[0])).ravel()
    fft *= float(d[0]) * float(d[1]) * float(d[2]) / float(d[0])
    assert np.abs(fft - brute).max() <= 1e-11 * np.abs(brute).max()
    f32 = orc.calc_potential_exact(q, dims, d)
    assert np.linalg.norm(f32 - brute) / np.linalg.norm(brute) < 2e-6
    probes = np.array([0, 17, 200, 323], np.int64)
    assert np.array_equal(orc.potential_probes_f64(q, dims, d, probes), brute[probes])


def test_treecode_exact_on_4x4(orc):
    """logN = 2: the leaf level is level 2, every far cell is a 1-cell box whose
    multipole is exact, so the tree-code equals the direct sum."""
    rng = np.random.default_rng(9)
    dims = (4, 4, 3)
    q = rng.normal(size=48).astype(F) * 1e14
    tc = orc.fmm_calc(q, dims, DX, 3)
    ex = orc.potential_exact_f64(q, dims, DX)
    assert np.abs(tc - ex).max() <= 2e-6 * np.abs(ex).max()


def test_treecode_p4_overflows_on_5nm_cells(orc):
    """Reference quirk, reproduced: the M2P factor 1/pow(r, l+1) is rounded to
    float (fmm_calc.cu:203); for l = 4 and r = 10 nm it is 1e40 > FLT_MAX, and
    inf times the zero l = 4 moment of a one-cell leaf box is NaN.  With 10 nm
    cells the same charges stay finite."""
    rng = np.random.default_rng(9)
    q = rng.normal(size=48).astype(F) * 1e14
    assert np.isnan(orc.fmm_calc(q, (4, 4, 3), DX, 4)).any()
    d10 = (F(1e-8),) * 3
    tc = orc.fmm_calc(q, (4, 4, 3), d10, 4)
    ex = orc.potential_exact_f64(q, (4, 4, 3), d10)
    assert np.abs(tc - ex).max() <= 2e-6 * np.abs(ex).max()


def _truncation_bound(q, xd, zd, P, dx):
    """sum over every (level, source box, layer) that reaches a target through
    its interaction list of A/(r-a) (a/r)^(P+1) (Greengard-Rokhlin), with A the
    box's absolute charge and a its charged-cell radius (fmm_calc.cu:113-223)."""
    logN = int(round(math.log2(xd)))
    q3 = q.reshape(zd, xd, xd).astype(float)
    bound = np.zeros((zd, xd, xd))
    yy, xx = np.mgrid[0:xd, 0:xd]
    for L in range(2, logN + 1):
        w = 2 ** (logN - L)
        nb = 2 ** L
        for by in range(nb):
            for bx in range(nb):
                cx, cy = w * bx + (w - 1) / 2, w * by + (w - 1) / 2
                pbx, pby = bx // 2, by // 2
                tbx, tby = xx // w, yy // w
                in_il = (np.abs(tbx // 2 - pbx) <= 1) & (np.abs(tby // 2 - pby) <= 1) & \
                        ((np.abs(tbx - bx) > 1) | (np.abs(tby - by) > 1))
                if not in_il.any():
                    continue
                for zc in range(zd):
                    blk = q3[zc, by * w:(by + 1) * w, bx * w:(bx + 1) * w]
                    A = np.abs(blk).sum() * dx ** 3
                    if A == 0:
                        continue
                    sy, sx = np.nonzero(blk)
                    a = np.sqrt(((sx + bx * w - cx) ** 2 + (sy + by * w - cy) ** 2).max()) * dx
                    for zp in range(zd):
                        r = np.sqrt(((xx - cx) ** 2 + (yy - cy) ** 2) * dx * dx + ((zp - zc) * dx) ** 2)
                        bound[zp][in_il] += (A / (r - a) * (a / r) ** (P + 1))[in_il]
    return bound.ravel()


@pytest.mark.parametrize("P", [1, 2, 3, 4])
def test_treecode_within_truncation_bound(orc, P):
    rng = np.random.default_rng(P)
    xd, zd = 16, 3
    q = rng.normal(size=xd * xd * zd).astype(F) * 1e14
    dims = (xd, xd, zd)
    d = (F(1e-8),) * 3   # 10 nm cells keep the P = 4 series finite in FP32
    tc = orc.fmm_calc(q, dims, d, P).astype(float)
    ex = orc.potential_exact_f64(q, dims, d)
    bound = _truncation_bound(q, xd, zd, P, float(d[0]))
    slack = 1e-6 * np.abs(ex).max()
    assert np.all(np.abs(tc - ex) <= bound + slack)


def test_treecode_error_falls_with_P(orc):
    rng = np.random.default_rng(2)
    dims = (32, 32, 3)
    q = rng.normal(size=32 * 32 * 3).astype(F) * 1e14
    ex = orc.potential_exact_f64(q, dims, DX)
    errs = [np.linalg.norm(orc.fmm_calc(q, dims, DX, P) - ex) / np.linalg.norm(ex) for P in range(0, 4)]
    assert all(errs[i + 1] < errs[i] for i in range(3)), errs


def _macrospin(orc, dt, nsteps, theta0=1.0, H=1e5):
    """one cell, constant H along z, demag off (exchange of a lone cell is
    parallel to M and exerts no torque)."""
    M = np.array([math.sin(theta0), 0.0, math.cos(theta0)], F) * F(8.6e5)
    mat = np.ones(1, np.uint8)
    hext = (0.0, 0.0, H)
    d = DX
    Hf, _, _ = orc.hfield(M, (1, 1, 1), d, 0, F(4e-7 * np.pi), F(8.6e5), F(1.3e-11), demag=0, external=1,
                          hext=hext)
    for _ in range(nsteps):
        M, _ = orc.rk4_step(M, Hf, mat, (1, 1, 1), d, F(dt), 0, F(4e-7 * np.pi), F(8.6e5), F(1.3e-11), F(0.5),
                            F(2.211e5), demag=0, external=1, hext=hext, combine64=1)
        Hf, _, _ = orc.hfield(M, (1, 1, 1), d, 0, F(4e-7 * np.pi), F(8.6e5), F(1.3e-11), demag=0, external=1,
                              hext=hext)
    return M.astype(float) / 8.6e5


def test_macrospin_closed_form_and_order(orc):
    """LL with H || z: tan(theta/2) = tan(theta0/2) exp(-alpha gamma H t),
    phi = gamma H t; RK4 global error falls as dt^4."""
    gamma, alpha, H, t_end = 2.211e5, 0.5, 1e5, 2e-10
    th = 2 * math.atan(math.tan(0.5) * math.exp(-alpha * gamma * H * t_end))
    ph = gamma * H * t_end
    exact = np.array([math.sin(th) * math.cos(ph), math.sin(th) * math.sin(ph), math.cos(th)])
    errs = []
    for nsteps in (20, 40):   # coarse enough that truncation dominates FP32 storage
        m = _macrospin(orc, t_end / nsteps, nsteps)
        errs.append(np.abs(m - exact).max())
        assert abs(np.linalg.norm(m) - 1) < 1e-6
    assert errs[1] < 1e-5
    assert 12 < errs[0] / errs[1] < 20, errs   # 16 for a 4th-order method


def test_llg_orthogonal_and_renormalised(orc):
    rng = np.random.default_rng(4)
    o = np.zeros(3, F)
    for _ in range(200):
        m = (rng.normal(size=3) * 8.6e5).astype(F)
        h = (rng.normal(size=3) * 1e5).astype(F)
        orc.lib().or_llg_mprime(m, h, 0.5, 2.211e5, 8.6e5, o)
        assert abs(np.dot(o.astype(float), m.astype(float))) <= 1e-5 * np.linalg.norm(o) * np.linalg.norm(m)


def refgpu_order_kats():
    """Known answers for the reference CUDA kernel's summation order
    (potential_calc.cu:18-116), derived by hand.  On a line of cells with
    dx = dy = dz = 1 the term of source i at target 0 is charge_i / i, so
    charge_i = c_i * i makes every term exactly c_i; the FP32 sum of
    c = (2^24, 1, 1, -2^24) then depends on the order: exact 2, sequential
    (((2^24 + 1) + 1) - 2^24 = 0 under round-to-even), and the kernel's
    order gives 1.
      * N = 8 (tstride 1): partial t holds term t; the halving tree adds
        ((p0 + p4) + (p2 + p6)) + ((p1 + p5) + (p3 + p7)); with terms
        p1 = 2^24, p5 = 1, p3 = 1, p7 = -2^24: (2^24 + 1) -> 2^24,
        (1 - 2^24) exact, total 1.
      * N = 2048 (tstride 2): partial 0 = term 1, partial t = term 2t + term
        2t+1; with terms 1 = 1, 2 = 2^24, 3 = 1, 1026 = -2^24: partial 1 =
        2^24 + 1 -> 2^24, partial 513 = -2^24, the tree's first level pairs
        1 with 513 (0 exactly) and the total is partial 0 = 1.
    Returns [(N, charge, expected phi at target 0, sequential, exact)]."""
    big = float(2 ** 24)
    out = []
    c = {1: big, 5: 1.0, 3: 1.0, 7: -big}
    q = np.zeros(8, F)
    for i, v in c.items():
        q[i] = F(v * i)
    out.append((8, q, 1.0, 0.0, 2.0))
    c = {1: 1.0, 2: big, 3: 1.0, 1026: -big}
    q = np.zeros(2048, F)
    for i, v in c.items():
        q[i] = F(v * i)
    out.append((2048, q, 1.0, 0.0, 2.0))
    return out


def test_refgpu_order_known_answers(orc):
    one = (F(1.0), F(1.0), F(1.0))
    for n, q, expect, seq, exact in refgpu_order_kats():
        dims = (n, 1, 1)
        got = orc.calc_potential_exact_refgpu(q, dims, one)
        assert got[0] == expect, (n, got[0])
        # the CPU branch adds in source order: the sequential value
        assert orc.calc_potential_exact(q, dims, one)[0] == seq
        assert orc.potential_exact_f64(q, dims, one)[0] == exact


def test_refgpu_order_accuracy_and_closed_forms(orc):
    """FP32 vs FP64 gap on random charges (N above and below 1024 and not a
    multiple of it), the two-charge closed form, self exclusion."""
    rng = np.random.default_rng(8)
    for dims, d in (((7, 5, 3), DX), ((33, 17, 3), (F(5e-9), F(4e-9), F(3e-9))), ((32, 32, 1), DX)):
        n = int(np.prod(dims))
        q = rng.normal(size=n).astype(F) * F(1e14)
        got = orc.calc_potential_exact_refgpu(q, dims, d)
        ex = orc.potential_exact_f64(q, dims, d)
        assert np.abs(got - ex).max() <= 2e-6 * np.abs(ex).max(), dims
    q = np.zeros(5, F)
    q[0], q[3] = 2e14, -1e14
    phi = orc.calc_potential_exact_refgpu(q, (5, 1, 1), DX)
    assert np.allclose(phi, orc.potential_exact_f64(q, (5, 1, 1), DX), rtol=2e-6)
    one = np.zeros(24, F)
    one[7] = 1e14
    assert orc.calc_potential_exact_refgpu(one, (4, 3, 2), DX)[7] == 0.0


def test_treecode_sample_is_the_full_walk(orc):
    """The CPU baseline's timed sample and the full tree-code share one body:
    stride 1 over every level of every charge layer in one thread adds up to
    or_fmm_calc bit for bit; several threads change only the order of the
    float additions."""
    rng = np.random.default_rng(17)
    dims, d, P = (32, 32, 3), DX, 3
    q = (rng.normal(size=int(np.prod(dims))) * 1e14).astype(F)
    full = orc.fmm_calc(q, dims, d, P)
    L = orc.lib()
    for nthreads, exact in ((1, True), (4, False)):
        acc = np.zeros_like(q)
        for zc in range(dims[2]):
            assert L.or_fmm_bfs_sample(q, acc, *dims, *d, P, zc, 2, 5, 1, nthreads) == 0
        if exact:
            assert np.array_equal(acc, full)
        else:
            assert np.allclose(acc, full, rtol=1e-5, atol=1e-6 * np.abs(full).max())
    # a stride samples every stride-th box of the requested levels only
    part = np.zeros_like(q)
    assert L.or_fmm_bfs_sample(q, part, *dims, *d, P, 0, 3, 3, 2, 1) == 0
    assert np.any(part != 0) and not np.array_equal(part, full)
    assert L.or_fmm_bfs_sample(q, part, *dims, *d, P, 0, 1, 3, 1, 1) == 1    # levels start at 2
