"""The FP64 regular-grid FMM oracle: operator identities and convergence."""
import math

import numpy as np
import pytest

from oracle import fmm_oracle as fo


def _far_eval(Mc, P, w, t):
    """phi(t) = sum M_n^m I_n^m(t) for multipole coefficients scaled by w."""
    I = fo.solid_I(P, t)[:, :, 0]
    return sum((Mc[n, P + m] * w ** n * I[n, P + m]) for n in range(P + 1) for m in range(-n, n + 1)).real


def test_solid_harmonic_addition_theorems():
    rng = np.random.default_rng(0)
    s = rng.normal(size=3) * 0.3
    r = rng.normal(size=3) + np.array([3.0, 0, 0])
    N = 30
    Rs, Ir = fo.solid_R(N, s)[:, :, 0], fo.solid_I(N, r)[:, :, 0]
    val = sum(np.conj(Rs[n, N + m]) * Ir[n, N + m] for n in range(N + 1) for m in range(-n, n + 1))
    assert abs(val.real - 1 / np.linalg.norm(r - s)) < 1e-12
    a, b = rng.normal(size=3), rng.normal(size=3)
    Ra, Rb, Rab = (fo.solid_R(6, v)[:, :, 0] for v in (a, b, a + b))
    for n in range(6):
        for m in range(-n, n + 1):
            acc = sum(Ra[j, 6 + k] * Rb[n - j, 6 + m - k] for j in range(n + 1) for k in range(-j, j + 1)
                      if abs(m - k) <= n - j)
            assert abs(acc - Rab[n, 6 + m]) < 1e-12


@pytest.fixture(scope="module")
def ops():
    return fo.build_ops(P=6, w=4, zd=2, Lmax=5, Lleaf=3)


def test_p2m_then_evaluate_matches_direct(ops):
    P, w = 6, 4
    rng = np.random.default_rng(1)
    q = rng.normal(size=w * w)
    Mc = fo._from_real(ops["P2M"] @ q, ops["lst"], P)
    off = np.arange(w) - (w - 1) / 2
    cx, cy = np.meshgrid(off, off, indexing="xy")
    pts = np.stack([cx.ravel(), cy.ravel(), 0 * cx.ravel()], 1)
    for t in ([13.0, 7.0, 1.0], [-9.0, 11.0, 0.0], [0.0, 16.0, 2.0]):
        t = np.array(t)
        exact = np.sum(q / np.linalg.norm(t - pts, axis=1))
        far = _far_eval(Mc, P, w, t[None])
        assert abs(far - exact) < 1e-5 * abs(exact) + 1e-9


def test_m2m_equals_parent_p2m():
    """M2M of the four children's P2M == P2M of the parent box of 2w cells."""
    P, w = 6, 4
    op_w = fo.build_ops(P, w, 1, 5, 3)
    op_2w = fo.build_ops(P, 2 * w, 1, 5, 2)
    rng = np.random.default_rng(2)
    q = rng.normal(size=(2 * w, 2 * w))
    parent = op_2w["P2M"] @ q.ravel()
    acc = np.zeros(op_w["K"])
    for ci, (px, py) in enumerate(fo.CHILD):
        child = q[py * w:(py + 1) * w, px * w:(px + 1) * w].ravel()
        acc += op_w["M2M"][ci] @ (op_w["P2M"] @ child)
    assert np.allclose(acc, parent, atol=1e-12 * np.abs(parent).max())


def test_m2l_l2p_vs_direct_and_l2l_consistency():
    P, w, Lmax, Lleaf = 8, 4, 5, 3
    op = fo.build_ops(P, w, 2, Lmax, Lleaf)
    rng = np.random.default_rng(3)
    q = rng.normal(size=w * w)
    off = np.arange(w) - (w - 1) / 2
    cx, cy = np.meshgrid(off, off, indexing="xy")
    src = np.stack([cx.ravel(), cy.ravel(), 0 * cx.ravel()], 1)
    for (ox, oy, dz) in [(2, 0, 0), (3, -2, 1), (-3, 3, -1)]:
        T = op["M2L"][(Lleaf, ox, oy, dz)]
        Lr = T @ (op["P2M"] @ q)
        phi = op["L2P"] @ Lr       # at target cells (target box is the origin box)
        # source box sits at (ox*w, oy*w) cells, layer -dz relative to the target layer
        s = src + np.array([ox * w, oy * w, -dz])
        exact = np.array([np.sum(q / np.linalg.norm(tp - s, axis=1)) for tp in src])
        assert np.abs(phi - exact).max() < 3e-3 * np.abs(exact).max(), (ox, oy, dz)
    # L2L: the parent local evaluated at child cells == child local evaluated there
    Lp = rng.normal(size=op["K"]) * 0.1
    big = fo.build_ops(P, 2 * w, 1, Lmax, Lleaf - 1)
    phi_parent = big["L2P"] @ Lp
    for ci, (px, py) in enumerate(fo.CHILD):
        phi_child = op["L2P"] @ (op["L2L"][ci] @ Lp)
        blk = phi_parent.reshape(2 * w, 2 * w)[py * w:(py + 1) * w, px * w:(px + 1) * w].ravel()
        assert np.allclose(phi_child, blk, atol=1e-12 * np.abs(blk).max() + 1e-15)


@pytest.mark.parametrize("w", [2, 4])
def test_fmm_oracle_converges_to_direct(w):
    rng = np.random.default_rng(4)
    zd, S = 3, 32
    q = rng.normal(size=(zd, S, S))
    ex = fo.direct_fft(q)
    errs = []
    for P in (2, 4, 6):
        f = fo.FmmOracle(S, S, zd, P, w=w)
        errs.append(np.linalg.norm(f.potential_units(q) - ex) / np.linalg.norm(ex))
    assert errs[0] > errs[1] > errs[2]
    assert errs[2] < 2e-4


def test_fmm_oracle_ragged_grid_and_anisotropic_cells():
    rng = np.random.default_rng(5)
    zd, yd, xd = 2, 21, 27
    q = rng.normal(size=(zd, yd, xd))
    f = fo.FmmOracle(xd, yd, zd, 7, w=4, az=0.6)
    ex = fo.direct_fft(q, az=0.6)
    err = np.linalg.norm(f.potential_units(q) - ex) / np.linalg.norm(ex)
    assert err < 5e-4


def test_direct_fft_vs_brute_force(orc):
    rng = np.random.default_rng(6)
    q = rng.normal(size=(2, 6, 7)).astype(np.float32)
    d = (np.float32(1.0), np.float32(1.0), np.float32(0.7))
    brute = orc.potential_exact_f64(q.ravel(), (7, 6, 2), d)
    az = float(d[2])
    fft = fo.direct_fft(q.astype(float), az=az).ravel() * az
    assert np.abs(fft - brute).max() < 1e-12 * np.abs(brute).max()
