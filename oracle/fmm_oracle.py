"""FP64 numpy oracle for the regular-grid FMM and the exact direct sum.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

* ``direct_fft(q, ay, az)`` -- the exact O(N^2) sum phi_i = sum_{j != i} q_j/|r_i - r_j|
  (distances in units of dx, r = (dx, dy*ay, dz*az)) computed as a zero-padded
  FFT convolution with the kernel G(0) = 0, G(r) = 1/|r|.  Mathematically
  identical to the direct sum; FP64 rounding only.
* ``solid_R/solid_I`` -- regular / irregular solid harmonics in the
  normalisation of mms-gpu_amd/csrc/fmm_operators.cpp, written independently
  here (vectorised Legendre recurrences).
* ``FmmOracle`` -- the same algorithm as the GPU path (per-layer quadtrees,
  leaf P2M, M2M, M2L over the interaction list, L2L, L2P, near-field P2P over
  the 3x3 neighbour leaves of every layer) in FP64, with the operators built
  here from the translation theorems.

Pinning: operator identities (P2M+M2P vs 1/r, M2M(P2M(child)) == P2M(parent),
M2L+L2P vs direct multipole evaluation, L2L(L) evaluated == L evaluated),
FmmOracle vs direct_fft with error falling geometrically in P, and
direct_fft vs the brute-force FP64 direct sum of mms_oracle.c (tests/).
"""
import math
from math import factorial as fac

import numpy as np


# ------------------------------------------------------------ harmonics --
def _legendre_cs(nmax, x):
    x = np.asarray(x, float)
    s = np.sqrt(np.maximum(0.0, 1 - x * x))
    P = np.zeros((nmax + 1, nmax + 1) + x.shape)
    P[0, 0] = 1.0
    for m in range(1, nmax + 1):
        P[m, m] = -(2 * m - 1) * s * P[m - 1, m - 1]
    for m in range(nmax):
        P[m + 1, m] = (2 * m + 1) * x * P[m, m]
    for m in range(nmax + 1):
        for n in range(m + 2, nmax + 1):
            P[n, m] = ((2 * n - 1) * x * P[n - 1, m] - (n + m - 1) * P[n - 2, m]) / (n - m)
    return P


def _harm(nmax, r, kind):
    r = np.atleast_2d(np.asarray(r, float))
    x, y, z = r[:, 0], r[:, 1], r[:, 2]
    rr = np.sqrt(x * x + y * y + z * z)
    ct = np.where(rr > 0, z / np.where(rr > 0, rr, 1.0), 1.0)
    ph = np.arctan2(y, x)
    P = _legendre_cs(nmax, ct)
    out = np.zeros((nmax + 1, 2 * nmax + 1, len(rr)), complex)
    for n in range(nmax + 1):
        for m in range(n + 1):
            e = np.exp(1j * m * ph)
            if kind == "R":
                v = rr ** n * P[n, m] * e / fac(n + m)
            else:
                v = fac(n - m) * P[n, m] * e / rr ** (n + 1)
            out[n, nmax + m] = v
            out[n, nmax - m] = (-1) ** m * np.conj(v)
    return out


def solid_R(nmax, r):
    """R_n^m(r) = r^n P_n^m(cos t) e^{imp}/(n+m)!, table [n, m+nmax, point]."""
    return _harm(nmax, r, "R")


def solid_I(nmax, r):
    """I_n^m(r) = (n-m)! P_n^m(cos t) e^{imp}/r^{n+1}, table [n, m+nmax, point]."""
    return _harm(nmax, r, "I")


def coef_list(P):
    return [(n, m) for n in range(P + 1) for m in range(n + 1) if (n + m) % 2 == 0]


def ncoef(P):
    return (P + 1) * (P + 2) // 2


def _from_real(v, lst, P):
    C = np.zeros((P + 1, 2 * P + 1), complex)
    i = 0
    for n, m in lst:
        re = v[i]
        i += 1
        im = 0.0
        if m > 0:
            im = v[i]
            i += 1
        C[n, P + m] = re + 1j * im
        C[n, P - m] = (-1) ** m * (re - 1j * im)
    return C


def _to_real(C, lst, P):
    out = []
    for n, m in lst:
        out.append(C[n, P + m].real)
        if m > 0:
            out.append(C[n, P + m].imag)
    return np.array(out)


# ------------------------------------------------------------- operators --
CHILD = [(0, 0), (1, 0), (0, 1), (1, 1)]


def build_ops(P, w, zd, Lmax, Lleaf, ay=1.0, az=1.0):
    """Operators in the scaled real basis: P2M [K, w*w], L2P [w*w, K],
    M2M/L2L [4, K, K], M2L {(level, ox, oy, dz): [K, K]}."""
    lst = coef_list(P)
    K = ncoef(P)
    off = np.arange(w) - (w - 1) / 2.0
    cx, cy = np.meshgrid(off, off * ay, indexing="xy")
    pts = np.stack([cx.ravel(), cy.ravel(), 0 * cx.ravel()], 1) / w
    R = solid_R(P, pts)
    P2M = np.zeros((K, w * w))
    L2P = np.zeros((w * w, K))
    for c in range(w * w):
        P2M[:, c] = _to_real(np.conj(R[:, :, c]), lst, P)
        row = []
        for n, m in lst:
            z = R[n, P + m, c]
            row += [z.real] if m == 0 else [2 * z.real, -2 * z.imag]
        L2P[c] = row
    M2M = np.zeros((4, K, K))
    L2L = np.zeros((4, K, K))
    for ci, (px, py) in enumerate(CHILD):
        d = np.array([[(px - 0.5) * 0.5, (py - 0.5) * 0.5 * ay, 0.0]])
        Rd = solid_R(P, d)[:, :, 0]
        for e in range(K):
            v = np.zeros(K)
            v[e] = 1
            C = _from_real(v, lst, P)
            Mp = np.zeros_like(C)
            Lc = np.zeros_like(C)
            for n in range(P + 1):
                for m in range(-n, n + 1):
                    s = 0
                    for j in range(n + 1):
                        for k in range(-j, j + 1):
                            if abs(m - k) <= n - j:
                                s += np.conj(Rd[j, P + k]) * C[n - j, P + m - k] * 0.5 ** (n - j)
                    Mp[n, P + m] = s
                    s = 0
                    for j in range(n, P + 1):
                        for k in range(-j, j + 1):
                            if abs(k - m) <= j - n:
                                s += C[j, P + k] * Rd[j - n, P + k - m]
                    Lc[n, P + m] = s * 0.5 ** n
            M2M[ci, :, e] = _to_real(Mp, lst, P)
            L2L[ci, :, e] = _to_real(Lc, lst, P)
    M2L = {}
    offs = [(ox, oy) for oy in range(-3, 4) for ox in range(-3, 4) if max(abs(ox), abs(oy)) >= 2]
    # each real basis vector has at most two nonzero complex entries (l, +-m)
    basis = []
    for e in range(K):
        v = np.zeros(K)
        v[e] = 1
        C = _from_real(v, lst, P)
        nz = [(l, m, C[l, P + m]) for l in range(P + 1) for m in range(-l, l + 1) if C[l, P + m] != 0]
        basis.append(nz)
    for L in range(2, Lleaf + 1):
        wl = 2.0 ** (Lmax - L)
        for dz in range(-(zd - 1), zd):
            D = np.array([[-ox, -oy * ay, dz * az / wl] for ox, oy in offs])
            I = solid_I(2 * P, D)
            for oi, (ox, oy) in enumerate(offs):
                T = np.zeros((K, K))
                for e in range(K):
                    Lc = np.zeros((P + 1, 2 * P + 1), complex)
                    for j, k in lst:
                        s = sum(c * I[l + j, 2 * P + m - k, oi] for l, m, c in basis[e])
                        Lc[j, P + k] = (-1) ** (j + k) * s / wl
                    T[:, e] = _to_real(Lc, lst, P)
                M2L[(L, ox, oy, dz)] = T
    return dict(P2M=P2M, L2P=L2P, M2M=M2M, L2L=L2L, M2L=M2L, K=K, lst=lst)


# -------------------------------------------------------------- direct --
def direct_fft(q, ay=1.0, az=1.0):
    """Exact sum_{j != i} q_j / |r_i - r_j| on a (zd, yd, xd) grid, FP64."""
    zd, yd, xd = q.shape
    gz = np.arange(-(zd - 1), zd) * az
    gy = np.arange(-(yd - 1), yd) * ay
    gx = np.arange(-(xd - 1), xd)
    Z, Y, X = np.meshgrid(gz, gy, gx, indexing="ij")
    r = np.sqrt(Z * Z + Y * Y + X * X)
    G = np.where(r > 0, 1.0 / np.where(r > 0, r, 1.0), 0.0)
    sh = tuple(int(2 ** math.ceil(math.log2(s))) for s in (3 * zd - 1, 3 * yd - 1, 3 * xd - 1))
    F = np.fft.rfftn(G, sh, axes=(0, 1, 2)) * np.fft.rfftn(q, sh, axes=(0, 1, 2))
    out = np.fft.irfftn(F, sh, axes=(0, 1, 2))
    return out[zd - 1:2 * zd - 1, yd - 1:2 * yd - 1, xd - 1:2 * xd - 1]


# ------------------------------------------------------------------ FMM --
class FmmOracle:
    """FP64 regular-grid FMM of order P with w x w leaf boxes."""

    def __init__(self, xd, yd, zd, P, w=4, ay=1.0, az=1.0):
        side = max(xd, yd)
        self.Lmax = int(math.ceil(math.log2(side))) if side > 1 else 0
        self.S = 2 ** self.Lmax
        self.Lleaf = self.Lmax - int(round(math.log2(w)))
        if self.Lleaf < 2:
            raise ValueError("grid too small for this leaf width")
        self.xd, self.yd, self.zd, self.P, self.w, self.ay, self.az = xd, yd, zd, P, w, ay, az
        self.ops = build_ops(P, w, zd, self.Lmax, self.Lleaf, ay, az)
        # near-field P2P operator: target (iy, ix) in a leaf, source (sy, sx) in
        # the 3w x 3w window starting one leaf before it; one matrix per dz
        G = np.zeros((2 * zd - 1, w * w, 9 * w * w))
        for dzi in range(2 * zd - 1):
            dz = dzi - (zd - 1)
            for iy in range(w):
                for ix in range(w):
                    for sy in range(3 * w):
                        for sx in range(3 * w):
                            rx, ry, rz = sx - w - ix, (sy - w - iy) * ay, dz * az
                            r2 = rx * rx + ry * ry + rz * rz
                            G[dzi, iy * w + ix, sy * 3 * w + sx] = 1 / math.sqrt(r2) if r2 > 0 else 0.0
        self.G = G

    def multipoles(self, q):
        zd, w, S = self.zd, self.w, self.S
        nb = 2 ** self.Lleaf
        qp = np.zeros((zd, S, S))
        qp[:, :self.yd, :self.xd] = q
        qb = qp.reshape(zd, nb, w, nb, w).transpose(0, 1, 3, 2, 4).reshape(zd, nb, nb, w * w)
        M = {self.Lleaf: qb @ self.ops["P2M"].T}
        for L in range(self.Lleaf - 1, 1, -1):
            Mc = M[L + 1]
            Mp = np.zeros((zd, 2 ** L, 2 ** L, self.ops["K"]))
            for ci, (px, py) in enumerate(CHILD):
                Mp += Mc[:, py::2, px::2, :] @ self.ops["M2M"][ci].T
            M[L] = Mp
        return M, qp

    def locals(self, M):
        zd, K = self.zd, self.ops["K"]
        Lc = {}
        for L in range(2, self.Lleaf + 1):
            n = 2 ** L
            Lloc = np.zeros((zd, n, n, K))
            Mpad = np.zeros((zd, n + 6, n + 6, K))
            Mpad[:, 3:n + 3, 3:n + 3] = M[L]
            for py in range(2):
                for px in range(2):
                    for oy in range(-2 - py, 4 - py):
                        for ox in range(-2 - px, 4 - px):
                            if max(abs(ox), abs(oy)) < 2:
                                continue
                            for zt in range(zd):
                                for zs in range(zd):
                                    T = self.ops["M2L"][(L, ox, oy, zt - zs)]
                                    src = Mpad[zs, 3 + py + oy:3 + n + oy:2, 3 + px + ox:3 + n + ox:2]
                                    Lloc[zt, py::2, px::2] += src @ T.T
            if L > 2:
                for ci, (px, py) in enumerate(CHILD):
                    Lloc[:, py::2, px::2] += Lc[L - 1] @ self.ops["L2L"][ci].T
            Lc[L] = Lloc
        return Lc

    def potential_units(self, q):
        """sum rho/|r| with r in units of dx (multiply by dV/dx for phi)."""
        q = np.asarray(q, float).reshape(self.zd, self.yd, self.xd)
        zd, w, S = self.zd, self.w, self.S
        nb = 2 ** self.Lleaf
        M, qp = self.multipoles(q)
        Lc = self.locals(M)
        far = Lc[self.Lleaf] @ self.ops["L2P"].T
        phi = far.reshape(zd, nb, nb, w, w).transpose(0, 1, 3, 2, 4).reshape(zd, S, S)
        # near field: windows of 3w x 3w around every leaf
        qpad = np.zeros((zd, S + 2 * w, S + 2 * w))
        qpad[:, w:w + S, w:w + S] = qp
        win = np.lib.stride_tricks.sliding_window_view(qpad, (3 * w, 3 * w), axis=(1, 2))[:, ::w, ::w]
        win = win.reshape(zd, nb, nb, 9 * w * w)
        near = np.zeros((zd, nb, nb, w * w))
        for zt in range(zd):
            for zs in range(zd):
                near[zt] += win[zs] @ self.G[zs - zt + zd - 1].T
        phi += near.reshape(zd, nb, nb, w, w).transpose(0, 1, 3, 2, 4).reshape(zd, S, S)
        return phi[:, :self.yd, :self.xd]
