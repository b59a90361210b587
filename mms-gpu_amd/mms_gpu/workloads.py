"""Seeded synthetic inputs: the film configurations the library is measured on.

All randomness comes from a counter-based SplitMix64 stream (Steele, Lea and
Flood, "Fast splittable pseudorandom number generators", OOPSLA 2014):
value i of stream s under seed k is splitmix64(k + (s * 2^32 + i + 1) * golden).
Uniform doubles take the top 53 bits.  Seed 2010 is the reference's seed
(runjob.sh:32).

Units follow the reference: M in A/m (|M| = Ms on magnetic cells, 0 elsewhere),
cells 5 nm cubes, index z*ydim*xdim + y*xdim + x, vectors x,y,z per cell.
"""
import math

import numpy as np

MU_0 = np.float32(4 * np.pi * 1e-7)
MS = np.float32(8.6e5)
AEXCH = np.float32(1.3e-11)
ALFA = np.float32(0.5)
GAMMA = np.float32(2.211e5)
CELL = np.float32(5e-9)
SEED = 2010

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed, stream, count):
    """count uint64 values of `stream` under `seed`."""
    with np.errstate(over="ignore"):
        ctr = (np.uint64(stream) << np.uint64(32)) + np.arange(1, count + 1, dtype=np.uint64)
        z = np.uint64(seed) + ctr * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def uniform(seed, stream, count):
    """count doubles in [0, 1)."""
    return (splitmix64(seed, stream, count) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def random_unit(seed, stream, count):
    """count unit vectors uniform on the sphere: s = 2u1-1, phi = 2 pi u2."""
    u = uniform(seed, stream, 2 * count).reshape(count, 2)
    s = 2 * u[:, 0] - 1
    ph = 2 * np.pi * u[:, 1]
    r = np.sqrt(np.maximum(0.0, 1 - s * s))
    return np.stack([r * np.cos(ph), r * np.sin(ph), s], axis=1)


class Workload:
    """A film: dims, cell size, material flags, initial M, field, steps."""

    def __init__(self, name, xdim, ydim, zdim, material, M, *, dt, steps, hext=(0.0, 0.0, 0.0),
                 solver="direct", P=0, dx=CELL):
        self.name = name
        self.xdim, self.ydim, self.zdim = xdim, ydim, zdim
        self.dx = self.dy = self.dz = np.float32(dx)
        self.material = np.ascontiguousarray(material, np.uint8).ravel()
        self.M = np.ascontiguousarray(M, np.float32).ravel()
        self.dt = np.float32(dt)
        self.steps = steps
        self.hext = tuple(float(v) for v in hext)
        self.solver = solver
        self.P = P

    @property
    def dims(self):
        return (self.xdim, self.ydim, self.zdim)

    @property
    def d(self):
        return (self.dx, self.dy, self.dz)

    @property
    def n(self):
        return self.xdim * self.ydim * self.zdim

    @property
    def n_magnetic(self):
        return int(self.material.sum())


def _finish(mask3, m_unit):
    """M = Ms * m on magnetic cells, 0 elsewhere (float32 AoS)."""
    M = (m_unit.reshape(-1, 3) * float(MS)).astype(np.float32)
    M[mask3.ravel() == 0] = 0
    return M


def c1_square(seed=SEED):
    """32x32x3 film, every cell magnetic, random unit m; direct sum; dt 1e-13 x 1000."""
    xd, yd, zd = 32, 32, 3
    n = xd * yd * zd
    mat = np.ones(n, np.uint8)
    M = _finish(mat, random_unit(seed, 1, n))
    return Workload("c1_square_32x32x3", xd, yd, zd, mat, M, dt=1e-13, steps=1000, solver="direct")


def dots_mask(side=256, n_per_side=4, pitch=64, diameter=40):
    """In-plane mask of n_per_side^2 discs (cell centres within diameter/2)."""
    y, x = np.mgrid[0:side, 0:side]
    mask = np.zeros((side, side), np.uint8)
    centres = []
    for j in range(n_per_side):
        for i in range(n_per_side):
            cx, cy = pitch // 2 + pitch * i, pitch // 2 + pitch * j
            centres.append((cx, cy))
            # disc centred between cells (cx-0.5, cy-0.5) so it is symmetric
            r2 = (x + 0.5 - cx) ** 2 + (y + 0.5 - cy) ** 2
            mask[r2 <= (diameter / 2.0) ** 2] = 1
    return mask, centres


def c2_dots(seed=SEED):
    """256x256x3, 4x4 dots of diameter 40 cells on a 64-cell pitch, per-dot
    vortex (seeded chirality and polarity, core radius 2 cells) plus 1% angular
    noise; H_ext = 10 mT / mu0 along x; FMM P=4; dt 1e-13 x 1000."""
    side, zd = 256, 3
    mask2, centres = dots_mask(side)
    y, x = np.mgrid[0:side, 0:side]
    u = uniform(seed, 2, 2 * len(centres)).reshape(-1, 2)
    m = np.zeros((side, side, 3))
    for k, (cx, cy) in enumerate(centres):
        chir = 1.0 if u[k, 0] < 0.5 else -1.0
        pol = 1.0 if u[k, 1] < 0.5 else -1.0
        dxp, dyp = x + 0.5 - cx, y + 0.5 - cy
        r = np.sqrt(dxp * dxp + dyp * dyp)
        sel = (mask2 == 1) & (r < 32)
        ang = np.arctan2(dyp, dxp)
        mz = pol * np.exp(-(r / 2.0) ** 2)
        ip = np.sqrt(np.maximum(0.0, 1 - mz * mz))
        m[sel, 0] = (-chir * np.sin(ang) * ip)[sel]
        m[sel, 1] = (chir * np.cos(ang) * ip)[sel]
        m[sel, 2] = mz[sel]
    m3 = np.broadcast_to(m, (zd, side, side, 3)).reshape(-1, 3).copy()
    noise = random_unit(seed, 3, m3.shape[0]) * 0.01
    m3 = m3 + noise
    m3 /= np.maximum(np.linalg.norm(m3, axis=1, keepdims=True), 1e-30)
    mat = np.broadcast_to(mask2, (zd, side, side)).ravel().copy()
    M = _finish(mat, m3)
    hx = float(np.float32(10e-3) / MU_0)
    return Workload("c2_dots_256x256x3", side, side, zd, mat, M, dt=1e-13, steps=1000,
                    hext=(hx, 0.0, 0.0), solver="fmm", P=4)


def c3_domains(seed=SEED, side=1024, zd=4, domain=32):
    """1024x1024x4, every cell magnetic, 32x32-cell domains each with a random
    in-plane angle shared through z; FMM P=6; dt 5e-14 x 500."""
    nd = (side + domain - 1) // domain
    ang = 2 * np.pi * uniform(seed, 4, nd * nd).reshape(nd, nd)
    y, x = np.mgrid[0:side, 0:side]
    a = ang[y // domain, x // domain]
    m2 = np.stack([np.cos(a), np.sin(a), np.zeros_like(a)], axis=-1)
    m3 = np.broadcast_to(m2, (zd, side, side, 3)).reshape(-1, 3)
    n = side * side * zd
    mat = np.ones(n, np.uint8)
    M = _finish(mat, m3)
    return Workload(f"c3_domains_{side}x{side}x{zd}", side, side, zd, mat, M, dt=5e-14, steps=500,
                    solver="fmm", P=6)


def probes(seed, n_cells, count, stream=5):
    """count distinct seeded probe cell indices."""
    u = splitmix64(seed, stream, 4 * count)
    idx = np.unique((u % np.uint64(n_cells)).astype(np.int64))
    rng_order = np.argsort(splitmix64(seed, stream + 1, len(idx)))
    return np.sort(idx[rng_order][:count])


def charge_of(M, dims, d):
    """rho = div M with the reference stencil (vector_functions.cpp:8-28:
    central differences, out-of-grid neighbours 0, times 1/2), float32."""
    xd, yd, zd = dims
    m = np.asarray(M, np.float32).reshape(zd, yd, xd, 3)

    def sh(a, ax, step):
        out = np.zeros_like(a)
        src = [slice(None)] * 3
        dst = [slice(None)] * 3
        if step > 0:
            dst[ax], src[ax] = slice(0, -1), slice(1, None)
        else:
            dst[ax], src[ax] = slice(1, None), slice(0, -1)
        out[tuple(dst)] = a[tuple(src)]
        return out
    dx, dy, dz = (np.float32(v) for v in d)
    return (((sh(m[..., 0], 2, 1) - sh(m[..., 0], 2, -1)) / dx + (sh(m[..., 1], 1, 1) - sh(m[..., 1], 1, -1)) / dy
             + (sh(m[..., 2], 0, 1) - sh(m[..., 2], 0, -1)) / dz) * np.float32(0.5)).astype(np.float32).ravel()


def demag_field_error(film, w, nprobe=8192, seed=7):
    """Relative error of the demag field H = grad(phi)/4pi (reference central
    differences) of the film's current state at `nprobe` seeded interior
    probe cells (all six neighbours inside the grid): the film's solver
    against an FP64 direct sum evaluated on the device at each probe and its
    six neighbours.  Returns (relative L2 over the probes, max |dH| / max |H|,
    number of probes)."""
    ch = charge_of(film.get_m(), w.dims, w.d)
    phi = film.potential(ch)
    xd, yd, zd = w.dims
    ix, iy, iz = xd - 2, yd - 2, zd - 2
    pr = probes(seed, ix * iy * iz, nprobe)
    pz, py, px = 1 + pr // (ix * iy), 1 + (pr // ix) % iy, 1 + pr % ix
    offs = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]
    idx = np.stack([((pz + oz) * yd + (py + oy)) * xd + (px + ox) for ox, oy, oz in offs], 1)
    ex = film.potential_probes_f64(ch, idx.ravel()).reshape(-1, 6)
    got = phi[idx].astype(float)

    def grad(v):
        return np.stack([(v[:, 0] - v[:, 1]) / (2 * float(w.dx)), (v[:, 2] - v[:, 3]) / (2 * float(w.dy)),
                         (v[:, 4] - v[:, 5]) / (2 * float(w.dz))], 1) / (4 * np.pi)
    He, Hg = grad(ex), grad(got)
    return (float(np.linalg.norm(Hg - He) / np.linalg.norm(He)),
            float(np.abs(Hg - He).max() / np.abs(He).max()), int(len(pr)))


def single_domain_square(xdim, ydim, zdim):
    """The reference's default geometry (main.cpp:120-129,190-220): the 1-cell
    border of the mask is non-magnetic and only layers 1..zdim-2 are magnetic,
    with M = Ms along x (IC = 0, single domain)."""
    mask = np.zeros((ydim, xdim), np.uint8)
    mask[1:ydim - 1, 1:xdim - 1] = 1
    return layered_single_domain(mask, zdim)


def reference_initial_condition(mask2, zdim, ic, seed, Ms=MS):
    """The reference's initial magnetisation (main.cpp:167-227) on the 2D
    material mask `mask2` (1 = magnetic, [ydim][xdim]): layers 1..zdim-2
    carry material; in-plane angle phi = 0 (ic 0, single domain), the four
    quadrant angles -pi/4, pi/4, 3pi/4, -3pi/4 around (xdim/2, ydim/2) (ic 1,
    vortex), or frand_atob(0, 360) degrees from the C library's
    srand(seed)/rand() (ic 2, random; numerics.cu:125-130), drawn cell by
    cell in z, y, x order over the magnetic cells.  The unit vector is
    formed in double from the float angles and rounded to float, then
    scaled by Ms in float.  Returns (material [n] uint8, M [3n] float32)."""
    import ctypes
    libc = ctypes.CDLL("libc.so.6")
    libc.rand.restype = ctypes.c_int
    libc.srand.argtypes = [ctypes.c_uint]
    RAND_MAX = 2147483647
    if ic == 2:
        libc.srand(seed)
    ydim, xdim = mask2.shape
    mat = np.zeros((zdim, ydim, xdim), np.uint8)
    M = np.zeros((zdim, ydim, xdim, 3), np.float32)
    theta = np.float32(np.pi / 2)
    for z in range(1, zdim - 1):
        for y in range(ydim):
            for x in range(xdim):
                if not mask2[y, x]:
                    continue
                if ic == 0:
                    phi = np.float32(0.0)
                elif ic == 2:
                    r = libc.rand() / float(RAND_MAX)
                    r = 0.0 + float(np.float32(360.0) - np.float32(0.0)) * r
                    phi = np.float32(float(np.float32(r)) * np.pi / 180)
                elif ic == 1:
                    a, b = x - xdim // 2, y - ydim // 2
                    if a > 0 and b > 0:
                        phi = np.float32(-np.pi / 4)
                    elif a <= 0 and b > 0:
                        phi = np.float32(np.pi / 4)
                    elif a <= 0 and b <= 0:
                        phi = np.float32(3 * np.pi / 4)
                    else:
                        phi = np.float32(-3 * np.pi / 4)
                else:
                    raise ValueError("unknown initial condition")
                u = np.array([np.sin(float(theta)) * np.cos(float(phi)), np.sin(float(theta)) * np.sin(float(phi)),
                              np.cos(float(theta))], np.float32)
                M[z, y, x] = np.float32(Ms) * u
                mat[z, y, x] = 1
    return mat.ravel(), M.ravel()


def layered_single_domain(mask2, zdim):
    ydim, xdim = mask2.shape
    mat = np.zeros((zdim, ydim, xdim), np.uint8)
    mat[1:zdim - 1] = mask2
    M = np.zeros((zdim, ydim, xdim, 3), np.float32)
    # Ms * (sin(pi/2) cos 0, sin(pi/2) sin 0, cos(pi/2)) in the reference's float arithmetic
    theta, phi = np.float32(np.pi / 2), np.float32(0.0)
    v = np.array([np.sin(np.float64(theta)) * np.cos(np.float64(phi)),
                  np.sin(np.float64(theta)) * np.sin(np.float64(phi)),
                  np.cos(np.float64(theta))], np.float32)
    M[mat == 1] = (MS * v).astype(np.float32)
    return mat.ravel(), M.ravel()


# ----------------------------------------------------------- film sweep --
def film_mask(kind, side):
    """In-plane mask of one sweep film: 'disc' (cells within side/2 of the
    centre), 'square' (every cell) or 'U' (square minus a notch of width
    side/3 and depth side/2 from the top edge)."""
    y, x = np.mgrid[0:side, 0:side]
    if kind == "disc":
        c = side / 2.0
        return (((x + 0.5 - c) ** 2 + (y + 0.5 - c) ** 2) <= c * c).astype(np.uint8)
    m = np.ones((side, side), np.uint8)
    if kind == "U":
        w = max(1, side // 3)
        x0 = (side - w) // 2
        m[side - side // 2:, x0:x0 + w] = 0
    return m


def c5_films(seed=SEED, nfilm=512, smin=16, smax=96, zd=3):
    """Hysteresis-style sweep: nfilm independent films, side ~ U{smin..smax},
    masks uniform over {disc, square, U}, random unit m on magnetic cells,
    film f in a uniform field (-50 + 100 f/(nfilm-1)) mT / mu0 along x;
    direct sum below 4096 cells, FMM P=4 above; dt 1e-13 x 2000."""
    u = uniform(seed, 6, 2 * nfilm).reshape(nfilm, 2)
    films = []
    kinds = ("disc", "square", "U")
    for f in range(nfilm):
        side = smin + int(u[f, 0] * (smax - smin + 1))
        kind = kinds[int(u[f, 1] * 3)]
        mask = film_mask(kind, side)
        mat = np.broadcast_to(mask, (zd, side, side)).ravel().copy()
        n = side * side * zd
        m = random_unit(seed, 1000 + f, n)
        M = _finish(mat, m)
        hx = float(np.float32((-50 + 100 * f / max(1, nfilm - 1)) * 1e-3) / MU_0)
        solver = "direct" if n < 4096 else "fmm"
        films.append(Workload(f"film{f}_{kind}_{side}x{side}x{zd}", side, side, zd, mat, M, dt=1e-13,
                              steps=2000, hext=(hx, 0.0, 0.0), solver=solver, P=4 if solver == "fmm" else 0))
    return films
