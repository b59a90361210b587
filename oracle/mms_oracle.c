/*
 * mms_oracle.c -- CPU oracle for the field-evaluation + LLG RK4 path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: it may be loaded by
 * tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and
 * by nothing else.  The product library (libmmsgpu.so) never links or calls it.
 *
 * Two kinds of functions live here:
 *   (1) FP32 restatements of the reference CPU algorithms, following the
 *       reference's loop order, skip rules and float/double promotions
 *       (each function cites the reference file:line it restates);
 *   (2) FP64 oracles written for this project (the direct-sum potential in
 *       double precision, on all targets or on probe cells) that pin accuracy.
 *
 * Pinning (how each function is checked, see tests/):
 *   or_divergence_3d / or_gradient_3d / or_add_exchange_field
 *       bit-exact against the reference vector_functions.cpp compiled from its
 *       own sources (oracle/_ref), plus hand-computed stencils on tiny grids.
 *   or_associated_legendre / or_spherical_harmonic
 *       bit-exact against the reference numerics.cu compiled from its own
 *       sources (oracle/_ref), plus textbook P_l^m tables.
 *   or_tree_lists
 *       identical to the reference Box.cu interaction/neighbour lists
 *       (oracle/_ref), plus the 27/8-element counting invariants.
 *   or_calc_potential_exact (FP32) / or_potential_*_f64
 *       two-charge closed form, self-exclusion, linearity, FP32 vs FP64 gap,
 *       brute force vs an independent FFT convolution (tests/).
 *   or_calc_potential_exact_refgpu
 *       known answers derived by hand from the kernel's documented order
 *       (charges with exact terms whose FP32 sum depends on the order: chunk
 *       sequence and tree pairing give one value, sequential and exact sums
 *       others), FP32 vs FP64 gap, closed forms (tests/).
 *   or_fmm_calc (reference tree-code, P <= 4)
 *       vs FP64 direct sum within the order-P truncation bound; error falls
 *       with P; reduces to the exact sum when every box is a leaf neighbour.
 *   or_llg_mprime / or_hfield / or_rk4_step
 *       LLG_Mprime bit-exact against the reference ode_functions.cpp object
 *       (oracle/_ref); macrospin closed form; M.dM/dt = 0; |M| = Ms after
 *       renormalisation; RK4 order-4 convergence; the reference's own logged
 *       step-1 energies and torques (the log-files directory).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdint.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef float fptype;

/* ------------------------------------------------------------------------ */
/* small Vector3 / Cmpx helpers with the reference's float semantics         */
/* Vector3.cu:66-91, Cmpx.cu:32-72                                           */
/* ------------------------------------------------------------------------ */
typedef struct { float x, y, z; } v3;
typedef struct { float re, im; } cx;

static inline v3 v3_make(float x, float y, float z) { v3 r = {x, y, z}; return r; }
static inline v3 v3_add(v3 a, v3 b) { return v3_make(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 v3_sub(v3 a, v3 b) { return v3_make(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 v3_scale(v3 a, float k) { return v3_make(k * a.x, k * a.y, k * a.z); }
static inline float v3_dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 v3_cross(v3 a, v3 b) {
    return v3_make(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline float v3_mag(v3 a) { return sqrtf(a.x * a.x + a.y * a.y + a.z * a.z); }
static inline float v3_colat(v3 a) { return acosf(a.z / sqrtf(a.x * a.x + a.y * a.y + a.z * a.z)); }
static inline float v3_azim(v3 a) { return atan2f(a.y, a.x); }

static inline cx cx_make(float re, float im) { cx r = {re, im}; return r; }
static inline void cx_addeq(cx *a, cx b) { a->re += b.re; a->im += b.im; }
static inline void cx_muleq(cx *a, cx b) {
    float x = a->re * b.re - a->im * b.im;
    float y = a->re * b.im + a->im * b.re;
    a->re = x; a->im = y;
}
static inline void cx_scaleeq(cx *a, float k) { a->re *= k; a->im *= k; }

/* ------------------------------------------------------------------------ */
/* numerics.cu:8-35 factorial (LUT 0..8, recursion beyond)                    */
/* ------------------------------------------------------------------------ */
static int or_factorial(int x) {
    static const int lut[9] = {1, 1, 2, 6, 24, 120, 720, 5040, 40320};
    if (x >= 0 && x <= 8) return lut[x];
    return (x == 0) ? 1 : x * or_factorial(x - 1);
}

/* numerics.cu:58-104 associated_legendre: P_l^m(x), Condon-Shortley sign,
 * hard-coded for l <= 4, NaN above.  The double-typed literals (1/2.0 etc.)
 * promote the product chain to double before the final float return; that is
 * restated with explicit casts. */
float or_associated_legendre(int l, int m, float x) {
    switch (l) {
    case 0:
        return m == 0 ? 1.0f : 0.0f;
    case 1:
        if (m == 0) return x;
        if (m == 1) return -sqrtf(1 - x * x);
        return 0.0f;
    case 2:
        if (m == 0) return (float)(0.5 * (double)(3 * x * x - 1));
        if (m == 1) return -3 * x * sqrtf(1 - x * x);
        if (m == 2) return 3 * (1 - x * x);
        return 0.0f;
    case 3:
        if (m == 0) return (float)(0.5 * (double)x * (double)(5 * x * x - 3));
        if (m == 1) return (float)(-1.5 * (double)(5 * x * x - 1) * (double)sqrtf(1 - x * x));
        if (m == 2) return 15 * x * (1 - x * x);
        if (m == 3) return -15 * powf(1 - x * x, 1.5f);
        return 0.0f;
    case 4:
        if (m == 0) return (float)(0.125 * (double)(x * x * (35 * x * x - 30) + 3));
        if (m == 1) return (float)(-2.5 * (double)x * (double)(7 * x * x - 3) * (double)sqrtf(1 - x * x));
        if (m == 2) return (float)(7.5 * (double)(7 * x * x - 1) * (double)(1 - x * x));
        if (m == 3) return -105 * x * powf(1 - x * x, 1.5f);
        if (m == 4) return 105 * (1 - x * x) * (1 - x * x);
        return 0.0f;
    default:
        return NAN;
    }
}

/* numerics.cu:107-114 spherical_harmonic: P_l^|m|(cos theta) e^{i m phi}.
 * cos/sin of float arguments resolve to the double ::cos/::sin there (only the
 * C double overloads are in the global namespace), so they are taken in double
 * and rounded to float on return. */
static cx or_sph(int l, int m, float theta, float phi) {
    float mag = or_associated_legendre(l, (int)fabsf((float)m), (float)cos((double)theta));
    float ang = m * phi;
    return cx_make((float)((double)mag * cos((double)ang)), (float)((double)mag * sin((double)ang)));
}

void or_spherical_harmonic(int l, int m, float theta, float phi, float *re, float *im) {
    cx c = or_sph(l, m, theta, phi);
    *re = c.re; *im = c.im;
}

/* ------------------------------------------------------------------------ */
/* vector_functions.cpp:8-28 divergence_3d (AoS Vector3 in, scalar out)      */
/* ------------------------------------------------------------------------ */
void or_divergence_3d(const float *V, int xdim, int ydim, int zdim,
                      float dx, float dy, float dz, float *S) {
    const v3 *M = (const v3 *)V;
    for (int z = 0; z < zdim; z++)
        for (int y = 0; y < ydim; y++)
            for (int x = 0; x < xdim; x++) {
                float x2 = (x != xdim - 1) ? M[(z)*ydim * xdim + (y)*xdim + (x + 1)].x : 0;
                float x0 = (x != 0) ? M[(z)*ydim * xdim + (y)*xdim + (x - 1)].x : 0;
                float y2 = (y != ydim - 1) ? M[(z)*ydim * xdim + (y + 1) * xdim + (x)].y : 0;
                float y0 = (y != 0) ? M[(z)*ydim * xdim + (y - 1) * xdim + (x)].y : 0;
                float z2 = (z != zdim - 1) ? M[(z + 1) * ydim * xdim + (y)*xdim + (x)].z : 0;
                float z0 = (z != 0) ? M[(z - 1) * ydim * xdim + (y)*xdim + (x)].z : 0;
                S[z * ydim * xdim + y * xdim + x] =
                    (float)((double)((x2 - x0) / dx + (y2 - y0) / dy + (z2 - z0) / dz) / 2.0);
            }
}

/* vector_functions.cpp:31-54 gradient_3d (overwrites V) */
void or_gradient_3d(const float *S, int xdim, int ydim, int zdim,
                    float dx, float dy, float dz, float cm, float *Vout) {
    v3 *V = (v3 *)Vout;
    for (int z = 0; z < zdim; z++)
        for (int y = 0; y < ydim; y++)
            for (int x = 0; x < xdim; x++) {
                float x2 = (x != xdim - 1) ? S[(z)*ydim * xdim + (y)*xdim + (x + 1)] : 0;
                float x0 = (x != 0) ? S[(z)*ydim * xdim + (y)*xdim + (x - 1)] : 0;
                float y2 = (y != ydim - 1) ? S[(z)*ydim * xdim + (y + 1) * xdim + (x)] : 0;
                float y0 = (y != 0) ? S[(z)*ydim * xdim + (y - 1) * xdim + (x)] : 0;
                float z2 = (z != zdim - 1) ? S[(z + 1) * ydim * xdim + (y)*xdim + (x)] : 0;
                float z0 = (z != 0) ? S[(z - 1) * ydim * xdim + (y)*xdim + (x)] : 0;
                V[z * ydim * xdim + y * xdim + x].x = (x2 - x0) / (2 * dx) * cm;
                V[z * ydim * xdim + y * xdim + x].y = (y2 - y0) / (2 * dy) * cm;
                V[z * ydim * xdim + y * xdim + x].z = (z2 - z0) / (2 * dz) * cm;
            }
}

/* vector_functions.cpp:57-103 add_exchange_field (H += 2A/(mu Ms^2) lap M) */
void or_add_exchange_field(const float *Min, float Ms, float Aexch, float mu,
                           int xdim, int ydim, int zdim, float dx, float dy, float dz,
                           float *Hio) {
    const v3 *M = (const v3 *)Min;
    v3 *H = (v3 *)Hio;
    const float cm = 2 * Aexch / (mu * Ms * Ms);
#define MI(zz, yy, xx) M[(zz)*ydim * xdim + (yy)*xdim + (xx)]
    for (int z = 0; z < zdim; z++)
        for (int y = 0; y < ydim; y++)
            for (int x = 0; x < xdim; x++) {
                float L[3];
                for (int c = 0; c < 3; c++) {
#define CMP(v) (c == 0 ? (v).x : (c == 1 ? (v).y : (v).z))
                    float ctr = CMP(MI(z, y, x));
                    L[c] = (((x != xdim - 1) ? CMP(MI(z, y, x + 1)) : 0) +
                            ((x != 0) ? CMP(MI(z, y, x - 1)) : 0) - (2 * ctr)) / (dx * dx) +
                           (((y != ydim - 1) ? CMP(MI(z, y + 1, x)) : 0) +
                            ((y != 0) ? CMP(MI(z, y - 1, x)) : 0) - (2 * ctr)) / (dy * dy) +
                           (((z != zdim - 1) ? CMP(MI(z + 1, y, x)) : 0) +
                            ((z != 0) ? CMP(MI(z - 1, y, x)) : 0) - (2 * ctr)) / (dz * dz);
#undef CMP
                }
                v3 add = v3_scale(v3_make(L[0], L[1], L[2]), cm);
                H[z * ydim * xdim + y * xdim + x] = v3_add(H[z * ydim * xdim + y * xdim + x], add);
            }
#undef MI
}

/* ------------------------------------------------------------------------ */
/* fmm_calc.cu:412-461 calc_potential_exact, CPU branch (FP32).  Sources are  */
/* walked outer (z_, y_, x_) skipping q == 0; each target accumulates its     */
/* sources in linear order, so the result does not depend on thread count.   */
/* ------------------------------------------------------------------------ */
int or_calc_potential_exact(const float *charge, int xdim, int ydim, int zdim,
                            float dx, float dy, float dz, float *potential) {
    const int n = xdim * ydim * zdim;
    for (int i = 0; i < n; i++) potential[i] = 0;
    for (int z_ = 0; z_ < zdim; z_++)
        for (int y_ = 0; y_ < ydim; y_++)
            for (int x_ = 0; x_ < xdim; x_++) {
                float dV = dx * dy * dz;
                float q = charge[z_ * ydim * xdim + y_ * xdim + x_] * dV;
                if (q == 0) continue;
                for (int z = 0; z < zdim; z++) {
#pragma omp parallel for schedule(static)
                    for (int y = 0; y < ydim; y++)
                        for (int x = 0; x < xdim; x++) {
                            if (z == z_ && y == y_ && x == x_) continue;
                            v3 R = v3_make((x - x_) * dx, (y - y_) * dy, (z - z_) * dz);
                            potential[z * ydim * xdim + y * xdim + x] += q / v3_mag(R);
                        }
                }
            }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* potential_calc.cu:18-116 calc_potential_exact_kernel (the reference's     */
/* CUDA direct sum), restated on the CPU: for target bi, thread t of the      */
/* 1024-thread block sums sources [t tstride, (t+1) tstride) in index order,  */
/* tstride = ceilf(N / 1024.f) (potential_calc.cu:26,57-75); the 1024         */
/* partials are added by the halving tree s[t] += s[t + h], h = 512 .. 1      */
/* (potential_calc.cu:82-111).  Every operation is rounded on its own (the    */
/* oracle is built with -ffp-contract=off).  The block/target assignment of   */
/* the kernel (bstride) does not change any value and is not restated.        */
/* ------------------------------------------------------------------------ */
int or_calc_potential_exact_refgpu(const float *charge, int xdim, int ydim, int zdim,
                                   float dx, float dy, float dz, float *potential) {
    const int N = xdim * ydim * zdim;
    const int tstride = (int)ceilf(N / (fptype)1024);
#pragma omp parallel for schedule(dynamic, 16)
    for (int bi = 0; bi < N; bi++) {
        float sdata[1024];
        const int xs = bi % xdim;
        const int ys = ((bi - xs) / xdim) % ydim;
        const int zs = (bi - xs - ys * xdim) / (xdim * ydim);
        for (int t = 0; t < 1024; t++) {
            sdata[t] = 0;
            const int ti1 = t * tstride;
            if (ti1 >= N) continue;
            float pot = 0;
            for (int i = ti1; i < ti1 + tstride; i++) {
                if (i >= N) continue;
                float dV = dx * dy * dz;
                float q = charge[i] * dV;
                int x_ = i % xdim;
                int y_ = ((i - x_) / xdim) % ydim;
                int z_ = (i - x_ - y_ * xdim) / (xdim * ydim);
                if (bi != i) {
                    v3 R = v3_make((xs - x_) * dx, (ys - y_) * dy, (zs - z_) * dz);
                    pot += q / v3_mag(R);
                }
            }
            sdata[t] = pot;
        }
        for (int h = 512; h > 0; h >>= 1)
            for (int t = 0; t < h; t++) sdata[t] += sdata[t + h];
        potential[bi] = sdata[0];
    }
    return 0;
}

/* FP64 direct sum (this project's accuracy oracle): phi_i = sum_{j != i}
 * rho_j dV / |r_i - r_j|, every quantity in double, targets in parallel. */
static double pot_one_f64(const float *charge, int xdim, int ydim, int zdim,
                          double dx, double dy, double dz, int tx, int ty, int tz) {
    double dV = dx * dy * dz, acc = 0.0;
    for (int z_ = 0; z_ < zdim; z_++) {
        double rz = (double)(tz - z_) * dz;
        for (int y_ = 0; y_ < ydim; y_++) {
            double ry = (double)(ty - y_) * dy;
            const float *row = charge + ((size_t)z_ * ydim + y_) * xdim;
            double part = 0.0;
            for (int x_ = 0; x_ < xdim; x_++) {
                double rx = (double)(tx - x_) * dx;
                double r2 = rx * rx + ry * ry + rz * rz;
                if (r2 > 0.0) part += (double)row[x_] / sqrt(r2);
            }
            acc += part;
        }
    }
    return acc * dV;
}

void or_potential_exact_f64(const float *charge, int xdim, int ydim, int zdim,
                            double dx, double dy, double dz, double *potential) {
    const long n = (long)xdim * ydim * zdim;
#pragma omp parallel for schedule(dynamic, 16)
    for (long i = 0; i < n; i++) {
        int tx = (int)(i % xdim), ty = (int)((i / xdim) % ydim), tz = (int)(i / ((long)xdim * ydim));
        potential[i] = pot_one_f64(charge, xdim, ydim, zdim, dx, dy, dz, tx, ty, tz);
    }
}

void or_potential_probes_f64(const float *charge, int xdim, int ydim, int zdim,
                             double dx, double dy, double dz,
                             const int64_t *probes, int nprobe, double *out) {
#pragma omp parallel for schedule(dynamic, 1)
    for (int p = 0; p < nprobe; p++) {
        long i = (long)probes[p];
        int tx = (int)(i % xdim), ty = (int)((i / xdim) % ydim), tz = (int)(i / ((long)xdim * ydim));
        out[p] = pot_one_f64(charge, xdim, ydim, zdim, dx, dy, dz, tx, ty, tz);
    }
}

/* ------------------------------------------------------------------------ */
/* Box.cu quadtree, restated on (level, x, y) arithmetic.                    */
/* Box index at a level is the Morton code with x in bit 0 of each base-4     */
/* digit (Box.cu:150-163, 59-71); the 8 neighbours are listed in the order of */
/* Box.cu:100-107 and exist iff inside [0, 2^level)^2 (Box.cu:108-118);       */
/* the interaction list is the children (order 0..3) of the parent's         */
/* neighbours, in neighbour order, minus own neighbours (Box.cu:121-145).     */
/* ------------------------------------------------------------------------ */
static inline void box_xy(int level, unsigned idx, int *x, int *y) {
    int xx = 0, yy = 0;
    for (int b = 0; b < level; b++) {
        xx |= ((idx >> (2 * b)) & 1) << b;
        yy |= ((idx >> (2 * b + 1)) & 1) << b;
    }
    *x = xx; *y = yy;
}
static inline unsigned box_idx(int level, int x, int y) {
    unsigned idx = 0;
    for (int b = 0; b < level; b++)
        idx |= (((unsigned)(x >> b) & 1u) << (2 * b)) | (((unsigned)(y >> b) & 1u) << (2 * b + 1));
    return idx;
}
static const int NBX[8] = {-1, 0, 1, -1, 1, -1, 0, 1};
static const int NBY[8] = {-1, -1, -1, 0, 0, 1, 1, 1};

/* neighbours: nb[8] box indices or -1.  returns count */
static int box_neighbors(int level, int x, int y, int *nb) {
    int n = 0, side = 1 << level;
    for (int j = 0; j < 8; j++) {
        int X = x + NBX[j], Y = y + NBY[j];
        if (X >= 0 && X < side && Y >= 0 && Y < side) { nb[j] = (int)box_idx(level, X, Y); n++; }
        else nb[j] = -1;
    }
    return n;
}

/* interaction list: il[27] box indices (same level), -1 padded. returns count */
static int box_interactions(int level, int x, int y, int *il) {
    for (int i = 0; i < 27; i++) il[i] = -1;
    if (level <= 1) return 0;
    int own[8], par[8];
    box_neighbors(level, x, y, own);
    box_neighbors(level - 1, x >> 1, y >> 1, par);
    int c1 = 0;
    for (int j = 0; j < 8; j++) {
        if (par[j] < 0) continue;
        int px, py;
        box_xy(level - 1, (unsigned)par[j], &px, &py);
        for (int k = 0; k < 4; k++) {
            int cxb = 2 * px + (k & 1), cyb = 2 * py + (k >> 1);
            int cidx = (int)box_idx(level, cxb, cyb);
            int found = 0;
            for (int q = 0; q < 8; q++) if (own[q] == cidx) { found = 1; break; }
            if (!found && c1 < 27) il[c1++] = cidx;
        }
    }
    return c1;
}

/* exported for tests: lists of every box at a level, 8 neighbour + 27
 * interaction indices per box (-1 padded) */
void or_tree_lists(int level, int *neighbors /* 4^level*8 */, int *interactions /* 4^level*27 */) {
    unsigned nbox = 1u << (2 * level);
    for (unsigned b = 0; b < nbox; b++) {
        int x, y;
        box_xy(level, b, &x, &y);
        box_neighbors(level, x, y, neighbors + 8 * b);
        box_interactions(level, x, y, interactions + 27 * b);
    }
}

/* Box.cu:37-38 box centre in leaf-cell units (computed in double, stored float) */
static inline float box_centre(int limit, int level, int x) {
    unsigned w = (unsigned)powf(2, (float)(limit - level));
    return (float)(w * (unsigned)x + (w - 1) * 0.5);
}

/* ------------------------------------------------------------------------ */
/* fmm_calc.cu:20-289 fmm_bfs (CPU branch, use_gpu = 0) + fmm_calc.cu:295-408 */
/* The reference tree-code: for each charge layer zc, walk levels 2..logN in  */
/* BFS (index) order; P2M from raw charges at every level; evaluate each      */
/* non-empty box's multipole at every cell of its interaction list in every  */
/* target layer (M2P); at the 1-cell leaves add the near-field P2P.          */
/* ------------------------------------------------------------------------ */
/* one source box b of `level` (charge layer zc) added into potential */
static void or_fmm_box(const float *charge, float *potential, int limit, int P, int xdim, int ydim, int zdim,
                       float dx, float dy, float dz, int zc, int level, unsigned b, cx *mpc) {
    const int K = (P + 1) * (P + 1);
    float width = (float)pow(2, limit - level);
    int bx, by;
    box_xy(level, b, &bx, &by);
    float ncx = box_centre(limit, level, bx), ncy = box_centre(limit, level, by);
    float q = 0;
    memset(mpc, 0, sizeof(cx) * K);
    float charge_found = 0;
    int yy1 = (int)ceilf(ncy - width / 2), yy2 = (int)floorf(ncy + width / 2);
    for (int yy = yy1; yy <= yy2; yy++) {
        for (int xx = (int)ceilf(ncx - width / 2); xx <= (int)floorf(ncx + width / 2); xx++) {
            float dV = dx * dy * dz;
            q = charge[yy * xdim + xx] * dV;
            if (q != 0) {
                charge_found = 1;
                cx r_ = cx_make((xx - ncx) * dx, (yy - ncy) * dy);
                float rmag = sqrtf(r_.re * r_.re + r_.im * r_.im);
                float rang = atan2f(r_.im, r_.re);
                for (int l = 0; l <= P; l++)
                    for (int m = -l; m <= l; m++) {
                        cx sph = or_sph(l, m, (float)(M_PI / 2), rang);
                        sph.im = -sph.im; /* conjugate */
                        cx_scaleeq(&sph, (float)((double)q * pow((double)rmag, (double)l)));
                        cx_addeq(&mpc[l * l + l + m], sph);
                    }
            }
        }
    }
    if (!charge_found) return;

    int il[27];
    box_interactions(level, bx, by, il);
    for (int i = 0; i < 27; i++) {
        if (il[i] < 0) continue;
        int ix, iy;
        box_xy(level, (unsigned)il[i], &ix, &iy);
        float icx = box_centre(limit, level, ix), icy = box_centre(limit, level, iy);
        for (int yy = (int)ceilf(icy - width / 2); yy <= (int)floorf(icy + width / 2); yy++)
            for (int xx = (int)ceilf(icx - width / 2); xx <= (int)floorf(icx + width / 2); xx++)
                for (int zp = 0; zp < zdim; zp++) {
                    v3 r = v3_make((xx - ncx) * dx, (yy - ncy) * dy, (zp - zc) * dz);
                    float colat = v3_colat(r), azim = v3_azim(r), rmag = v3_mag(r);
                    cx sum_lm = cx_make(0, 0);
                    for (int l = 0; l <= P; l++) {
                        cx sum_m = cx_make(0, 0);
                        for (int m = -l; m <= l; m++) {
                            cx sph = or_sph(l, m, colat, azim);
                            int am = abs(m);
                            cx_scaleeq(&sph, (float)((1.0 * or_factorial(l - am)) / or_factorial(l + am)));
                            cx_muleq(&sph, mpc[l * l + l + m]);
                            cx_addeq(&sum_m, sph);
                        }
                        cx_scaleeq(&sum_m, (float)(1 / pow((double)rmag, (double)(l + 1))));
                        cx_addeq(&sum_lm, sum_m);
                    }
                    potential[zp * ydim * xdim + yy * xdim + xx] += sum_lm.re;
                }
    }

    if (level == limit) {
        int nb[8];
        box_neighbors(level, bx, by, nb);
        for (int zp = 0; zp < zdim; zp++) {
            if (zp != zc) {
                v3 r = v3_make(0, 0, (zp - zc) * dz);
                potential[zp * ydim * xdim + (int)(ncy * xdim + ncx)] += q / v3_mag(r);
            }
            for (int i = 0; i < 8; i++) {
                if (nb[i] < 0) continue;
                int ix, iy;
                box_xy(level, (unsigned)nb[i], &ix, &iy);
                float icx = box_centre(limit, level, ix), icy = box_centre(limit, level, iy);
                v3 r = v3_make((icx - ncx) * dx, (icy - ncy) * dy, (zp - zc) * dz);
                potential[zp * ydim * xdim + (int)(icy * xdim + icx)] += q / v3_mag(r);
            }
        }
    }
}

/* The BFS walk over levels lev_lo..lev_hi, every `stride`-th box, of charge
 * layer zc.  nthreads == 1 keeps the reference's serial order; more threads
 * split the boxes (each with its own potential buffer, added at the end),
 * which changes only the order of the float additions (used to time it). */
static int or_fmm_levels(const float *charge, float *potential, int limit, int P, int xdim, int ydim, int zdim,
                         float dx, float dy, float dz, int zc, int lev_lo, int lev_hi, int stride, int nthreads) {
    const int K = (P + 1) * (P + 1);
    const size_t n = (size_t)zdim * ydim * xdim;
    if (nthreads <= 1) {
        cx *mpc = (cx *)malloc(sizeof(cx) * K);
        if (!mpc) return 1;
        for (int level = lev_lo; level <= lev_hi; level++) {
            unsigned nbox = 1u << (2 * level);
            for (unsigned b = 0; b < nbox; b += (unsigned)stride)
                or_fmm_box(charge, potential, limit, P, xdim, ydim, zdim, dx, dy, dz, zc, level, b, mpc);
        }
        free(mpc);
        return 0;
    }
    int fail = 0;
#pragma omp parallel num_threads(nthreads)
    {
        cx *mpc = (cx *)malloc(sizeof(cx) * K);
        float *pot = (float *)calloc(n, sizeof(float));
        if (!mpc || !pot) {
#pragma omp atomic write
            fail = 1;
        }
        for (int level = lev_lo; level <= lev_hi; level++) {
            long nbox = 1L << (2 * level);
#pragma omp for schedule(dynamic, 1)
            for (long b = 0; b < nbox; b += stride)
                if (mpc && pot)
                    or_fmm_box(charge, pot, limit, P, xdim, ydim, zdim, dx, dy, dz, zc, level, (unsigned)b, mpc);
        }
#pragma omp critical
        if (pot)
            for (size_t i = 0; i < n; i++) potential[i] += pot[i];
        free(mpc);
        free(pot);
    }
    return fail;
}

/* fmm_calc.cu:295-408.  Requires xdim == ydim == 2^logN (fmm_calc.cu:305).
 * Returns 1 on an unsupported shape (the reference would index out of range). */
int or_fmm_calc(const float *charge, float *potential, int xdim, int ydim, int zdim,
                float dx, float dy, float dz, int P) {
    unsigned logN = (unsigned)ceil(log2f((float)(xdim * ydim)) / log2f(4));
    if (xdim != ydim || (1 << logN) != xdim || P < 0 || P > 4) return 1;
    for (int i = 0; i < zdim * ydim * xdim; i++) potential[i] = 0;
    for (int zc = 0; zc < zdim; zc++)
        if (or_fmm_levels(charge + zc * ydim * xdim, potential, (int)logN, P, xdim, ydim, zdim, dx, dy, dz, zc, 2,
                          (int)logN, 1, 1))
            return 1;
    return 0;
}

/* The same walk restricted to one charge layer zc, levels lev_lo..lev_hi and
 * every `stride`-th box, on `nthreads` threads, added into potential: the
 * bounded sample of the CPU baseline (stride 1 over every level and layer in
 * one thread is or_fmm_calc exactly). */
int or_fmm_bfs_sample(const float *charge, float *potential, int xdim, int ydim, int zdim,
                      float dx, float dy, float dz, int P, int zc, int lev_lo, int lev_hi, int stride,
                      int nthreads) {
    unsigned logN = (unsigned)ceil(log2f((float)(xdim * ydim)) / log2f(4));
    if (xdim != ydim || (1 << logN) != xdim || P < 0 || P > 4 || lev_lo < 2 || lev_hi > (int)logN ||
        lev_lo > lev_hi || stride < 1 || zc < 0 || zc >= zdim)
        return 1;
    return or_fmm_levels(charge + (size_t)zc * ydim * xdim, potential, (int)logN, P, xdim, ydim, zdim, dx, dy, dz,
                         zc, lev_lo, lev_hi, stride, nthreads);
}

/* ------------------------------------------------------------------------ */
/* ode_functions.cpp:11-18 LLG_Mprime (Landau-Lifshitz form, M in A/m)        */
/* ------------------------------------------------------------------------ */
static inline v3 llg(v3 M, v3 H, float alfa, float gamma, float Ms) {
    v3 MxH = v3_cross(M, H);
    return v3_sub(v3_scale(MxH, -gamma), v3_scale(v3_cross(M, MxH), alfa * gamma / Ms));
}
void or_llg_mprime(const float *M, const float *H, float alfa, float gamma, float Ms, float *out) {
    v3 r = llg(v3_make(M[0], M[1], M[2]), v3_make(H[0], H[1], H[2]), alfa, gamma, Ms);
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}

/* external field: the reference ramp (ode_functions.cpp:61-66) when hext is
 * NULL, else a constant vector in A/m (the dot film's and the sweep's uniform fields) */
static v3 ext_field(float t, float mu_0, const float *hext) {
    if (hext) return v3_make(hext[0], hext[1], hext[2]);
    double amp = (t < 1e-9) ? (-0.3 / 1e-9 * (double)t + 0.3) : 0.0;
    return v3_scale(v3_make(1, 0, 0), (float)((double)(1 / mu_0) * amp));
}

/* potential solver selector for or_hfield: 0 = exact FP32 (reference CPU
 * branch), 1 = reference tree-code (P <= 4) */
/* ode_functions.cpp:21-70 Hfield */
int or_hfield(const float *M, float *Hout, float *charge, float *potential, float t,
              int xdim, int ydim, int zdim, float dx, float dy, float dz, int P,
              float mu_0, float Ms, float Aexch, int demag, int exchange, int external,
              int use_fmm, const float *hext) {
    int status = 0;
    const int n = zdim * ydim * xdim;
    v3 *H = (v3 *)Hout;
    for (int i = 0; i < n; i++) H[i] = v3_make(0, 0, 0);
    if (demag) {
        or_divergence_3d(M, xdim, ydim, zdim, dx, dy, dz, charge);
        if (use_fmm) status |= or_fmm_calc(charge, potential, xdim, ydim, zdim, dx, dy, dz, P);
        else status |= or_calc_potential_exact(charge, xdim, ydim, zdim, dx, dy, dz, potential);
        or_gradient_3d(potential, xdim, ydim, zdim, dx, dy, dz, (float)(1 / (4.0 * M_PI)), Hout);
    }
    if (exchange) or_add_exchange_field(M, Ms, Aexch, mu_0, xdim, ydim, zdim, dx, dy, dz, Hout);
    if (external) {
        v3 he = ext_field(t, mu_0, hext);
        for (int i = 0; i < n; i++) H[i] = v3_add(H[i], he);
    }
    return status ? 1 : 0;
}

/* ode_functions.cpp:76-148 rk4_step, FP32 throughout as in the reference.
 * combine64 != 0 selects this project's precision split instead: stage
 * combination, accumulation and renormalisation in double (the slope and
 * stage magnetisations handed to Hfield stay float). */
int or_rk4_step(float t, float dt, float *t2, const unsigned char *material,
                const float *Min, float *M2out, float *Hio, float *slope_io, float *this_slope_io,
                float *charge, float *potential, int xdim, int ydim, int zdim,
                float dx, float dy, float dz, int P, float mu_0, float Ms, float Aexch,
                float alfa, float gamma, int demag, int exchange, int external, int use_fmm,
                int normalize, const float *hext, int combine64) {
    const int n = zdim * ydim * xdim;
    const v3 *M = (const v3 *)Min;
    v3 *M2 = (v3 *)M2out, *H = (v3 *)Hio, *slope = (v3 *)slope_io, *ts = (v3 *)this_slope_io;
    double *acc = NULL;
    if (combine64) {
        acc = (double *)calloc((size_t)3 * n, sizeof(double));
        if (!acc) return 1;
    }
    for (int i = 0; i < n; i++) slope[i] = v3_make(0, 0, 0);
    static const float stage_half[3] = {0.5f, 0.5f, 1.0f};
    static const float weight[4] = {1.0f, 2.0f, 2.0f, 1.0f};
    for (int s = 0; s < 4; s++) {
        if (s > 0)
            or_hfield(M2out, Hio, charge, potential, t, xdim, ydim, zdim, dx, dy, dz, P, mu_0, Ms,
                      Aexch, demag, exchange, external, use_fmm, hext);
        for (int i = 0; i < n; i++) {
            if (!material[i]) continue;
            ts[i] = llg(s == 0 ? M[i] : M2[i], H[i], alfa, gamma, Ms);
            if (!combine64) {
                if (s < 3) {
                    M2[i] = v3_add(M[i], v3_scale(v3_scale(ts[i], stage_half[s]), dt));
                    slope[i] = v3_add(slope[i], s == 0 ? ts[i] : v3_scale(ts[i], weight[s]));
                } else {
                    slope[i] = v3_add(slope[i], ts[i]);
                    M2[i] = v3_add(M[i], v3_scale(v3_scale(slope[i], dt), (float)(1 / 6.0)));
                    if (normalize) M2[i] = v3_scale(M2[i], Ms / v3_mag(M2[i]));
                }
            } else {
                double k[3] = {ts[i].x, ts[i].y, ts[i].z};
                double m0[3] = {M[i].x, M[i].y, M[i].z};
                double *a = acc + 3 * (size_t)i;
                for (int c = 0; c < 3; c++) a[c] += (double)weight[s] * k[c];
                slope[i] = v3_make((float)a[0], (float)a[1], (float)a[2]);
                double mn[3];
                if (s < 3) {
                    double h = (s < 2 ? 0.5 : 1.0) * (double)dt;
                    for (int c = 0; c < 3; c++) mn[c] = m0[c] + h * k[c];
                } else {
                    for (int c = 0; c < 3; c++) mn[c] = m0[c] + (double)dt / 6.0 * a[c];
                    if (normalize) {
                        double mag = sqrt(mn[0] * mn[0] + mn[1] * mn[1] + mn[2] * mn[2]);
                        for (int c = 0; c < 3; c++) mn[c] *= (double)Ms / mag;
                    }
                }
                M2[i] = v3_make((float)mn[0], (float)mn[1], (float)mn[2]);
            }
        }
    }
    if (acc) free(acc);
    *t2 = t + dt;
    return 0;
}

/* Energy and torque of ode_functions.cpp:234-240,268-278 (serial reductions
 * in float, as the reference) */
void or_energy_torque(const unsigned char *material, const float *Min, const float *Hin, int n,
                      float mu_0, float Ms, float dx, float dy, float dz,
                      float *E, float *torque) {
    const v3 *M = (const v3 *)Min, *H = (const v3 *)Hin;
    float e = 0, tq = 0;
    for (int i = 0; i < n; i++) {
        if (!material[i]) continue;
        e += v3_dot(M[i], H[i]);
        float this_t = v3_mag(v3_cross(M[i], H[i]));
        tq = (tq > this_t) ? tq : this_t;
    }
    e = (float)((double)e * (-0.5 * mu_0 * (dx * dy * dz) / 1.6e-19));
    tq *= (1 / Ms / Ms);
    *E = e;
    *torque = tq;
}

int or_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
