// stencil_llg.hip -- HBM-bound per-cell kernels of the RK4 LLG step on gfx950:
//   div_kernel     rho = div M                          (vector_functions.cpp:8-28)
//   stage_kernel   H = grad(phi)/4pi + exchange + H_ext (vector_functions.cpp:31-103,
//                  ode_functions.cpp:21-70), dM/dt (ode_functions.cpp:11-18) and the
//                  RK4 stage combine (ode_functions.cpp:76-148) fused in one pass,
//                  stage combination and |M| = Ms renormalisation in FP64.
//   hfield_kernel  H only (for the Hfield entry point and diagnostics)
//   energy_kernel  sum M.H and max |M x H| (ode_functions.cpp:234-240,268-278)
//
// Layout: M stage values are FP32 SoA planes [3][n] (x plane, y plane, z plane),
// the RK4 base M_n and the slope accumulator are FP64 SoA [3][n].  Each thread
// owns one (x,y) column and walks z, keeping the z-1/z/z+1 values in registers;
// in-plane neighbours come from the caches (blocks are 64 x 4 columns, or
// 256 x 8 for the four-cells-per-thread stage kernel, so the y-neighbour rows
// of a block are resident in L1/L2).
// Out-of-grid neighbours read as 0 (the reference's zero padding: smeared
// surface charge, Dirichlet-zero exchange).
#include <cstdlib>
#include "mms_internal.hpp"

namespace mms {

namespace {

constexpr int BX = 64, BY = 4;
// rows per block of the four-cells-per-thread stage kernel (measured on the
// 1024 x 1024 x 4 film: 2, 4, 8 and 16 rows of 64 threads gave 54.9, 54.9,
// 51.6 and 58.6 us per launch; 128 x 4 and 32 x 16 threads 53.1 and 52.9)
constexpr int BY4 = 8;

// XCD-aware tile order for a 1D grid of 8 * ceil(ntiles / 8) workgroups.
// Workgroup i runs on XCD i mod 8 (round-robin dispatch); it takes logical
// tile (i mod 8) * per + i / 8, so each XCD walks one contiguous band of
// tile rows in order and the y-halo rows of a tile are still in that XCD's
// L2 from the tile row above.  Returns false for the padding workgroups.
__device__ __forceinline__ bool xcd_tile(int ntx, int ntiles, int &tx, int &ty) {
    const int per = (ntiles + 7) >> 3;
    const int i = blockIdx.x;
    const int t = (i & 7) * per + (i >> 3);
    if (t >= ntiles) return false;
    tx = t % ntx;
    ty = t / ntx;
    return true;
}

// Batched films: a tile is 256 consecutive (x, y) columns of one film in
// row-major order, starting at column t.y, so only each film's last tile has
// idle threads; out-of-film threads get x = xd (no work).
__device__ __forceinline__ void batch_column(const int4 &t, const BatchFilm &f, int &x, int &y) {
    const int col = t.y + (int)(threadIdx.y * BX + threadIdx.x);
    const bool ok = col < f.xd * f.yd;
    y = ok ? col / f.xd : 0;
    x = ok ? col - y * f.xd : f.xd;
}

template <int ZD>
__device__ __forceinline__ void div_column(const float *__restrict__ m, float *__restrict__ rho, int x, int y, int xd,
                                           int yd, int zd_rt, long n, float dx, float dy, float dz) {
    const int zd = ZD ? ZD : zd_rt;
    if (x >= xd || y >= yd) return;
    const long plane = (long)xd * yd;
    const float *mx = m, *my = m + n, *mz = m + 2 * n;
    const long c0 = (long)y * xd + x;
    float zprev = 0.0f, zcur = mz[c0];
#pragma unroll
    for (int z = 0; z < zd; z++) {
        const long c = (long)z * plane + c0;
        const float znext = (z != zd - 1) ? mz[c + plane] : 0.0f;
        const float x2 = (x != xd - 1) ? mx[c + 1] : 0.0f;
        const float x0 = (x != 0) ? mx[c - 1] : 0.0f;
        const float y2 = (y != yd - 1) ? my[c + xd] : 0.0f;
        const float y0 = (y != 0) ? my[c - xd] : 0.0f;
        rho[c] = ((x2 - x0) / dx + (y2 - y0) / dy + (znext - zprev) / dz) * 0.5f;
        zprev = zcur;
        zcur = znext;
    }
}

template <int ZD>
__global__ __launch_bounds__(BX * BY) void div_kernel(const float *__restrict__ m, float *__restrict__ rho, int xd,
                                                      int yd, int zd, float dx, float dy, float dz) {
    const int ntx = (xd + BX - 1) / BX, nty = (yd + BY - 1) / BY;
    int tx, ty;
    if (!xcd_tile(ntx, ntx * nty, tx, ty)) return;
    div_column<ZD>(m, rho, tx * BX + threadIdx.x, ty * BY + threadIdx.y, xd, yd, zd, (long)xd * yd * zd, dx, dy, dz);
}

// Four consecutive x cells per thread (xd a multiple of 4): 16-byte loads and
// stores, each cell by div_column's expression, so the charges are bitwise
// the same.  The tile is 4 BX x BY cells.
template <int ZD>
__global__ __launch_bounds__(BX * BY) void div4_kernel(const float *__restrict__ m, float *__restrict__ rho, int xd,
                                                       int yd, int zd_rt, float dx, float dy, float dz) {
    typedef float f4 __attribute__((ext_vector_type(4)));
    const int ntx = (xd + 4 * BX - 1) / (4 * BX), nty = (yd + BY - 1) / BY;
    int tx, ty;
    if (!xcd_tile(ntx, ntx * nty, tx, ty)) return;
    const int x = 4 * (tx * BX + threadIdx.x), y = ty * BY + threadIdx.y;
    if (x >= xd || y >= yd) return;
    const int zd = ZD ? ZD : zd_rt;
    const long plane = (long)xd * yd, n = plane * zd;
    const float *mx = m, *my = m + n, *mz = m + 2 * n;
    const long c0 = (long)y * xd + x;
    const bool xl = (x != 0), xr = (x + 4 != xd), yu = (y != yd - 1), yl = (y != 0);
    const f4 zero = f4{0.f, 0.f, 0.f, 0.f};
    f4 zprev = zero, zcur = *reinterpret_cast<const f4 *>(mz + c0);
#pragma unroll
    for (int z = 0; z < zd; z++) {
        const long c = (long)z * plane + c0;
        const f4 znext = (z != zd - 1) ? *reinterpret_cast<const f4 *>(mz + c + plane) : zero;
        const f4 xc = *reinterpret_cast<const f4 *>(mx + c);
        const float xm = xl ? mx[c - 1] : 0.0f, xp = xr ? mx[c + 4] : 0.0f;
        const f4 y2 = yu ? *reinterpret_cast<const f4 *>(my + c + xd) : zero;
        const f4 y0 = yl ? *reinterpret_cast<const f4 *>(my + c - xd) : zero;
        const float x2[4] = {xc[1], xc[2], xc[3], xp}, x0[4] = {xm, xc[0], xc[1], xc[2]};
        f4 r;
#pragma unroll
        for (int i = 0; i < 4; i++) r[i] = ((x2[i] - x0[i]) / dx + (y2[i] - y0[i]) / dy + (znext[i] - zprev[i]) / dz) * 0.5f;
        *reinterpret_cast<f4 *>(rho + c) = r;
        zprev = zcur;
        zcur = znext;
    }
}

template <int ZD>
__global__ __launch_bounds__(BX * BY) void div_batch_kernel(const float *__restrict__ m, float *__restrict__ rho,
                                                            const int4 *__restrict__ tiles,
                                                            const BatchFilm *__restrict__ films, long pstride,
                                                            float dx, float dy, float dz) {
    const int4 t = tiles[blockIdx.x];
    if (t.x < 0) return;  // padding slot of the XCD-ordered tile list
    const BatchFilm f = films[t.x];
    int x, y;
    batch_column(t, f, x, y);
    div_column<ZD>(m + f.off, rho + f.off, x, y, f.xd, f.yd, f.zd, pstride, dx, dy, dz);
}


struct StageArgs {
    const float *phi;      // potential (null when demag off or h_given)
    const float *m_in;     // FP32 SoA stage magnetisation
    float *m_out;          // FP32 SoA next stage magnetisation
    double *mn;            // FP64 SoA base M_n (updated at stage 3)
    const float *m_base;   // FP32 SoA copy of M_n for the stage values of stages 1, 2
                           // (stage 0: the input itself; null: the FP64 base)
    float *acc;            // FP32 SoA slope accumulator
    double *mn_out;        // FP64 SoA destination of M_{n+1} at stage 3 (null = mn)
    const uint8_t *matl;   // material flags (null = all magnetic)
    const double *t_dev;   // simulation time (device)
    const float *m_keep;   // FP32 SoA values of the non-material cells' output (null: copy the input)
    const float *h_given;  // AoS field to use instead of computing it (stage 0 of rk4_step)
    float *h_out;          // optional AoS output of the field used
    float *k_out;          // optional AoS output of this stage's dM/dt
    float *slope_out;      // optional AoS output of the float slope sum
    int xd, yd, zd;
    float dx, dy, dz;
    float inv4pi;          // (float)(1/(4 pi))
    float exch_cm;         // 2 A / (mu0 Ms^2)
    int demag, exchange, external;
    float hx, hy, hz;      // constant external field (A/m), external == 2
    float inv_mu0;         // (float)(1/mu0), external == 1
    float alfa, gamma, Ms;
    int stage;             // 0..3
    double dt;
    int normalize;
};

// External field of the reference ramp (ode_functions.cpp:61): the ramp is
// evaluated in double and only the scalar multiplier is rounded to float.
__device__ inline float ramp_hx(double t, float inv_mu0) {
    double amp = (t < 1e-9) ? (-0.3 / 1e-9 * t + 0.3) : 0.0;
    return (float)((double)inv_mu0 * amp);
}

__device__ inline void llg_rhs(float Mx, float My, float Mz, float Hx, float Hy, float Hz,
                               float alfa, float gamma, float Ms, float &kx, float &ky, float &kz) {
    // dM/dt = -gamma M x H - (alfa gamma / Ms) M x (M x H)   (ode_functions.cpp:11-18)
    const float cx = My * Hz - Mz * Hy, cy = Mz * Hx - Mx * Hz, cz = Mx * Hy - My * Hx;
    const float ddx = My * cz - Mz * cy, ddy = Mz * cx - Mx * cz, ddz = Mx * cy - My * cx;
    const float a = alfa * gamma / Ms;
    kx = cx * (-gamma) - ddx * a;
    ky = cy * (-gamma) - ddy * a;
    kz = cz * (-gamma) - ddz * a;
}

// One (x, y) column of one film: film extent xd x yd x zd, component planes
// `n` cells apart; the arrays start at the film's first cell.  The arrays are
// distinct buffers (__restrict__), so the loads of the next z level can be
// issued ahead of the current level's stores.  ST >= 0: the plain resident
// step at stage ST (field computed here, no extra outputs, M_{n+1} written
// over the base), compiled without the per-stage branches; ST = -1: every
// option at run time.
template <int ZD, int ST>
__device__ __forceinline__ void stage_column_r(const StageArgs &A, const float *__restrict__ m_in_,
                                               float *__restrict__ m_out_, double *__restrict__ mn_,
                                               float *__restrict__ acc_, double *__restrict__ mn_out_,
                                               const float *__restrict__ phi_, const uint8_t *__restrict__ matl_,
                                               const float *__restrict__ mb_, const float *__restrict__ mk_,
                                               const float *__restrict__ h_given_,
                                               float *__restrict__ h_out_, float *__restrict__ k_out_,
                                               float *__restrict__ slope_out_, int x, int y, int xd, int yd,
                                               int zd_rt, long n, float hx, float hy, float hz) {
    const int zd = ZD ? ZD : zd_rt;
    const long plane = (long)xd * yd;
    const long c0 = (long)y * xd + x;
    const float *mx = m_in_, *my = m_in_ + n, *mz = m_in_ + 2 * n;
    const double t = A.t_dev ? *A.t_dev : 0.0;
    float hex = hx, hey = hy, hez = hz;
    if (A.external == 1) { hex = ramp_hx(t, A.inv_mu0); hey = 0.0f; hez = 0.0f; }
    // reciprocals of the uniform denominators, once per thread: a multiply
    // per term instead of an IEEE division (within an ulp of the divided form)
    const float rdx2 = 1.0f / (A.dx * A.dx), rdy2 = 1.0f / (A.dy * A.dy), rdz2 = 1.0f / (A.dz * A.dz);
    const float gx = A.inv4pi / (2.0f * A.dx), gy = A.inv4pi / (2.0f * A.dy), gz = A.inv4pi / (2.0f * A.dz);
    constexpr bool PLAIN = ST >= 0;
    const int stage = PLAIN ? ST : A.stage;
    if constexpr (PLAIN) {
        h_given_ = nullptr; h_out_ = nullptr; k_out_ = nullptr; slope_out_ = nullptr; mn_out_ = nullptr; mk_ = nullptr;
    }
    const bool need_field = (h_given_ == nullptr);
    const bool xr = (x != xd - 1), xl = (x != 0), yu = (y != yd - 1), yl = (y != 0);
    const int ixr = xr ? 1 : 0, ixl = xl ? 1 : 0;
    const long iyu = yu ? xd : 0, iyl = yl ? xd : 0;

    // z-rotation registers
    float pprev = 0.0f, pcur = 0.0f;
    float mxp = 0.0f, myp = 0.0f, mzp = 0.0f;
    float mxc = mx[c0], myc = my[c0], mzc = mz[c0];
    if (need_field && A.demag) pcur = phi_[c0];

#pragma unroll
    for (int z = 0; z < zd; z++) {
        const long c = (long)z * plane + c0;
        const bool ztop = (z != zd - 1);
        float mxn = 0.0f, myn = 0.0f, mzn = 0.0f, pnext = 0.0f;
        if (ztop) { mxn = mx[c + plane]; myn = my[c + plane]; mzn = mz[c + plane]; }
        float Hx, Hy, Hz;
        if (need_field) {
            Hx = 0.0f; Hy = 0.0f; Hz = 0.0f;
            if (A.demag) {
                // neighbours outside the film read the cell itself and are
                // then replaced by 0: unconditional loads, no divergent branches
                if (ztop) pnext = phi_[c + plane];
                const float x2 = xr ? phi_[c + ixr] : 0.0f;
                const float x0 = xl ? phi_[c - ixl] : 0.0f;
                const float y2 = yu ? phi_[c + iyu] : 0.0f;
                const float y0 = yl ? phi_[c - iyl] : 0.0f;
                Hx = (x2 - x0) * gx;
                Hy = (y2 - y0) * gy;
                Hz = (pnext - pprev) * gz;
            }
            if (A.exchange) {
                float Lx, Ly, Lz;
                {
                    const float xa = xr ? mx[c + ixr] : 0.0f, xb = xl ? mx[c - ixl] : 0.0f;
                    const float ya = yu ? mx[c + iyu] : 0.0f, yb = yl ? mx[c - iyl] : 0.0f;
                    Lx = (xa + xb - 2.0f * mxc) * rdx2 + (ya + yb - 2.0f * mxc) * rdy2 + (mxn + mxp - 2.0f * mxc) * rdz2;
                }
                {
                    const float xa = xr ? my[c + ixr] : 0.0f, xb = xl ? my[c - ixl] : 0.0f;
                    const float ya = yu ? my[c + iyu] : 0.0f, yb = yl ? my[c - iyl] : 0.0f;
                    Ly = (xa + xb - 2.0f * myc) * rdx2 + (ya + yb - 2.0f * myc) * rdy2 + (myn + myp - 2.0f * myc) * rdz2;
                }
                {
                    const float xa = xr ? mz[c + ixr] : 0.0f, xb = xl ? mz[c - ixl] : 0.0f;
                    const float ya = yu ? mz[c + iyu] : 0.0f, yb = yl ? mz[c - iyl] : 0.0f;
                    Lz = (xa + xb - 2.0f * mzc) * rdx2 + (ya + yb - 2.0f * mzc) * rdy2 + (mzn + mzp - 2.0f * mzc) * rdz2;
                }
                Hx += A.exch_cm * Lx; Hy += A.exch_cm * Ly; Hz += A.exch_cm * Lz;
            }
            if (A.external) { Hx += hex; Hy += hey; Hz += hez; }
        } else {
            Hx = h_given_[3 * c]; Hy = h_given_[3 * c + 1]; Hz = h_given_[3 * c + 2];
        }
        if (h_out_) { h_out_[3 * c] = Hx; h_out_[3 * c + 1] = Hy; h_out_[3 * c + 2] = Hz; }

        const bool magnetic = matl_ ? (matl_[c] != 0) : true;
        float *ox = m_out_, *oy = m_out_ + n, *oz = m_out_ + 2 * n;
        if (magnetic) {
            float kx, ky, kz;
            llg_rhs(mxc, myc, mzc, Hx, Hy, Hz, A.alfa, A.gamma, A.Ms, kx, ky, kz);
            if (k_out_) { k_out_[3 * c] = kx; k_out_[3 * c + 1] = ky; k_out_[3 * c + 2] = kz; }
            // the FP64 base and the slope accumulator are touched once per
            // stage: non-temporal loads and stores keep them from evicting
            // the neighbour-reused FP32 m and phi from the caches.  The slope
            // k1 + 2 k2 + 2 k3 + k4 is summed in FP32 as in the reference
            // (ode_functions.cpp:105-140); the combine M + dt/6 slope and the
            // renormalisation are FP64.
            // Stages 1-3 round their stage value M_n + h k to FP32, so they
            // take M_n from its FP32 copy (the stage-0 input, 12 bytes
            // instead of 24); the final combine at stage 4 reads the FP64 base.
            double *bx = mn_, *by = mn_ + n, *bz = mn_ + 2 * n;
            float *ax = acc_, *ay = acc_ + n, *az = acc_ + 2 * n;
            double b0, b1, b2;
            if (stage == 3 || !mb_) {
                b0 = __builtin_nontemporal_load(&bx[c]); b1 = __builtin_nontemporal_load(&by[c]);
                b2 = __builtin_nontemporal_load(&bz[c]);
            } else if (stage == 0) {
                b0 = mxc; b1 = myc; b2 = mzc;
            } else {
                b0 = __builtin_nontemporal_load(&mb_[c]); b1 = __builtin_nontemporal_load(&mb_[n + c]);
                b2 = __builtin_nontemporal_load(&mb_[2 * n + c]);
            }
            float s0, s1, s2;
            if (stage == 0) { s0 = kx; s1 = ky; s2 = kz; }
            else { const float wk = (stage == 3) ? 1.0f : 2.0f;  // w k is exact
                   s0 = __builtin_nontemporal_load(&ax[c]) + wk * kx;
                   s1 = __builtin_nontemporal_load(&ay[c]) + wk * ky;
                   s2 = __builtin_nontemporal_load(&az[c]) + wk * kz; }
            if (slope_out_) { slope_out_[3 * c] = s0; slope_out_[3 * c + 1] = s1; slope_out_[3 * c + 2] = s2; }
            double n0, n1, n2;
            if (stage < 3) {
                const double h = (stage < 2 ? 0.5 : 1.0) * A.dt;
                n0 = b0 + h * (double)kx; n1 = b1 + h * (double)ky; n2 = b2 + h * (double)kz;
                __builtin_nontemporal_store(s0, &ax[c]);
                __builtin_nontemporal_store(s1, &ay[c]);
                __builtin_nontemporal_store(s2, &az[c]);
            } else {
                const double h = A.dt / 6.0;
                n0 = b0 + h * (double)s0; n1 = b1 + h * (double)s1; n2 = b2 + h * (double)s2;
                if (A.normalize) {
                    const double f = (double)A.Ms / sqrt(n0 * n0 + n1 * n1 + n2 * n2);
                    n0 *= f; n1 *= f; n2 *= f;
                }
                double *o64 = mn_out_ ? mn_out_ : mn_;
                __builtin_nontemporal_store(n0, &o64[c]);
                __builtin_nontemporal_store(n1, &o64[n + c]);
                __builtin_nontemporal_store(n2, &o64[2 * n + c]);
            }
            ox[c] = (float)n0; oy[c] = (float)n1; oz[c] = (float)n2;
        } else {
            // non-material cells: the reference leaves M2 and this_slope as
            // the caller passed them and the slope sum at 0
            // (ode_functions.cpp:91-95,102)
            if (mk_) { ox[c] = mk_[c]; oy[c] = mk_[n + c]; oz[c] = mk_[2 * n + c]; }
            else { ox[c] = mxc; oy[c] = myc; oz[c] = mzc; }
            if (slope_out_) { slope_out_[3 * c] = 0.0f; slope_out_[3 * c + 1] = 0.0f; slope_out_[3 * c + 2] = 0.0f; }
        }
        pprev = pcur; pcur = pnext;
        mxp = mxc; myp = myc; mzp = mzc;
        mxc = mxn; myc = myn; mzc = mzn;
    }
}

template <int ZD, int ST>
__device__ __forceinline__ void stage_column(const StageArgs &A, int x, int y, int xd, int yd, int zd_rt, long off,
                                             long n, float hx, float hy, float hz) {
    if (x >= xd || y >= yd) return;
    stage_column_r<ZD, ST>(A, A.m_in + off, A.m_out + off, A.mn + off, A.acc + off, A.mn_out ? A.mn_out + off : nullptr,
                       A.phi ? A.phi + off : nullptr, A.matl ? A.matl + off : nullptr,
                       A.m_base ? A.m_base + off : nullptr, A.m_keep ? A.m_keep + off : nullptr, A.h_given, A.h_out,
                       A.k_out, A.slope_out, x, y, xd, yd,
                       zd_rt, n, hx, hy, hz);
}

template <int ZD, int ST>
__global__ __launch_bounds__(BX * BY) void stage_kernel(StageArgs A) {
    const int ntx = (A.xd + BX - 1) / BX, nty = (A.yd + BY - 1) / BY;
    int tx, ty;
    if (!xcd_tile(ntx, ntx * nty, tx, ty)) return;
    stage_column<ZD, ST>(A, tx * BX + threadIdx.x, ty * BY + threadIdx.y, A.xd, A.yd, A.zd, 0, (long)A.xd * A.yd * A.zd,
                     A.hx, A.hy, A.hz);
}

// The plain resident step at stage ST with four consecutive x cells per
// thread (xd a multiple of 4, 16-byte aligned arrays): 16-byte loads and
// stores, and each cell by stage_column_r's expressions in the same order,
// so the results are bitwise the same.  A group of four with a non-magnetic
// cell stores its magnetic cells one by one and leaves the others' slope
// and FP64 base untouched, as the one-cell form does.
template <int ZD, int ST>
__device__ __forceinline__ void stage_column4(const StageArgs &A, const float *__restrict__ m_in_,
                                              float *__restrict__ m_out_, double *__restrict__ mn_,
                                              float *__restrict__ acc_, const float *__restrict__ phi_,
                                              const uint8_t *__restrict__ matl_, const float *__restrict__ mb_,
                                              int x, int y, int xd, int yd, int zd_rt, long n) {
    typedef float f4 __attribute__((ext_vector_type(4)));
    typedef double d2 __attribute__((ext_vector_type(2)));
    constexpr int stage = ST;
    const int zd = ZD ? ZD : zd_rt;
    const long plane = (long)xd * yd;
    const long c0 = (long)y * xd + x;
    const float *mx = m_in_, *my = m_in_ + n, *mz = m_in_ + 2 * n;
    const double t = A.t_dev ? *A.t_dev : 0.0;
    float hex = A.hx, hey = A.hy, hez = A.hz;
    if (A.external == 1) { hex = ramp_hx(t, A.inv_mu0); hey = 0.0f; hez = 0.0f; }
    const float rdx2 = 1.0f / (A.dx * A.dx), rdy2 = 1.0f / (A.dy * A.dy), rdz2 = 1.0f / (A.dz * A.dz);
    const float gx = A.inv4pi / (2.0f * A.dx), gy = A.inv4pi / (2.0f * A.dy), gz = A.inv4pi / (2.0f * A.dz);
    const bool xr = (x + 4 != xd), xl = (x != 0), yu = (y != yd - 1), yl = (y != 0);
    const f4 zero = f4{0.f, 0.f, 0.f, 0.f};
    auto ld4 = [](const float *p) { return *reinterpret_cast<const f4 *>(p); };
    // the x neighbours of the four cells: cells x - 1 .. x + 4 of one row
    auto row6 = [&](const float *p, long c, float (&v)[6]) {
        const f4 cv = ld4(p + c);
        v[0] = xl ? p[c - 1] : 0.0f;
        v[1] = cv[0]; v[2] = cv[1]; v[3] = cv[2]; v[4] = cv[3];
        v[5] = xr ? p[c + 4] : 0.0f;
    };
    f4 pprev = zero, pcur = zero;
    f4 mxp = zero, myp = zero, mzp = zero;
    f4 mxc = ld4(mx + c0), myc = ld4(my + c0), mzc = ld4(mz + c0);
    if (A.demag) pcur = ld4(phi_ + c0);
#pragma unroll
    for (int z = 0; z < zd; z++) {
        const long c = (long)z * plane + c0;
        const bool ztop = (z != zd - 1);
        f4 mxn = zero, myn = zero, mzn = zero, pnext = zero;
        if (ztop) { mxn = ld4(mx + c + plane); myn = ld4(my + c + plane); mzn = ld4(mz + c + plane); }
        f4 Hx = zero, Hy = zero, Hz = zero;
        if (A.demag) {
            if (ztop) pnext = ld4(phi_ + c + plane);
            float pr[6];
            row6(phi_, c, pr);
            const f4 y2 = yu ? ld4(phi_ + c + xd) : zero, y0 = yl ? ld4(phi_ + c - xd) : zero;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                // the film's edge cells take 0 for the missing neighbour
                const float x2 = (i < 3 || xr) ? pr[i + 2] : 0.0f, x0 = (i > 0 || xl) ? pr[i] : 0.0f;
                Hx[i] = (x2 - x0) * gx;
                Hy[i] = (y2[i] - y0[i]) * gy;
                Hz[i] = (pnext[i] - pprev[i]) * gz;
            }
        }
        if (A.exchange) {
            float rx[6], ry[6], rz[6];
            row6(mx, c, rx); row6(my, c, ry); row6(mz, c, rz);
            const f4 xya = yu ? ld4(mx + c + xd) : zero, xyb = yl ? ld4(mx + c - xd) : zero;
            const f4 yya = yu ? ld4(my + c + xd) : zero, yyb = yl ? ld4(my + c - xd) : zero;
            const f4 zya = yu ? ld4(mz + c + xd) : zero, zyb = yl ? ld4(mz + c - xd) : zero;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                float Lx, Ly, Lz;
                {
                    const float xa = (i < 3 || xr) ? rx[i + 2] : 0.0f, xb = (i > 0 || xl) ? rx[i] : 0.0f;
                    Lx = (xa + xb - 2.0f * mxc[i]) * rdx2 + (xya[i] + xyb[i] - 2.0f * mxc[i]) * rdy2 +
                         (mxn[i] + mxp[i] - 2.0f * mxc[i]) * rdz2;
                }
                {
                    const float xa = (i < 3 || xr) ? ry[i + 2] : 0.0f, xb = (i > 0 || xl) ? ry[i] : 0.0f;
                    Ly = (xa + xb - 2.0f * myc[i]) * rdx2 + (yya[i] + yyb[i] - 2.0f * myc[i]) * rdy2 +
                         (myn[i] + myp[i] - 2.0f * myc[i]) * rdz2;
                }
                {
                    const float xa = (i < 3 || xr) ? rz[i + 2] : 0.0f, xb = (i > 0 || xl) ? rz[i] : 0.0f;
                    Lz = (xa + xb - 2.0f * mzc[i]) * rdx2 + (zya[i] + zyb[i] - 2.0f * mzc[i]) * rdy2 +
                         (mzn[i] + mzp[i] - 2.0f * mzc[i]) * rdz2;
                }
                Hx[i] += A.exch_cm * Lx; Hy[i] += A.exch_cm * Ly; Hz[i] += A.exch_cm * Lz;
            }
        }
        if (A.external) {
#pragma unroll
            for (int i = 0; i < 4; i++) { Hx[i] += hex; Hy[i] += hey; Hz[i] += hez; }
        }
        bool mag[4] = {true, true, true, true};
        if (matl_) {
            const uchar4 mt = *reinterpret_cast<const uchar4 *>(matl_ + c);
            mag[0] = mt.x != 0; mag[1] = mt.y != 0; mag[2] = mt.z != 0; mag[3] = mt.w != 0;
        }
        const bool all = mag[0] && mag[1] && mag[2] && mag[3];
        float *ox = m_out_, *oy = m_out_ + n, *oz = m_out_ + 2 * n;
        double *bx = mn_, *by = mn_ + n, *bz = mn_ + 2 * n;
        float *ax = acc_, *ay = acc_ + n, *az = acc_ + 2 * n;
        f4 kx, ky, kz;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            float k0, k1, k2;
            llg_rhs(mxc[i], myc[i], mzc[i], Hx[i], Hy[i], Hz[i], A.alfa, A.gamma, A.Ms, k0, k1, k2);
            kx[i] = k0; ky[i] = k1; kz[i] = k2;
        }
        // M_n: FP64 base at stage 3, FP32 copy (or this stage's input at stage 0) before
        double b0[4], b1[4], b2[4];
        if (stage == 3 || !mb_) {
            const d2 u0 = __builtin_nontemporal_load(reinterpret_cast<const d2 *>(bx + c));
            const d2 u1 = __builtin_nontemporal_load(reinterpret_cast<const d2 *>(bx + c + 2));
            const d2 v0 = __builtin_nontemporal_load(reinterpret_cast<const d2 *>(by + c));
            const d2 v1 = __builtin_nontemporal_load(reinterpret_cast<const d2 *>(by + c + 2));
            const d2 w0 = __builtin_nontemporal_load(reinterpret_cast<const d2 *>(bz + c));
            const d2 w1 = __builtin_nontemporal_load(reinterpret_cast<const d2 *>(bz + c + 2));
            b0[0] = u0[0]; b0[1] = u0[1]; b0[2] = u1[0]; b0[3] = u1[1];
            b1[0] = v0[0]; b1[1] = v0[1]; b1[2] = v1[0]; b1[3] = v1[1];
            b2[0] = w0[0]; b2[1] = w0[1]; b2[2] = w1[0]; b2[3] = w1[1];
        } else if (stage == 0) {
#pragma unroll
            for (int i = 0; i < 4; i++) { b0[i] = mxc[i]; b1[i] = myc[i]; b2[i] = mzc[i]; }
        } else {
            const f4 u = __builtin_nontemporal_load(reinterpret_cast<const f4 *>(mb_ + c));
            const f4 v = __builtin_nontemporal_load(reinterpret_cast<const f4 *>(mb_ + n + c));
            const f4 w = __builtin_nontemporal_load(reinterpret_cast<const f4 *>(mb_ + 2 * n + c));
#pragma unroll
            for (int i = 0; i < 4; i++) { b0[i] = u[i]; b1[i] = v[i]; b2[i] = w[i]; }
        }
        f4 s0, s1, s2;
        if (stage == 0) { s0 = kx; s1 = ky; s2 = kz; }
        else {
            const float wk = (stage == 3) ? 1.0f : 2.0f;  // w k is exact
            const f4 a0 = __builtin_nontemporal_load(reinterpret_cast<const f4 *>(ax + c));
            const f4 a1 = __builtin_nontemporal_load(reinterpret_cast<const f4 *>(ay + c));
            const f4 a2 = __builtin_nontemporal_load(reinterpret_cast<const f4 *>(az + c));
#pragma unroll
            for (int i = 0; i < 4; i++) { s0[i] = a0[i] + wk * kx[i]; s1[i] = a1[i] + wk * ky[i]; s2[i] = a2[i] + wk * kz[i]; }
        }
        double n0[4], n1[4], n2[4];
        if (stage < 3) {
            const double h = (stage < 2 ? 0.5 : 1.0) * A.dt;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                n0[i] = b0[i] + h * (double)kx[i]; n1[i] = b1[i] + h * (double)ky[i]; n2[i] = b2[i] + h * (double)kz[i];
            }
        } else {
            const double h = A.dt / 6.0;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                n0[i] = b0[i] + h * (double)s0[i]; n1[i] = b1[i] + h * (double)s1[i]; n2[i] = b2[i] + h * (double)s2[i];
                if (A.normalize) {
                    const double f = (double)A.Ms / sqrt(n0[i] * n0[i] + n1[i] * n1[i] + n2[i] * n2[i]);
                    n0[i] *= f; n1[i] *= f; n2[i] *= f;
                }
            }
        }
        f4 o0, o1, o2;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            o0[i] = mag[i] ? (float)n0[i] : mxc[i];
            o1[i] = mag[i] ? (float)n1[i] : myc[i];
            o2[i] = mag[i] ? (float)n2[i] : mzc[i];
        }
        *reinterpret_cast<f4 *>(ox + c) = o0;
        *reinterpret_cast<f4 *>(oy + c) = o1;
        *reinterpret_cast<f4 *>(oz + c) = o2;
        if (all) {
            if (stage < 3) {
                __builtin_nontemporal_store(s0, reinterpret_cast<f4 *>(ax + c));
                __builtin_nontemporal_store(s1, reinterpret_cast<f4 *>(ay + c));
                __builtin_nontemporal_store(s2, reinterpret_cast<f4 *>(az + c));
            } else {
                __builtin_nontemporal_store(d2{n0[0], n0[1]}, reinterpret_cast<d2 *>(bx + c));
                __builtin_nontemporal_store(d2{n0[2], n0[3]}, reinterpret_cast<d2 *>(bx + c + 2));
                __builtin_nontemporal_store(d2{n1[0], n1[1]}, reinterpret_cast<d2 *>(by + c));
                __builtin_nontemporal_store(d2{n1[2], n1[3]}, reinterpret_cast<d2 *>(by + c + 2));
                __builtin_nontemporal_store(d2{n2[0], n2[1]}, reinterpret_cast<d2 *>(bz + c));
                __builtin_nontemporal_store(d2{n2[2], n2[3]}, reinterpret_cast<d2 *>(bz + c + 2));
            }
        } else {
#pragma unroll
            for (int i = 0; i < 4; i++) {
                if (!mag[i]) continue;
                if (stage < 3) {
                    __builtin_nontemporal_store(s0[i], &ax[c + i]);
                    __builtin_nontemporal_store(s1[i], &ay[c + i]);
                    __builtin_nontemporal_store(s2[i], &az[c + i]);
                } else {
                    __builtin_nontemporal_store(n0[i], &bx[c + i]);
                    __builtin_nontemporal_store(n1[i], &by[c + i]);
                    __builtin_nontemporal_store(n2[i], &bz[c + i]);
                }
            }
        }
        pprev = pcur; pcur = pnext;
        mxp = mxc; myp = myc; mzp = mzc;
        mxc = mxn; myc = myn; mzc = mzn;
    }
}

template <int ZD, int ST>
__global__ __launch_bounds__(BX * BY4) void stage4_kernel(StageArgs A) {
    const int ntx = (A.xd + 4 * BX - 1) / (4 * BX), nty = (A.yd + BY4 - 1) / BY4;
    int tx, ty;
    if (!xcd_tile(ntx, ntx * nty, tx, ty)) return;
    const int x = 4 * (tx * BX + threadIdx.x), y = ty * BY4 + threadIdx.y;
    if (x >= A.xd || y >= A.yd) return;
    stage_column4<ZD, ST>(A, A.m_in, A.m_out, A.mn, A.acc, A.phi, A.matl, A.m_base, x, y, A.xd, A.yd, A.zd,
                          (long)A.xd * A.yd * A.zd);
}

// Batched films: block b handles tile tiles[b] = (film, tile x, tile y) of a
// film whose extent, cell offset and external field come from films[].
template <int ZD, int ST>
__global__ __launch_bounds__(BX * BY) void stage_batch_kernel(StageArgs A, const int4 *__restrict__ tiles,
                                                              const BatchFilm *__restrict__ films, long pstride) {
    const int4 t = tiles[blockIdx.x];
    if (t.x < 0) return;  // padding slot of the XCD-ordered tile list
    const BatchFilm f = films[t.x];
    int x, y;
    batch_column(t, f, x, y);
    stage_column<ZD, ST>(A, x, y, f.xd, f.yd, f.zd, f.off, pstride, f.hx, f.hy, f.hz);
}

__global__ __launch_bounds__(BX * BY) void hfield_kernel(StageArgs A) {
    const int x = blockIdx.x * BX + threadIdx.x, y = blockIdx.y * BY + threadIdx.y;
    const int xd = A.xd, yd = A.yd, zd = A.zd;
    if (x >= xd || y >= yd) return;
    const long plane = (long)xd * yd, n = plane * zd;
    const long c0 = (long)y * xd + x;
    const float *mx = A.m_in, *my = A.m_in + n, *mz = A.m_in + 2 * n;
    const double t = A.t_dev ? *A.t_dev : 0.0;
    float hex = A.hx, hey = A.hy, hez = A.hz;
    if (A.external == 1) { hex = ramp_hx(t, A.inv_mu0); hey = 0.0f; hez = 0.0f; }
    const float dx2 = A.dx * A.dx, dy2 = A.dy * A.dy, dz2 = A.dz * A.dz;
    for (int z = 0; z < zd; z++) {
        const long c = (long)z * plane + c0;
        float Hx = 0.0f, Hy = 0.0f, Hz = 0.0f;
        if (A.demag) {
            const float x2 = (x != xd - 1) ? A.phi[c + 1] : 0.0f;
            const float x0 = (x != 0) ? A.phi[c - 1] : 0.0f;
            const float y2 = (y != yd - 1) ? A.phi[c + xd] : 0.0f;
            const float y0 = (y != 0) ? A.phi[c - xd] : 0.0f;
            const float z2 = (z != zd - 1) ? A.phi[c + plane] : 0.0f;
            const float z0 = (z != 0) ? A.phi[c - plane] : 0.0f;
            Hx = (x2 - x0) / (2.0f * A.dx) * A.inv4pi;
            Hy = (y2 - y0) / (2.0f * A.dy) * A.inv4pi;
            Hz = (z2 - z0) / (2.0f * A.dz) * A.inv4pi;
        }
        if (A.exchange) {
            float L[3];
            const float *comp[3] = {mx, my, mz};
            for (int q = 0; q < 3; q++) {
                const float *v = comp[q];
                const float ctr = v[c];
                const float xa = (x != xd - 1) ? v[c + 1] : 0.0f, xb = (x != 0) ? v[c - 1] : 0.0f;
                const float ya = (y != yd - 1) ? v[c + xd] : 0.0f, yb = (y != 0) ? v[c - xd] : 0.0f;
                const float za = (z != zd - 1) ? v[c + plane] : 0.0f, zb = (z != 0) ? v[c - plane] : 0.0f;
                L[q] = (xa + xb - 2.0f * ctr) / dx2 + (ya + yb - 2.0f * ctr) / dy2 + (za + zb - 2.0f * ctr) / dz2;
            }
            Hx += A.exch_cm * L[0]; Hy += A.exch_cm * L[1]; Hz += A.exch_cm * L[2];
        }
        if (A.external) { Hx += hex; Hy += hey; Hz += hez; }
        A.h_out[3 * c] = Hx; A.h_out[3 * c + 1] = Hy; A.h_out[3 * c + 2] = Hz;
    }
}

// Block partials of E = sum M.H and T = max |M x H| over material cells.
// H is read from h_aos (computed beforehand by hfield_kernel).
__global__ __launch_bounds__(256) void energy_kernel(const float *__restrict__ m, const float *__restrict__ h_aos,
                                                     const uint8_t *__restrict__ matl, long n,
                                                     double *__restrict__ partial) {
    __shared__ double se[256];
    __shared__ double st[256];
    double e = 0.0, tq = 0.0;
    for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256L) {
        if (matl && !matl[i]) continue;
        const double Mx = m[i], My = m[n + i], Mz = m[2 * n + i];
        const double Hx = h_aos[3 * i], Hy = h_aos[3 * i + 1], Hz = h_aos[3 * i + 2];
        e += Mx * Hx + My * Hy + Mz * Hz;
        const double cx = My * Hz - Mz * Hy, cy = Mz * Hx - Mx * Hz, cz = Mx * Hy - My * Hx;
        tq = fmax(tq, sqrt(cx * cx + cy * cy + cz * cz));
    }
    se[threadIdx.x] = e;
    st[threadIdx.x] = tq;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            se[threadIdx.x] += se[threadIdx.x + s];
            st[threadIdx.x] = fmax(st[threadIdx.x], st[threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) { partial[2 * blockIdx.x] = se[0]; partial[2 * blockIdx.x + 1] = st[0]; }
}

// Block partials of (sum Mx, sum My, sum Mz, sum |M|, count) over material
// cells, for the time loop's log line (ode_functions.cpp:313-330).
__global__ __launch_bounds__(256) void mavg_kernel(const float *__restrict__ m, const uint8_t *__restrict__ matl,
                                                   long n, double *__restrict__ partial) {
    __shared__ double red[5][256];
    double a[5] = {0, 0, 0, 0, 0};
    for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256L) {
        if (matl && !matl[i]) continue;
        const double x = m[i], y = m[n + i], z = m[2 * n + i];
        a[0] += x; a[1] += y; a[2] += z; a[3] += sqrt(x * x + y * y + z * z); a[4] += 1.0;
    }
    for (int q = 0; q < 5; q++) red[q][threadIdx.x] = a[q];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int q = 0; q < 5; q++) red[q][threadIdx.x] += red[q][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int q = 0; q < 5; q++) partial[5 * blockIdx.x + q] = red[q][0];
}

__global__ void soa_from_aos_kernel(const float *__restrict__ aos, float *__restrict__ s32, double *__restrict__ s64, long n) {
    for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256L) {
        const float a = aos[3 * i], b = aos[3 * i + 1], c = aos[3 * i + 2];
        if (s32) { s32[i] = a; s32[n + i] = b; s32[2 * n + i] = c; }
        if (s64) { s64[i] = a; s64[n + i] = b; s64[2 * n + i] = c; }
    }
}
__global__ void aos_from_soa_kernel(const float *__restrict__ s32, float *__restrict__ aos, long n) {
    for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256L) {
        aos[3 * i] = s32[i]; aos[3 * i + 1] = s32[n + i]; aos[3 * i + 2] = s32[2 * n + i];
    }
}
__global__ void zero_nonmagnetic_kernel(float *__restrict__ m, const uint8_t *__restrict__ matl, long n) {
    for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256L)
        if (!matl[i]) { m[i] = 0.0f; m[n + i] = 0.0f; m[2 * n + i] = 0.0f; }
}
__global__ void add_time_kernel(double *t, double dt) { *t += dt; }
__global__ void fill_f32_kernel(float *p, float v, long n) {
    for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256L) p[i] = v;
}

inline int grid1d(long n) {
    long b = (n + 255) / 256;
    return (int)(b < 8192 ? (b < 1 ? 1 : b) : 8192);
}

StageArgs make_args(const Grid &g, const Material &mat) {
    StageArgs A{};
    A.xd = g.xd; A.yd = g.yd; A.zd = g.zd;
    A.dx = g.dx; A.dy = g.dy; A.dz = g.dz;
    A.inv4pi = (float)(1 / (4.0 * M_PI));
    A.exch_cm = 2 * mat.Aexch / (mat.mu_0 * mat.Ms * mat.Ms);
    A.demag = mat.demag; A.exchange = mat.exchange; A.external = mat.external;
    A.hx = mat.hext[0]; A.hy = mat.hext[1]; A.hz = mat.hext[2];
    A.inv_mu0 = 1 / mat.mu_0;
    A.alfa = mat.alfa; A.gamma = mat.gamma; A.Ms = mat.Ms;
    A.normalize = 1;
    return A;
}

// The plain resident step runs the stage kernel compiled for its stage; a
// stage with extra inputs or outputs runs the run-time form.
template <int ZD>
void launch_stage_zd(dim3 grid, dim3 block, hipStream_t s, const StageArgs &A, bool plain) {
    if (!plain) {
        hipLaunchKernelGGL((stage_kernel<ZD, -1>), grid, block, 0, s, A);
        return;
    }
    switch (A.stage) {
    case 0: hipLaunchKernelGGL((stage_kernel<ZD, 0>), grid, block, 0, s, A); break;
    case 1: hipLaunchKernelGGL((stage_kernel<ZD, 1>), grid, block, 0, s, A); break;
    case 2: hipLaunchKernelGGL((stage_kernel<ZD, 2>), grid, block, 0, s, A); break;
    default: hipLaunchKernelGGL((stage_kernel<ZD, 3>), grid, block, 0, s, A);
    }
}

template <int ZD>
void launch_stage4_zd(dim3 grid, dim3 block, hipStream_t s, const StageArgs &A) {
    switch (A.stage) {
    case 0: hipLaunchKernelGGL((stage4_kernel<ZD, 0>), grid, block, 0, s, A); break;
    case 1: hipLaunchKernelGGL((stage4_kernel<ZD, 1>), grid, block, 0, s, A); break;
    case 2: hipLaunchKernelGGL((stage4_kernel<ZD, 2>), grid, block, 0, s, A); break;
    default: hipLaunchKernelGGL((stage4_kernel<ZD, 3>), grid, block, 0, s, A);
    }
}

template <int ZD>
void launch_stage_batch_zd(dim3 grid, dim3 block, hipStream_t s, const StageArgs &A, const int4 *tiles,
                           const BatchFilm *films, long pstride) {
    switch (A.stage) {
    case 0: hipLaunchKernelGGL((stage_batch_kernel<ZD, 0>), grid, block, 0, s, A, tiles, films, pstride); break;
    case 1: hipLaunchKernelGGL((stage_batch_kernel<ZD, 1>), grid, block, 0, s, A, tiles, films, pstride); break;
    case 2: hipLaunchKernelGGL((stage_batch_kernel<ZD, 2>), grid, block, 0, s, A, tiles, films, pstride); break;
    default: hipLaunchKernelGGL((stage_batch_kernel<ZD, 3>), grid, block, 0, s, A, tiles, films, pstride);
    }
}

// workgroups of the XCD-ordered 1D grid: 8 * ceil(tiles / 8)
unsigned xcd_grid(const Grid &g) {
    const long ntiles = (long)((g.xd + BX - 1) / BX) * ((g.yd + BY - 1) / BY);
    return (unsigned)(8 * ((ntiles + 7) / 8));
}

}  // namespace

int launch_div(const Grid &g, const float *m, float *rho, hipStream_t s, Profiler *prof) {
    KTimer kt(prof, K_DIV);
    dim3 block(BX, BY);
    if (g.xd % 4 == 0 && ((reinterpret_cast<uintptr_t>(m) | reinterpret_cast<uintptr_t>(rho)) & 15) == 0) {
        const long ntiles = (long)((g.xd + 4 * BX - 1) / (4 * BX)) * ((g.yd + BY - 1) / BY);
        dim3 grid((unsigned)(8 * ((ntiles + 7) / 8)));
        switch (g.zd) {
        case 3: hipLaunchKernelGGL(div4_kernel<3>, grid, block, 0, s, m, rho, g.xd, g.yd, g.zd, g.dx, g.dy, g.dz); break;
        case 4: hipLaunchKernelGGL(div4_kernel<4>, grid, block, 0, s, m, rho, g.xd, g.yd, g.zd, g.dx, g.dy, g.dz); break;
        default: hipLaunchKernelGGL(div4_kernel<0>, grid, block, 0, s, m, rho, g.xd, g.yd, g.zd, g.dx, g.dy, g.dz);
        }
        MMS_HIP(hipGetLastError());
        return 0;
    }
    dim3 grid(xcd_grid(g));
    switch (g.zd) {
    case 3: hipLaunchKernelGGL(div_kernel<3>, grid, block, 0, s, m, rho, g.xd, g.yd, g.zd, g.dx, g.dy, g.dz); break;
    case 4: hipLaunchKernelGGL(div_kernel<4>, grid, block, 0, s, m, rho, g.xd, g.yd, g.zd, g.dx, g.dy, g.dz); break;
    default: hipLaunchKernelGGL(div_kernel<0>, grid, block, 0, s, m, rho, g.xd, g.yd, g.zd, g.dx, g.dy, g.dz);
    }
    MMS_HIP(hipGetLastError());
    return 0;
}

int launch_stage_ex(const Grid &g, const Material &mat, int stage, double dt, const float *phi,
                    const float *m_in, float *m_out, double *mn, float *acc, const uint8_t *matl,
                    const double *t_dev, const StageExtra &ex, hipStream_t s, Profiler *prof) {
    KTimer kt(prof, K_STAGE);
    StageArgs A = make_args(g, mat);
    A.phi = phi; A.m_in = m_in; A.m_out = m_out; A.mn = mn; A.acc = acc; A.matl = matl;
    A.t_dev = t_dev; A.stage = stage; A.dt = dt;
    A.h_given = ex.h_given; A.h_out = ex.h_out; A.k_out = ex.k_out; A.slope_out = ex.slope_out;
    A.normalize = ex.normalize; A.mn_out = ex.mn_out; A.m_base = ex.m_base; A.m_keep = ex.m_keep;
    MMS_CHECK(stage != 0 || !ex.m_base || ex.m_base == m_in, "stage 0 takes M_n from its own input");
    MMS_CHECK(stage >= 0 && stage <= 3, "RK4 stage must be 0..3");
    dim3 grid(xcd_grid(g)), block(BX, BY);
    const bool plain = !ex.h_given && !ex.h_out && !ex.k_out && !ex.slope_out && !ex.mn_out && !ex.m_keep;
    const uintptr_t al = reinterpret_cast<uintptr_t>(m_in) | reinterpret_cast<uintptr_t>(m_out) |
                         reinterpret_cast<uintptr_t>(mn) | reinterpret_cast<uintptr_t>(acc) |
                         reinterpret_cast<uintptr_t>(phi) | reinterpret_cast<uintptr_t>(ex.m_base);
    // MMS_GPU_STAGE_ONE_CELL=1 keeps the one-cell form (the four-cell form's
    // results are bitwise the same; the switch lets tests show it)
    const char *one = std::getenv("MMS_GPU_STAGE_ONE_CELL");
    if (plain && !(one && one[0] == '1') && g.xd % 4 == 0 && (al & 15) == 0 &&
        (reinterpret_cast<uintptr_t>(matl) & 3) == 0) {
        const long ntiles = (long)((g.xd + 4 * BX - 1) / (4 * BX)) * ((g.yd + BY4 - 1) / BY4);
        dim3 grid4((unsigned)(8 * ((ntiles + 7) / 8))), block4(BX, BY4);
        switch (g.zd) {
        case 3: launch_stage4_zd<3>(grid4, block4, s, A); break;
        case 4: launch_stage4_zd<4>(grid4, block4, s, A); break;
        default: launch_stage4_zd<0>(grid4, block4, s, A);
        }
        MMS_HIP(hipGetLastError());
        return 0;
    }
    switch (g.zd) {
    case 3: launch_stage_zd<3>(grid, block, s, A, plain); break;
    case 4: launch_stage_zd<4>(grid, block, s, A, plain); break;
    default: launch_stage_zd<0>(grid, block, s, A, plain);
    }
    MMS_HIP(hipGetLastError());
    return 0;
}

int launch_stage(const Grid &g, const Material &mat, int stage, double dt, const float *phi,
                 const float *m_in, float *m_out, double *mn, const float *m_base, float *acc, const uint8_t *matl,
                 const double *t_dev, hipStream_t s, Profiler *prof) {
    StageExtra ex{nullptr, nullptr, nullptr, nullptr, 1, nullptr, m_base};
    return launch_stage_ex(g, mat, stage, dt, phi, m_in, m_out, mn, acc, matl, t_dev, ex, s, prof);
}

int launch_hfield(const Grid &g, const Material &mat, const float *phi, const float *m,
                  const double *t_dev, float *H_aos, hipStream_t s, Profiler *prof) {
    KTimer kt(prof, K_HFIELD);
    StageArgs A = make_args(g, mat);
    A.phi = phi; A.m_in = m; A.t_dev = t_dev; A.h_out = H_aos;
    dim3 grid((g.xd + BX - 1) / BX, (g.yd + BY - 1) / BY), block(BX, BY);
    hipLaunchKernelGGL(hfield_kernel, grid, block, 0, s, A);
    MMS_HIP(hipGetLastError());
    return 0;
}

int launch_div_batch(const float *m, float *rho, const int4 *tiles, int ntiles, const BatchFilm *films, long pstride,
                     float dx, float dy, float dz, int zd, hipStream_t s, Profiler *prof) {
    KTimer kt(prof, K_DIV);
    if (ntiles <= 0) return 0;
    if (zd == 3)
        hipLaunchKernelGGL(div_batch_kernel<3>, dim3(ntiles), dim3(BX, BY), 0, s, m, rho, tiles, films, pstride, dx, dy, dz);
    else
        hipLaunchKernelGGL(div_batch_kernel<0>, dim3(ntiles), dim3(BX, BY), 0, s, m, rho, tiles, films, pstride, dx, dy, dz);
    MMS_HIP(hipGetLastError());
    return 0;
}

int launch_stage_batch(const Grid &g, const Material &mat, int stage, double dt, const float *phi, const float *m_in,
                       float *m_out, double *mn, const float *m_base, float *acc, const uint8_t *matl,
                       const int4 *tiles, int ntiles, const BatchFilm *films, long pstride, hipStream_t s,
                       Profiler *prof) {
    KTimer kt(prof, K_STAGE);
    if (ntiles <= 0) return 0;
    MMS_CHECK(stage != 0 || !m_base || m_base == m_in, "stage 0 takes M_n from its own input");
    StageArgs A = make_args(g, mat);
    A.phi = phi; A.m_in = m_in; A.m_out = m_out; A.mn = mn; A.m_base = m_base; A.acc = acc; A.matl = matl;
    A.t_dev = nullptr; A.stage = stage; A.dt = dt; A.external = 2;
    MMS_CHECK(stage >= 0 && stage <= 3, "RK4 stage must be 0..3");
    if (g.zd == 3)
        launch_stage_batch_zd<3>(dim3(ntiles), dim3(BX, BY), s, A, tiles, films, pstride);
    else
        launch_stage_batch_zd<0>(dim3(ntiles), dim3(BX, BY), s, A, tiles, films, pstride);
    MMS_HIP(hipGetLastError());
    return 0;
}

int launch_energy_h(const float *m, const float *h_aos, const uint8_t *matl, long n,
                    double *partial, int nblk, hipStream_t s, Profiler *prof) {
    KTimer kt(prof, K_ENERGY);
    hipLaunchKernelGGL(energy_kernel, dim3(nblk), dim3(256), 0, s, m, h_aos, matl, n, partial);
    MMS_HIP(hipGetLastError());
    return 0;
}

int launch_mavg(const float *m, const uint8_t *matl, long n, double *partial, int nblk, hipStream_t s) {
    hipLaunchKernelGGL(mavg_kernel, dim3(nblk), dim3(256), 0, s, m, matl, n, partial);
    MMS_HIP(hipGetLastError());
    return 0;
}

int launch_soa_from_aos(const float *aos, float *soa32, double *soa64, long n, hipStream_t s) {
    hipLaunchKernelGGL(soa_from_aos_kernel, dim3(grid1d(n)), dim3(256), 0, s, aos, soa32, soa64, n);
    MMS_HIP(hipGetLastError());
    return 0;
}
int launch_aos_from_soa(const float *soa32, float *aos, long n, hipStream_t s) {
    hipLaunchKernelGGL(aos_from_soa_kernel, dim3(grid1d(n)), dim3(256), 0, s, soa32, aos, n);
    MMS_HIP(hipGetLastError());
    return 0;
}
int launch_zero_nonmagnetic(float *m, const uint8_t *matl, long n, hipStream_t s) {
    hipLaunchKernelGGL(zero_nonmagnetic_kernel, dim3(grid1d(n)), dim3(256), 0, s, m, matl, n);
    MMS_HIP(hipGetLastError());
    return 0;
}
int launch_add_time(double *t_dev, double dt, hipStream_t s) {
    hipLaunchKernelGGL(add_time_kernel, dim3(1), dim3(1), 0, s, t_dev, dt);
    MMS_HIP(hipGetLastError());
    return 0;
}
int launch_fill_f32(float *p, float v, long n, hipStream_t s) {
    hipLaunchKernelGGL(fill_f32_kernel, dim3(grid1d(n)), dim3(256), 0, s, p, v, n);
    MMS_HIP(hipGetLastError());
    return 0;
}

}  // namespace mms
