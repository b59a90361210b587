// fmm.hip -- regular-grid fast multipole potential on gfx950.
//
// Replaces the reference tree-code (fmm_calc.cu:20-408, fmm_kernel.cu:21-155),
// which recomputes multipoles from raw charges at every level and evaluates
// them directly at every target of the interaction list.  Here the same
// quadtree (per charge layer, interaction list = children of the parent's
// neighbours minus own neighbours, Box.cu:121-145) carries a full FMM:
//
//   P2M   leaf multipoles from charges              M = P2M[K x w^2] rho_box
//   M2M   children -> parent                       M_p = sum_c M2M_c M_c
//   M2L   interaction list -> local                L_t += T(level, offset, dz) M_s
//   L2L   parent local -> child                    L_c += L2L_c L_p   (fused into M2L)
//   L2P   local -> cells of the leaf               phi = L2P[w^2 x K] L
//   P2P   3x3 neighbour leaf boxes, all layers     phi += sum rho/r    (fused with L2P)
//
// On a regular grid every operator is translation invariant, so each is a
// small constant matrix built once in FP64 on the host (fmm_operators.cpp)
// and the device work is gathers plus dense contractions.  Launches:
//   p2m_leaf  (1) | m2m (Lleaf-2) | m2l+l2l (Lleaf-1) | p2p_l2p (1)
// M2L threads are grouped so that the child parity (px, py) and the target
// layer are uniform across a wavefront; the operator blocks are then read
// through the scalar cache and broadcast to all lanes.
#include "mms_internal.hpp"

namespace mms {
namespace {

template <int P> struct NC { static constexpr int K = (P + 1) * (P + 2) / 2; };

// ------------------------------------------------------------------ P2M --
template <int K, int W>
__global__ __launch_bounds__(256) void p2m_leaf_kernel(const float *__restrict__ rho, float *__restrict__ M,
                                                       const float *__restrict__ p2m, int xd, int yd, int nb) {
    const int bx = blockIdx.x * 256 + threadIdx.x, by = blockIdx.y, z = blockIdx.z;
    if (bx >= nb) return;
    const long plane = (long)xd * yd;
    float q[W * W];
#pragma unroll
    for (int iy = 0; iy < W; iy++) {
        const int y = by * W + iy;
#pragma unroll
        for (int ix = 0; ix < W; ix++) {
            const int x = bx * W + ix;
            q[iy * W + ix] = (x < xd && y < yd) ? rho[(long)z * plane + (long)y * xd + x] : 0.0f;
        }
    }
    float *out = M + (((long)z * nb + by) * nb + bx) * K;
#pragma unroll 4
    for (int k = 0; k < K; k++) {
        float a = 0.0f;
#pragma unroll
        for (int c = 0; c < W * W; c++) a = fmaf(p2m[k * W * W + c], q[c], a);
        out[k] = a;
    }
}

// ------------------------------------------------------------------ M2M --
template <int K>
__global__ __launch_bounds__(256) void m2m_kernel(const float *__restrict__ Mc, float *__restrict__ Mp,
                                                  const float *__restrict__ m2m, int nbp) {
    const int bx = blockIdx.x * 256 + threadIdx.x, by = blockIdx.y, z = blockIdx.z;
    if (bx >= nbp) return;
    const int nbc = 2 * nbp;
    float acc[K];
#pragma unroll
    for (int a = 0; a < K; a++) acc[a] = 0.0f;
#pragma unroll 1
    for (int ci = 0; ci < 4; ci++) {
        const int cx = 2 * bx + (ci & 1), cy = 2 * by + (ci >> 1);
        const float *src = Mc + (((long)z * nbc + cy) * nbc + cx) * K;
        float v[K];
#pragma unroll
        for (int b = 0; b < K; b++) v[b] = src[b];
        const float *T = m2m + (long)ci * K * K;
#pragma unroll
        for (int a = 0; a < K; a++)
#pragma unroll
            for (int b = 0; b < K; b++) acc[a] = fmaf(T[a * K + b], v[b], acc[a]);
    }
    float *out = Mp + (((long)z * nbp + by) * nbp + bx) * K;
#pragma unroll
    for (int a = 0; a < K; a++) out[a] = acc[a];
}

// ----------------------------------------------------------- M2L + L2L --
// grid: x,y over the parity sub-lattice (nb/2 each), z = zt*4 + parity
template <int K>
__global__ __launch_bounds__(256) void m2l_kernel(const float *__restrict__ M, float *__restrict__ L,
                                                  const float *__restrict__ Lpar, const float *__restrict__ T,
                                                  const float *__restrict__ l2l, int nb, int zd) {
    const int half = nb >> 1;
    const int i = blockIdx.x * 16 + (threadIdx.x & 15), j = blockIdx.y * 16 + (threadIdx.x >> 4);
    const int zt = blockIdx.z >> 2, par = blockIdx.z & 3;
    const int px = par & 1, py = par >> 1;
    if (i >= half || j >= half) return;
    const int bx = 2 * i + px, by = 2 * j + py;
    float acc[K];
#pragma unroll
    for (int a = 0; a < K; a++) acc[a] = 0.0f;
    if (Lpar) {  // L2L from the parent at (i, j) of the coarser level
        const float *src = Lpar + (((long)zt * half + j) * half + i) * K;
        float v[K];
#pragma unroll
        for (int b = 0; b < K; b++) v[b] = src[b];
        const float *O = l2l + (long)par * K * K;
#pragma unroll
        for (int a = 0; a < K; a++)
#pragma unroll
            for (int b = 0; b < K; b++) acc[a] = fmaf(O[a * K + b], v[b], acc[a]);
    }
    const long KK = (long)K * K;
#pragma unroll 1
    for (int zs = 0; zs < zd; zs++) {
        const int dzi = zt - zs + zd - 1;
#pragma unroll 1
        for (int oy = -2 - py; oy <= 3 - py; oy++) {
            const int sy = by + oy;
#pragma unroll 1
            for (int ox = -2 - px; ox <= 3 - px; ox++) {
                if (ox >= -1 && ox <= 1 && oy >= -1 && oy <= 1) continue;
                const int sx = bx + ox;
                const bool ok = (sx >= 0 && sx < nb && sy >= 0 && sy < nb);
                const float *src = M + (((long)zs * nb + (ok ? sy : 0)) * nb + (ok ? sx : 0)) * K;
                float v[K];
#pragma unroll
                for (int b = 0; b < K; b++) v[b] = ok ? src[b] : 0.0f;
                const float *O = T + ((long)dzi * 49 + (oy + 3) * 7 + (ox + 3)) * KK;
#pragma unroll
                for (int a = 0; a < K; a++)
#pragma unroll
                    for (int b = 0; b < K; b++) acc[a] = fmaf(O[a * K + b], v[b], acc[a]);
            }
        }
    }
    float *out = L + (((long)zt * nb + by) * nb + bx) * K;
#pragma unroll
    for (int a = 0; a < K; a++) out[a] = acc[a];
}

// -------------------------------------------------------- P2P + L2P --
// Block = TW x TW target cells of one layer (TW = 16); the charges of the
// tile plus a W-cell halo in every layer are staged in LDS together with the
// 1/r table over all (dx, dy, dz) offsets of the near field.
template <int K, int W>
__global__ __launch_bounds__(256) void p2p_l2p_kernel(const float *__restrict__ rho, float *__restrict__ phi,
                                                      const float *__restrict__ Lleaf, const float *__restrict__ l2p,
                                                      const float *__restrict__ table, int xd, int yd, int zd,
                                                      int nb, float scale) {
    constexpr int TW = 16, SW = TW + 2 * W, TS = 4 * W - 1;
    extern __shared__ float lds[];
    float *src = lds;                      // [zd][SW][SW]
    float *tab = lds + zd * SW * SW;       // [2zd-1][TS][TS]
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const int x0 = blockIdx.x * TW, y0 = blockIdx.y * TW, zt = blockIdx.z;
    const long plane = (long)xd * yd;
    for (int e = threadIdx.x; e < zd * SW * SW; e += 256) {
        const int zs = e / (SW * SW), r = e % (SW * SW), sy = r / SW, sx = r % SW;
        const int gx = x0 - W + sx, gy = y0 - W + sy;
        src[e] = (gx >= 0 && gx < xd && gy >= 0 && gy < yd) ? rho[(long)zs * plane + (long)gy * xd + gx] : 0.0f;
    }
    for (int e = threadIdx.x; e < (2 * zd - 1) * TS * TS; e += 256) tab[e] = table[e];
    __syncthreads();
    const int x = x0 + tx, y = y0 + ty;
    const int lx0 = (tx / W) * W, ly0 = (ty / W) * W;  // leaf origin within the tile
    float near = 0.0f;
#pragma unroll 1
    for (int zs = 0; zs < zd; zs++) {
        const float *s = src + zs * SW * SW;
        const float *tb = tab + (zs - zt + zd - 1) * TS * TS;
#pragma unroll 1
        for (int syi = 0; syi < 3 * W; syi++) {
            const int sy = ly0 + syi;                 // LDS row of the source
            const int dyi = sy - (ty + W) + 2 * W - 1;  // table row
            const float *srow = s + sy * SW + lx0;
            const float *trow = tb + dyi * TS + (lx0 - (tx + W) + 2 * W - 1);
#pragma unroll
            for (int sxi = 0; sxi < 3 * W; sxi++) near = fmaf(srow[sxi], trow[sxi], near);
        }
    }
    if (x >= xd || y >= yd) return;
    const int bx = x / W, by = y / W, c = (y % W) * W + (x % W);
    const float *Lv = Lleaf + (((long)zt * nb + by) * nb + bx) * K;
    const float *row = l2p + (long)c * K;
    float far = 0.0f;
#pragma unroll
    for (int k = 0; k < K; k++) far = fmaf(row[k], Lv[k], far);
    phi[(long)zt * plane + (long)y * xd + x] = (near + far) * scale;
}

// -------------------------------------------------------------- dispatch --
template <int P, int W>
int run_fmm(const Grid &g, FmmPlan &pl, const float *rho, float *phi, hipStream_t s, Profiler *prof) {
    constexpr int K = NC<P>::K;
    const int zd = g.zd;
    {
        KTimer kt(prof, K_P2M);
        const int nb = 1 << pl.Lleaf;
        dim3 grid((nb + 255) / 256, nb, zd);
        hipLaunchKernelGGL((p2m_leaf_kernel<K, W>), grid, dim3(256), 0, s, rho, pl.mult + pl.lvl_off[pl.Lleaf],
                           pl.d_p2m, g.xd, g.yd, nb);
        MMS_HIP(hipGetLastError());
    }
    for (int L = pl.Lleaf - 1; L >= 2; L--) {
        KTimer kt(prof, K_M2M);
        const int nbp = 1 << L;
        dim3 grid((nbp + 255) / 256, nbp, zd);
        hipLaunchKernelGGL((m2m_kernel<K>), grid, dim3(256), 0, s, pl.mult + pl.lvl_off[L + 1],
                           pl.mult + pl.lvl_off[L], pl.d_m2m, nbp);
        MMS_HIP(hipGetLastError());
    }
    for (int L = 2; L <= pl.Lleaf; L++) {
        KTimer kt(prof, K_M2L);
        const int nb = 1 << L, half = nb >> 1;
        dim3 grid((half + 15) / 16, (half + 15) / 16, zd * 4);
        const float *Lpar = (L > 2) ? pl.loc + pl.lvl_off[L - 1] : nullptr;
        hipLaunchKernelGGL((m2l_kernel<K>), grid, dim3(256), 0, s, pl.mult + pl.lvl_off[L], pl.loc + pl.lvl_off[L],
                           Lpar, pl.d_m2l + (long)(L - 2) * pl.m2l_level_stride, pl.d_l2l, nb, zd);
        MMS_HIP(hipGetLastError());
    }
    {
        KTimer kt(prof, K_P2P_L2P);
        constexpr int TW = 16, SW = TW + 2 * W, TS = 4 * W - 1;
        const size_t lds = sizeof(float) * ((size_t)zd * SW * SW + (size_t)(2 * zd - 1) * TS * TS);
        dim3 grid((g.xd + TW - 1) / TW, (g.yd + TW - 1) / TW, zd);
        const float scale = (float)((double)(g.dx * g.dy * g.dz) / (double)g.dx);
        hipLaunchKernelGGL((p2p_l2p_kernel<K, W>), grid, dim3(256), lds, s, rho, phi, pl.loc + pl.lvl_off[pl.Lleaf],
                           pl.d_l2p, pl.d_near, g.xd, g.yd, zd, 1 << pl.Lleaf, scale);
        MMS_HIP(hipGetLastError());
    }
    return 0;
}

template <int W>
int run_fmm_w(const Grid &g, FmmPlan &pl, const float *rho, float *phi, hipStream_t s, Profiler *prof) {
    switch (pl.P) {
    case 1: return run_fmm<1, W>(g, pl, rho, phi, s, prof);
    case 2: return run_fmm<2, W>(g, pl, rho, phi, s, prof);
    case 3: return run_fmm<3, W>(g, pl, rho, phi, s, prof);
    case 4: return run_fmm<4, W>(g, pl, rho, phi, s, prof);
    case 5: return run_fmm<5, W>(g, pl, rho, phi, s, prof);
    case 6: return run_fmm<6, W>(g, pl, rho, phi, s, prof);
    case 7: return run_fmm<7, W>(g, pl, rho, phi, s, prof);
    case 8: return run_fmm<8, W>(g, pl, rho, phi, s, prof);
    case 9: return run_fmm<9, W>(g, pl, rho, phi, s, prof);
    case 10: return run_fmm<10, W>(g, pl, rho, phi, s, prof);
    default: set_error("FMM expansion order must be 1..10"); return 1;
    }
}

}  // namespace

int fmm_potential(const Grid &g, FmmPlan &pl, const float *rho, float *phi, hipStream_t s, Profiler *prof) {
    MMS_CHECK(pl.mult != nullptr, "FMM plan not initialised");
    switch (pl.w) {
    case 2: return run_fmm_w<2>(g, pl, rho, phi, s, prof);
    case 4: return run_fmm_w<4>(g, pl, rho, phi, s, prof);
    case 8: return run_fmm_w<8>(g, pl, rho, phi, s, prof);
    default: set_error("FMM leaf width must be 2, 4 or 8"); return 1;
    }
}

int FmmPlan::init(const Grid &g, int P_, int w_, int precision_) {
    release();
    MMS_CHECK(P_ >= 1 && P_ <= 10, "FMM expansion order must be 1..10");
    MMS_CHECK(w_ == 2 || w_ == 4 || w_ == 8, "FMM leaf width must be 2, 4 or 8");
    MMS_CHECK(g.zd >= 1 && g.zd <= 16, "FMM supports 1..16 layers");
    P = P_; w = w_; precision = precision_; zd = g.zd;
    K = fmm_ncoef(P);
    const int side = g.xd > g.yd ? g.xd : g.yd;
    Lmax = 0;
    while ((1 << Lmax) < side) Lmax++;
    int lw = 0;
    while ((1 << lw) < w) lw++;
    Lleaf = Lmax - lw;
    MMS_CHECK(Lleaf >= 2, "grid too small for the FMM at this leaf width (needs >= 4 leaf boxes per side)");
    // level storage
    lvl_off.assign(Lleaf + 1, 0);
    long tot = 0;
    for (int L = 2; L <= Lleaf; L++) {
        lvl_off[L] = tot;
        tot += (long)zd * (1L << (2 * L)) * K;
    }
    MMS_HIP(hipMalloc(&mult, sizeof(float) * tot));
    MMS_HIP(hipMalloc(&loc, sizeof(float) * tot));
    // operators
    FmmHostOps ops;
    fmm_build_ops(P, w, (double)g.ay, ops);
    auto upload = [&](float **dst, const std::vector<double> &src) -> int {
        std::vector<float> f(src.begin(), src.end());
        MMS_HIP(hipMalloc(dst, sizeof(float) * f.size()));
        MMS_HIP(hipMemcpy(*dst, f.data(), sizeof(float) * f.size(), hipMemcpyHostToDevice));
        return 0;
    };
    if (upload(&d_p2m, ops.p2m) || upload(&d_l2p, ops.l2p) || upload(&d_m2m, ops.m2m) || upload(&d_l2l, ops.l2l))
        return 1;
    m2l_level_stride = (long)(2 * zd - 1) * 49 * K * K;
    std::vector<double> all;
    all.reserve((size_t)m2l_level_stride * (Lleaf - 1));
    for (int L = 2; L <= Lleaf; L++) {
        std::vector<double> T = ops.m2l_level(L, Lmax, zd, (double)g.ay, (double)g.az);
        all.insert(all.end(), T.begin(), T.end());
    }
    if (upload(&d_m2l, all)) return 1;
    // near-field 1/r table over offsets dx, dy in [-(2w-1), 2w-1], dz in [-(zd-1), zd-1]
    near_side = 4 * w - 1;
    std::vector<double> tab((size_t)(2 * zd - 1) * near_side * near_side);
    for (int dz = -(zd - 1); dz <= zd - 1; dz++)
        for (int dy = -(2 * w - 1); dy <= 2 * w - 1; dy++)
            for (int dx = -(2 * w - 1); dx <= 2 * w - 1; dx++) {
                const double ry = dy * (double)g.ay, rz = dz * (double)g.az;
                const double r2 = (double)dx * dx + ry * ry + rz * rz;
                tab[((size_t)(dz + zd - 1) * near_side + (dy + 2 * w - 1)) * near_side + (dx + 2 * w - 1)] =
                    r2 > 0 ? 1.0 / std::sqrt(r2) : 0.0;
            }
    if (upload(&d_near, tab)) return 1;
    return 0;
}

void FmmPlan::release() {
    float **bufs[] = {&mult, &loc, &d_p2m, &d_l2p, &d_m2m, &d_l2l, &d_m2l, &d_near};
    for (float **b : bufs) {
        if (*b) (void)hipFree(*b);
        *b = nullptr;
    }
    if (d_m2l_bf) (void)hipFree(d_m2l_bf);
    d_m2l_bf = nullptr;
}

}  // namespace mms
