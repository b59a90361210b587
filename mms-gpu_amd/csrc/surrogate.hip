// surrogate.hip -- inference of the learned far-field surrogate on gfx950.
//
// The network maps the charge density of every film layer to its scalar
// potential, layer by layer in the plane (the layer index is an input):
//   lift      3x3 conv, 3 -> 64 channels (charge / charge scale, material
//             mask, layer index / zd) + bias, GELU
//   conv1     3x3 conv 64 -> 64 + bias, GELU
//   conv2     3x3 conv 64 -> 64 + bias + spectral branch, GELU, where the
//             spectral branch of conv1's output h is
//               S = By^T Bx^T [ W_mode . (By Bx h) ]
//             over 32 x 32 retained real Fourier modes (Bx[k][x] = c (1,
//             cos(2 pi x/N), sin(2 pi x/N), ..., cos(2 pi 16 x/N)), c =
//             sqrt(2/N); By alike) with a 64 x 64 channel mixing matrix per
//             mode
//   conv3     3x3 conv 64 -> 64 + bias, GELU, with the 1x1 projection
//             64 -> 1 (+ bias) fused into its epilogue: the potential
// Activations are NHWC BF16 [z][y][x][64]; weights BF16; every product runs
// on the matrix cores (v_mfma_f32_16x16x32_bf16) with FP32 accumulation.
// Cells outside the film read as zero (zero padding).
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <vector>

#include "mms_internal.hpp"

namespace mms {
namespace {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ f32x4 mfma_bf16(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ unsigned short f2bf(float x) {  // round to nearest even
    unsigned u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float((unsigned)h << 16); }
// GELU, tanh form: 0.5 x (1 + tanh(u)) = x / (1 + exp(-2u)), u = sqrt(2/pi)
// (x + 0.044715 x^3), with the hardware exp2 and reciprocal (within a few
// ulp of the library tanh form; large |x| gives x or -0)
__device__ __forceinline__ float gelu(float x) {
    const float u2 = -1.5957691216f * (x + 0.044715f * x * x * x);
    return x * __builtin_amdgcn_rcpf(1.0f + __expf(u2));
}

constexpr int C = 64;     // hidden channels
constexpr int CP = 72;    // LDS pixel / weight-row stride in bf16 (16-byte aligned, spreads banks)
constexpr int TS = 16;    // output tile side (pixels)
constexpr int HS = TS + 2;

// --------------------------------------------------------------- lift --
// 3x3 conv of the three input planes (charge * rscale, mask, layer index)
// to 64 channels: one MFMA k step of 32 (27 taps x planes, 5 zero) per
// 16 x 16 output tile; wave w takes tile rows 4w .. 4w + 3 (64 pixels).
// wl: [64 co][32 k] bf16 (k = tap * 3 + plane), bias [64].
__global__ __launch_bounds__(256) void sur_lift_kernel(const float *__restrict__ rho, const uint8_t *__restrict__ matl,
                                                       const unsigned short *__restrict__ wl,
                                                       const float *__restrict__ bias, unsigned short *__restrict__ out,
                                                       int xd, int yd, int zd, float rscale) {
    __shared__ float in[3][HS][HS];
    const int z = blockIdx.z, x0 = blockIdx.x * TS, y0 = blockIdx.y * TS;
    const long plane = (long)xd * yd;
    for (int e = threadIdx.x; e < HS * HS; e += 256) {
        const int sy = e / HS, sx = e % HS, gx = x0 - 1 + sx, gy = y0 - 1 + sy;
        const bool ok = gx >= 0 && gx < xd && gy >= 0 && gy < yd;
        const long c = (long)z * plane + (long)gy * xd + gx;
        in[0][sy][sx] = ok ? rho[c] * rscale : 0.0f;
        in[1][sy][sx] = ok ? (matl ? (matl[c] ? 1.0f : 0.0f) : 1.0f) : 0.0f;
        in[2][sy][sx] = ok ? (float)z / (float)zd : 0.0f;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, j = lane & 15, g = lane >> 4;
    bf16x8 a[4];
#pragma unroll
    for (int r = 0; r < 4; r++) a[r] = *reinterpret_cast<const bf16x8 *>(wl + (16 * r + j) * 32 + 8 * g);
#pragma unroll
    for (int ct = 0; ct < 4; ct++) {
        const int py = 4 * wave + ct, px = j;  // pixel of this lane's column
        u16x8 bv;
#pragma unroll
        for (int t = 0; t < 8; t++) {
            const int k = 8 * g + t, tap = k / 3, pl = k % 3;
            bv[t] = (k < 27) ? f2bf(in[pl][py + tap / 3][px + tap % 3]) : (unsigned short)0;
        }
        const bf16x8 b = __builtin_bit_cast(bf16x8, bv);
        const int gx = x0 + px, gy = y0 + py;
        const bool ok = gx < xd && gy < yd;
        unsigned short *o = out + (((long)z * yd + gy) * xd + gx) * C;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            f32x4 acc = mfma_bf16(a[r], b, f32x4{0.f, 0.f, 0.f, 0.f});
            // D row 4 g + i of row tile r = output channel 16 r + 4 g + i of pixel j
            unsigned short v[4];
#pragma unroll
            for (int i = 0; i < 4; i++) v[i] = f2bf(gelu(acc[i] + bias[16 * r + 4 * g + i]));
            if (ok) *reinterpret_cast<uint2 *>(o + 16 * r + 4 * g) = make_uint2(v[0] | (unsigned)v[1] << 16,
                                                                              v[2] | (unsigned)v[3] << 16);
        }
    }
}

// ------------------------------------------------------------- conv 3x3 --
// 64 -> 64 channels on 16 x 16 pixel tiles.  A workgroup loads the nine
// 64 x 64 weight taps into LDS once (row stride 72 bf16) and then walks
// tiles blockIdx.x, blockIdx.x + gridDim.x, ... of layer blockIdx.z: the
// next tile's 18 x 18 x 64 input window is loaded into registers while the
// current one is multiplied, then stored to LDS.  Wave w computes D[64 co]
// [64 pixels of tile rows 4w..4w+3] over 18 k steps (tap, 32 input
// channels).  EPI 0: bias (+ add) + GELU, BF16 NHWC out; EPI 1: bias + GELU,
// then the 1x1 projection to one channel (weights wp, bias bp), times
// oscale, FP32 potential out.  w: [9 taps][64 co][64 ci] bf16.
constexpr int WCH = HS * HS * 8;                 // 16-byte chunks of a window
constexpr int WPT = (WCH + 255) / 256;           // per thread (11)

template <int EPI>
__global__ __launch_bounds__(256) void sur_conv_kernel(const unsigned short *__restrict__ in,
                                                       const unsigned short *__restrict__ w,
                                                       const float *__restrict__ bias,
                                                       const unsigned short *__restrict__ add,
                                                       unsigned short *__restrict__ out, const float *__restrict__ wp,
                                                       float bp, float *__restrict__ phi, float oscale, int xd,
                                                       int yd) {
    extern __shared__ __attribute__((aligned(16))) unsigned short lds[];
    unsigned short *win = lds;                 // [HS * HS][CP]
    unsigned short *wt = lds + HS * HS * CP;   // [9 * 64][CP]
    const int z = blockIdx.z, ntx = (xd + TS - 1) / TS, nty = (yd + TS - 1) / TS, ntile = ntx * nty;
    // the window chunks of tile t into registers (zero outside the film)
    auto load_win = [&](int t, uint4 (&v)[WPT]) {
        const int x0 = (t % ntx) * TS, y0 = (t / ntx) * TS;
#pragma unroll
        for (int q = 0; q < WPT; q++) {
            const int e = threadIdx.x + 256 * q, p = e >> 3, ch = e & 7, sy = p / HS, sx = p % HS;
            const int gx = x0 - 1 + sx, gy = y0 - 1 + sy;
            v[q] = make_uint4(0u, 0u, 0u, 0u);
            if (e < WCH && gx >= 0 && gx < xd && gy >= 0 && gy < yd)
                v[q] = *reinterpret_cast<const uint4 *>(in + (((long)z * yd + gy) * xd + gx) * C + 8 * ch);
        }
    };
    {   // weights: 18 chunks per thread, all loads in flight before the stores
        constexpr int NW = 9 * C * 8 / 256;
        uint4 v[NW];
#pragma unroll
        for (int q = 0; q < NW; q++) {
            const int e = threadIdx.x + 256 * q, row = e >> 3, ch = e & 7;
            v[q] = *reinterpret_cast<const uint4 *>(w + row * C + 8 * ch);
        }
#pragma unroll
        for (int q = 0; q < NW; q++) {
            const int e = threadIdx.x + 256 * q, row = e >> 3, ch = e & 7;
            *reinterpret_cast<uint4 *>(wt + row * CP + 8 * ch) = v[q];
        }
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, j = lane & 15, g = lane >> 4;
    uint4 nxt[WPT];
    int t = blockIdx.x;
    if (t < ntile) load_win(t, nxt);
    for (; t < ntile; t += gridDim.x) {
        __syncthreads();  // the previous tile's products are done with win
#pragma unroll
        for (int q = 0; q < WPT; q++) {
            const int e = threadIdx.x + 256 * q;
            if (e < WCH) *reinterpret_cast<uint4 *>(win + (e >> 3) * CP + 8 * (e & 7)) = nxt[q];
        }
        __syncthreads();
        if (t + (int)gridDim.x < ntile) load_win(t + gridDim.x, nxt);  // in flight during the products
        const int x0 = (t % ntx) * TS, y0 = (t / ntx) * TS;
        f32x4 acc[4][4];
#pragma unroll
        for (int r = 0; r < 4; r++)
#pragma unroll
            for (int c = 0; c < 4; c++) acc[r][c] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll 1
        for (int tap = 0; tap < 9; tap++) {
            const int dy = tap / 3, dx = tap % 3;
#pragma unroll
            for (int kb = 0; kb < 2; kb++) {
                bf16x8 a[4], b[4];
#pragma unroll
                for (int r = 0; r < 4; r++)
                    a[r] = *reinterpret_cast<const bf16x8 *>(wt + (tap * C + 16 * r + j) * CP + 32 * kb + 8 * g);
#pragma unroll
                for (int c = 0; c < 4; c++) {
                    const int py = 4 * wave + c, px = j;
                    b[c] = *reinterpret_cast<const bf16x8 *>(win + ((py + dy) * HS + px + dx) * CP + 32 * kb + 8 * g);
                }
#pragma unroll
                for (int r = 0; r < 4; r++)
#pragma unroll
                    for (int c = 0; c < 4; c++) acc[r][c] = mfma_bf16(a[r], b[c], acc[r][c]);
            }
        }
        // epilogue: lane (j, g) holds channels 16 r + 4 g + i of pixel (tile row 4 wave + c, column j)
#pragma unroll
        for (int c = 0; c < 4; c++) {
            const int gx = x0 + j, gy = y0 + 4 * wave + c;
            const bool ok = gx < xd && gy < yd;
            const long pix = ((long)z * yd + gy) * xd + gx;
            float proj = 0.0f;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                float v[4];
                uint2 ad = make_uint2(0u, 0u);
                if (EPI == 0 && add && ok) ad = *reinterpret_cast<const uint2 *>(add + pix * C + 16 * r + 4 * g);
                const unsigned short av[4] = {(unsigned short)(ad.x & 0xFFFF), (unsigned short)(ad.x >> 16),
                                              (unsigned short)(ad.y & 0xFFFF), (unsigned short)(ad.y >> 16)};
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const int co = 16 * r + 4 * g + i;
                    float tv = acc[r][c][i] + bias[co];
                    if (EPI == 0) tv += bf2f(av[i]);
                    v[i] = gelu(tv);
                    if (EPI == 1) proj += bf2f(f2bf(v[i])) * wp[co];
                }
                if (EPI == 0 && ok) {
                    unsigned short sv[4];
#pragma unroll
                    for (int i = 0; i < 4; i++) sv[i] = f2bf(v[i]);
                    *reinterpret_cast<uint2 *>(out + pix * C + 16 * r + 4 * g) =
                        make_uint2(sv[0] | (unsigned)sv[1] << 16, sv[2] | (unsigned)sv[3] << 16);
                }
            }
            if (EPI == 1) {
                // the 64 channels of pixel j are spread over lanes j, j + 16, j + 32, j + 48
                proj += __shfl_xor(proj, 16);
                proj += __shfl_xor(proj, 32);
                if (g == 0 && ok) phi[pix] = (proj + bp) * oscale;
            }
        }
    }
}

// ----------------------------------------------------- batched GEMM --
// C[b][m][n] = sum_k A[b][m][k] B[b][k][n] with arbitrary element strides;
// A and B of type TA, TB (float or bf16 bits), staged to BF16 in LDS; FP32
// accumulation on the matrix cores; C stored as TC.  Tile 64 x 64, k step
// 32, four waves of 32 x 32.
struct GemmArgs {
    const void *A; long sam, sak, sab;
    const void *B; long sbk, sbn, sbb;
    void *C; long scm, scn, scb;
    int M, N, K;
};
template <typename T> __device__ __forceinline__ unsigned short to_bf(T v);
template <> __device__ __forceinline__ unsigned short to_bf<float>(float v) { return f2bf(v); }
template <> __device__ __forceinline__ unsigned short to_bf<unsigned short>(unsigned short v) { return v; }

template <typename TA, typename TB, typename TC>
__global__ __launch_bounds__(256) void sur_gemm_kernel(GemmArgs g) {
    constexpr int KS = 32, KP = KS + 8;
    __shared__ __attribute__((aligned(16))) unsigned short As[64][KP];
    __shared__ __attribute__((aligned(16))) unsigned short Bs[64][KP];  // [n][k]
    const TA *A = static_cast<const TA *>(g.A) + (long)blockIdx.z * g.sab;
    const TB *B = static_cast<const TB *>(g.B) + (long)blockIdx.z * g.sbb;
    TC *Cm = static_cast<TC *>(g.C) + (long)blockIdx.z * g.scb;
    const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, j = lane & 15, gq = lane >> 4;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    f32x4 acc[2][2];
#pragma unroll
    for (int r = 0; r < 2; r++)
#pragma unroll
        for (int c = 0; c < 2; c++) acc[r][c] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < g.K; k0 += KS) {
        // 64 x 32 of A and 32 x 64 of B: 8 elements of each per thread
        TA va[8];
        TB vb[8];
#pragma unroll
        for (int t = 0; t < 8; t++) {
            const int e = threadIdx.x + 256 * t, am = e / KS, ak = e % KS;
            const int m = m0 + am, k = k0 + ak;
            va[t] = (m < g.M && k < g.K) ? A[(long)m * g.sam + (long)k * g.sak] : (TA)0;
            const int bn = e % 64, bk = e / 64, n = n0 + bn, kk = k0 + bk;
            vb[t] = (n < g.N && kk < g.K) ? B[(long)kk * g.sbk + (long)n * g.sbn] : (TB)0;
        }
        __syncthreads();
#pragma unroll
        for (int t = 0; t < 8; t++) {
            const int e = threadIdx.x + 256 * t;
            As[e / KS][e % KS] = to_bf<TA>(va[t]);
            Bs[e % 64][e / 64] = to_bf<TB>(vb[t]);
        }
        __syncthreads();
        bf16x8 a[2], b[2];
#pragma unroll
        for (int r = 0; r < 2; r++) a[r] = *reinterpret_cast<const bf16x8 *>(&As[wm + 16 * r + j][8 * gq]);
#pragma unroll
        for (int c = 0; c < 2; c++) b[c] = *reinterpret_cast<const bf16x8 *>(&Bs[wn + 16 * c + j][8 * gq]);
#pragma unroll
        for (int r = 0; r < 2; r++)
#pragma unroll
            for (int c = 0; c < 2; c++) acc[r][c] = mfma_bf16(a[r], b[c], acc[r][c]);
    }
#pragma unroll
    for (int r = 0; r < 2; r++)
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int m = m0 + wm + 16 * r + 4 * gq + i, n = n0 + wn + 16 * c + j;
                if (m < g.M && n < g.N) {
                    if constexpr (sizeof(TC) == 4) Cm[(long)m * g.scm + (long)n * g.scn] = acc[r][c][i];
                    else Cm[(long)m * g.scm + (long)n * g.scn] = f2bf(acc[r][c][i]);
                }
            }
}

// per-mode channel mixing: G[z][mode][co] = sum_ci H[z][mode][ci] Wm[mode][co][ci]
// (one workgroup of 64 threads per (mode, layer); FP32 in and out, BF16 weights)
__global__ __launch_bounds__(64) void sur_mix_kernel(const float *__restrict__ H, const unsigned short *__restrict__ wm,
                                                     float *__restrict__ G, int nmode) {
    const int mode = blockIdx.x, z = blockIdx.y, co = threadIdx.x;
    const float *h = H + ((long)z * nmode + mode) * C;
    const unsigned short *wr = wm + ((long)mode * C + co) * C;
    float s = 0.0f;
#pragma unroll 8
    for (int ci = 0; ci < C; ci++) s += bf2f(wr[ci]) * h[ci];
    G[((long)z * nmode + mode) * C + co] = s;
}

// relative L2 residual partials of a against b: [nblk][2] = (sum (a-b)^2, sum b^2)
__global__ __launch_bounds__(256) void sur_residual_kernel(const float *__restrict__ a, const float *__restrict__ b,
                                                           long n, double *__restrict__ part) {
    __shared__ double s0[256], s1[256];
    double d2 = 0.0, b2 = 0.0;
    for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256L) {
        const double d = (double)a[i] - (double)b[i];
        d2 += d * d;
        b2 += (double)b[i] * (double)b[i];
    }
    s0[threadIdx.x] = d2;
    s1[threadIdx.x] = b2;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) { s0[threadIdx.x] += s0[threadIdx.x + h]; s1[threadIdx.x] += s1[threadIdx.x + h]; }
        __syncthreads();
    }
    if (threadIdx.x == 0) { part[2 * blockIdx.x] = s0[0]; part[2 * blockIdx.x + 1] = s1[0]; }
}

template <typename TA, typename TB, typename TC>
int gemm(const GemmArgs &g, int batch, hipStream_t s) {
    dim3 grid((g.N + 63) / 64, (g.M + 63) / 64, batch);
    hipLaunchKernelGGL((sur_gemm_kernel<TA, TB, TC>), grid, dim3(256), 0, s, g);
    MMS_HIP(hipGetLastError());
    return 0;
}

// SplitMix64 -> uniform in [-a, a)
struct Rng {
    uint64_t s;
    double next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return (double)(z >> 11) * (1.0 / 9007199254740992.0);
    }
    float sym(double a) { return (float)((2.0 * next() - 1.0) * a); }
};

unsigned short host_bf(float x) {
    unsigned u;
    memcpy(&u, &x, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

}  // namespace

// ------------------------------------------------------------ host plan --
// Weight vector layout (FP32, in this order; surrogate_weight_count):
//   lift   [64 co][3 planes][3 ky][3 kx] + bias [64]
//   conv1  [64 co][64 ci][3][3] + bias [64]
//   conv2  [64 co][64 ci][3][3] + bias [64]
//   mode   [32 ky][32 kx][64 co][64 ci]
//   conv3  [64 co][64 ci][3][3] + bias [64]
//   proj   [64] + bias [1]
long surrogate_weight_count() {
    return (long)C * 27 + C + 3L * (C * C * 9 + C) + 32L * 32 * C * C + C + 1;
}

int SurrogatePlan::init(const Grid &g_, const float *weights) {
    release();
    g = g_;
    const long n = g.n(), npix = n;
    nmode = 32 * 32;
    std::vector<float> wv((size_t)surrogate_weight_count());
    if (weights) {
        std::copy(weights, weights + wv.size(), wv.begin());
    } else {
        // seeded He-style initialisation (the network untrained)
        Rng r{20260u};
        size_t o = 0;
        for (int i = 0; i < C * 27; i++) wv[o++] = r.sym(sqrt(6.0 / 27));
        for (int i = 0; i < C; i++) wv[o++] = 0.0f;
        for (int l = 0; l < 2; l++) {
            for (int i = 0; i < C * C * 9; i++) wv[o++] = r.sym(sqrt(6.0 / (C * 9)));
            for (int i = 0; i < C; i++) wv[o++] = 0.0f;
        }
        for (long i = 0; i < 32L * 32 * C * C; i++) wv[o++] = r.sym(1.0 / C);
        for (int i = 0; i < C * C * 9; i++) wv[o++] = r.sym(sqrt(6.0 / (C * 9)));
        for (int i = 0; i < C; i++) wv[o++] = 0.0f;
        for (int i = 0; i < C; i++) wv[o++] = r.sym(sqrt(3.0 / C));
        wv[o++] = 0.0f;
    }
    size_t o = 0;
    // lift: [co][k = tap * 3 + plane], k padded to 32
    std::vector<unsigned short> lift((size_t)C * 32, 0);
    for (int co = 0; co < C; co++)
        for (int pl = 0; pl < 3; pl++)
            for (int tap = 0; tap < 9; tap++) lift[co * 32 + tap * 3 + pl] = host_bf(wv[o + (co * 3 + pl) * 9 + tap]);
    o += C * 27;
    std::vector<float> bl(wv.begin() + o, wv.begin() + o + C);
    o += C;
    auto conv = [&](std::vector<unsigned short> &w, std::vector<float> &b) {
        w.assign((size_t)9 * C * C, 0);  // [tap][co][ci]
        for (int co = 0; co < C; co++)
            for (int ci = 0; ci < C; ci++)
                for (int tap = 0; tap < 9; tap++) w[((size_t)tap * C + co) * C + ci] = host_bf(wv[o + (co * C + ci) * 9 + tap]);
        o += (size_t)C * C * 9;
        b.assign(wv.begin() + o, wv.begin() + o + C);
        o += C;
    };
    std::vector<unsigned short> w1, w2, w3, wm((size_t)nmode * C * C);
    std::vector<float> b1, b2, b3;
    conv(w1, b1);
    conv(w2, b2);
    for (size_t i = 0; i < wm.size(); i++) wm[i] = host_bf(wv[o + i]);  // [mode][co][ci]
    o += wm.size();
    conv(w3, b3);
    std::vector<float> wpv(wv.begin() + o, wv.begin() + o + C);
    o += C;
    bp = wv[o++];
    // real Fourier bases (BF16): Bx[k][x] over the padded width, By[k][y]
    // rows scaled by sqrt(2 / N), so that Bx Bx^T is close to the identity
    auto basis = [](int N) {
        std::vector<unsigned short> b((size_t)32 * N);
        const double sc = sqrt(2.0 / N);
        for (int k = 0; k < 32; k++)
            for (int x = 0; x < N; x++) {
                const int f = (k + 1) / 2;
                const double a = 2.0 * M_PI * f * x / N;
                const double v = (k == 0) ? 1.0 : ((k & 1) ? cos(a) : sin(a));
                b[(size_t)k * N + x] = host_bf((float)(sc * v));
            }
        return b;
    };
    std::vector<unsigned short> bx = basis(g.xd), by = basis(g.yd);
    auto up = [](void **dst, const void *src, size_t bytes) -> int {
        MMS_HIP(hipMalloc(dst, bytes));
        MMS_HIP(hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice));
        return 0;
    };
    if (up((void **)&d_lift, lift.data(), lift.size() * 2) || up((void **)&d_blift, bl.data(), bl.size() * 4) ||
        up((void **)&d_w1, w1.data(), w1.size() * 2) || up((void **)&d_b1, b1.data(), b1.size() * 4) ||
        up((void **)&d_w2, w2.data(), w2.size() * 2) || up((void **)&d_b2, b2.data(), b2.size() * 4) ||
        up((void **)&d_w3, w3.data(), w3.size() * 2) || up((void **)&d_b3, b3.data(), b3.size() * 4) ||
        up((void **)&d_wm, wm.data(), wm.size() * 2) || up((void **)&d_wp, wpv.data(), wpv.size() * 4) ||
        up((void **)&d_bx, bx.data(), bx.size() * 2) || up((void **)&d_by, by.data(), by.size() * 2))
        return 1;
    MMS_HIP(hipMalloc(&d_h0, sizeof(unsigned short) * npix * C));
    MMS_HIP(hipMalloc(&d_h1, sizeof(unsigned short) * npix * C));
    MMS_HIP(hipMalloc(&d_spec, sizeof(unsigned short) * npix * C));
    MMS_HIP(hipMalloc(&d_x1, sizeof(float) * (size_t)g.zd * g.yd * 32 * C));
    MMS_HIP(hipMalloc(&d_hm, sizeof(float) * (size_t)g.zd * nmode * C));
    MMS_HIP(hipMalloc(&d_gm, sizeof(float) * (size_t)g.zd * nmode * C));
    MMS_HIP(hipMalloc(&d_x2, sizeof(float) * (size_t)g.zd * g.yd * 32 * C));
    MMS_HIP(hipMalloc(&d_part, sizeof(double) * 2 * 1024));
    return 0;
}

void SurrogatePlan::release() {
    void *bufs[] = {d_lift, d_blift, d_w1, d_b1, d_w2, d_b2, d_w3, d_b3, d_wm, d_wp, d_bx, d_by,
                    d_h0,   d_h1,    d_spec, d_x1, d_hm, d_gm, d_x2, d_part};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    d_lift = d_w1 = d_w2 = d_w3 = d_wm = d_bx = d_by = nullptr;
    d_h0 = d_h1 = d_spec = nullptr;
    d_blift = d_b1 = d_b2 = d_b3 = d_wp = d_x1 = d_hm = d_gm = d_x2 = nullptr;
    d_part = nullptr;
}

int surrogate_potential(SurrogatePlan &sp, const float *rho, const uint8_t *matl, float *phi, float rscale,
                        float oscale, hipStream_t s, Profiler *prof) {
    const Grid &g = sp.g;
    MMS_CHECK(sp.d_h0 != nullptr, "surrogate not initialised");
    const int X = g.xd, Y = g.yd, Z = g.zd;
    const long npix = g.n();
    dim3 grid((X + TS - 1) / TS, (Y + TS - 1) / TS, Z), blk(256);
    const size_t conv_lds = sizeof(unsigned short) * ((size_t)HS * HS * CP + 9 * C * CP);
    static int attr = 0;
    if (!attr) {
        MMS_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&sur_conv_kernel<0>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)conv_lds));
        MMS_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&sur_conv_kernel<1>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)conv_lds));
        attr = 1;
    }
    // convolutions: one resident workgroup per CU per layer slice walks the tiles
    const int ntile = ((X + TS - 1) / TS) * ((Y + TS - 1) / TS);
    int per_layer = (256 + Z - 1) / Z;
    if (per_layer > ntile) per_layer = ntile;
    dim3 cgrid(per_layer, 1, Z);
    KTimer kt(prof, K_SURROGATE);
    hipLaunchKernelGGL(sur_lift_kernel, grid, blk, 0, s, rho, matl, sp.d_lift, sp.d_blift, sp.d_h0, X, Y, Z, rscale);
    MMS_HIP(hipGetLastError());
    hipLaunchKernelGGL((sur_conv_kernel<0>), cgrid, blk, conv_lds, s, sp.d_h0, sp.d_w1, sp.d_b1,
                       (const unsigned short *)nullptr, sp.d_h1, (const float *)nullptr, 0.0f, (float *)nullptr, 0.0f,
                       X, Y);
    MMS_HIP(hipGetLastError());
    // spectral branch of h1: x transform, y transform, mixing, inverse y, inverse x
    {
        GemmArgs a{};  // X1[z][y][kx][c] = sum_x Bx[kx][x] h1[z][y][x][c]
        a.A = sp.d_bx; a.sam = X; a.sak = 1; a.sab = 0;
        a.B = sp.d_h1; a.sbk = C; a.sbn = 1; a.sbb = (long)X * C;
        a.C = sp.d_x1; a.scm = C; a.scn = 1; a.scb = 32L * C;
        a.M = 32; a.N = C; a.K = X;
        if (gemm<unsigned short, unsigned short, float>(a, Z * Y, s)) return 1;
    }
    {
        GemmArgs a{};  // H[z][ky][(kx, c)] = sum_y By[ky][y] X1[z][y][(kx, c)]
        a.A = sp.d_by; a.sam = Y; a.sak = 1; a.sab = 0;
        a.B = sp.d_x1; a.sbk = 32L * C; a.sbn = 1; a.sbb = (long)Y * 32 * C;
        a.C = sp.d_hm; a.scm = 32L * C; a.scn = 1; a.scb = (long)sp.nmode * C;
        a.M = 32; a.N = 32 * C; a.K = Y;
        if (gemm<unsigned short, float, float>(a, Z, s)) return 1;
    }
    hipLaunchKernelGGL(sur_mix_kernel, dim3(sp.nmode, Z), dim3(64), 0, s, sp.d_hm, sp.d_wm, sp.d_gm, sp.nmode);
    MMS_HIP(hipGetLastError());
    {
        GemmArgs a{};  // X2[z][y][(kx, c)] = sum_ky By[ky][y] G[z][ky][(kx, c)]
        a.A = sp.d_by; a.sam = 1; a.sak = Y; a.sab = 0;
        a.B = sp.d_gm; a.sbk = 32L * C; a.sbn = 1; a.sbb = (long)sp.nmode * C;
        a.C = sp.d_x2; a.scm = 32L * C; a.scn = 1; a.scb = (long)Y * 32 * C;
        a.M = Y; a.N = 32 * C; a.K = 32;
        if (gemm<unsigned short, float, float>(a, Z, s)) return 1;
    }
    {
        GemmArgs a{};  // S[z][y][x][c] = sum_kx Bx[kx][x] X2[z][y][kx][c]  (BF16 NHWC)
        a.A = sp.d_bx; a.sam = 1; a.sak = X; a.sab = 0;
        a.B = sp.d_x2; a.sbk = C; a.sbn = 1; a.sbb = 32L * C;
        a.C = sp.d_spec; a.scm = C; a.scn = 1; a.scb = (long)X * C;
        a.M = X; a.N = C; a.K = 32;
        if (gemm<unsigned short, float, unsigned short>(a, Z * Y, s)) return 1;
    }
    hipLaunchKernelGGL((sur_conv_kernel<0>), cgrid, blk, conv_lds, s, sp.d_h1, sp.d_w2, sp.d_b2, sp.d_spec, sp.d_h0,
                       (const float *)nullptr, 0.0f, (float *)nullptr, 0.0f, X, Y);
    MMS_HIP(hipGetLastError());
    hipLaunchKernelGGL((sur_conv_kernel<1>), cgrid, blk, conv_lds, s, sp.d_h0, sp.d_w3, sp.d_b3,
                       (const unsigned short *)nullptr, (unsigned short *)nullptr, sp.d_wp, sp.bp, phi, oscale, X, Y);
    MMS_HIP(hipGetLastError());
    (void)npix;
    return 0;
}

int surrogate_residual(SurrogatePlan &sp, const float *a, const float *b, long n, double *r, hipStream_t s) {
    hipLaunchKernelGGL(sur_residual_kernel, dim3(1024), dim3(256), 0, s, a, b, n, sp.d_part);
    MMS_HIP(hipGetLastError());
    if (!r) return 0;
    std::vector<double> h(2048);
    MMS_HIP(hipMemcpyAsync(h.data(), sp.d_part, sizeof(double) * 2048, hipMemcpyDeviceToHost, s));
    MMS_HIP(hipStreamSynchronize(s));
    double d2 = 0, b2 = 0;
    for (int i = 0; i < 1024; i++) { d2 += h[2 * i]; b2 += h[2 * i + 1]; }
    *r = b2 > 0 ? sqrt(d2 / b2) : 0.0;
    return 0;
}

}  // namespace mms
