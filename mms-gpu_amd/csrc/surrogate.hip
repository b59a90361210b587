// surrogate.hip -- the learned far-field surrogate on gfx950: inference and
// training kernels.
//
// The network maps a film's charge density to its scalar potential, layer by
// layer in the plane; the layers see each other through the input planes:
//   lift      3x3 conv, 3 -> 64 channels + bias, GELU; the input planes of
//             layer z are its own charge, the summed charges of the layers at
//             distance 1 and the summed charges of the layers at distance
//             >= 2 of the same film, each times the charge scale
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
//
// Training keeps the FP32 parameters and saves the BF16 pre-activations of
// the forward pass.  The backward data pass of a 3x3 conv is the same conv
// kernel with flipped, transposed weights W'[tap][ci][co] = W[8 - tap][co][ci]
// and an epilogue that multiplies by GELU'(pre-activation); the weight
// gradient dW[tap][co][ci] = sum_p dz[p][co] a[p + off(tap)][ci] is a GEMM
// over pixels on the matrix cores; the spectral branch's backward pass is the
// same transform chain with the transposed mode mixing; Adam updates the FP32
// parameters, which are then rounded to the BF16 device copies.
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
// round to nearest even (v_cvt_pk_bf16_f32)
__device__ __forceinline__ unsigned short f2bf(float x) { return __builtin_bit_cast(unsigned short, (__bf16)x); }
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ unsigned pack2(float a, float b) {
    return __builtin_bit_cast(unsigned, __builtin_convertvector((f32x2{a, b}), bf16x2));
}
__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float((unsigned)h << 16); }
// GELU, tanh form: 0.5 x (1 + tanh(u)) = x s, s = 1 / (1 + exp(-2u)), u =
// sqrt(2/pi) (x + 0.044715 x^3), with the hardware exp2 and reciprocal
// (within a few ulp of the library tanh form; large |x| gives x or -0):
// exp(-2u) = 2^(x (a + b x^2)), a = -2 sqrt(2/pi) log2(e), b = 0.044715 a
__device__ __forceinline__ float gelu_s(float x) {
    const float q = x * fmaf(x * x, -0.10294324f, -2.3022082f);
    return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(q));
}
__device__ __forceinline__ float gelu(float x) { return x * gelu_s(x); }
// d/dx of the same form: s + 2 x s (1 - s) sqrt(2/pi) (1 + 3 * 0.044715 x^2)
__device__ __forceinline__ float gelu_grad(float x) {
    const float s = gelu_s(x);
    return s + 2.0f * x * s * (1.0f - s) * 0.7978845608f * (1.0f + 0.134145f * x * x);
}
__device__ __forceinline__ uint2 pack4(const float (&v)[4]) { return make_uint2(pack2(v[0], v[1]), pack2(v[2], v[3])); }
__device__ __forceinline__ void unpack4(uint2 u, float (&v)[4]) {
    v[0] = bf2f((unsigned short)(u.x & 0xFFFF));
    v[1] = bf2f((unsigned short)(u.x >> 16));
    v[2] = bf2f((unsigned short)(u.y & 0xFFFF));
    v[3] = bf2f((unsigned short)(u.y >> 16));
}

constexpr int C = 64;     // hidden channels
constexpr int CP = 80;    // LDS pixel / weight-row stride in bf16: 160-byte rows keep the 16-byte fragment reads free of bank conflicts (144-byte rows: 4 conflict cycles per read)
constexpr int TS = 16;    // output tile side (pixels)
constexpr int HS = TS + 2;

// FP32 parameters in device layout (surrogate_param_count):
//   lift [64 co][32 k] (k = tap * 3 + plane, 27 used) + bias [64]
//   conv1, conv2 [9 tap][64 co][64 ci] + bias [64] each
//   mode mixing [1024 mode][64 co][64 ci]
//   conv3 [9][64][64] + bias [64], projection [64] + bias [1]
constexpr long CONVN = 9L * C * C;
constexpr long OFF_L = 0, OFF_BL = OFF_L + 32L * C, OFF_W1 = OFF_BL + C, OFF_W2 = OFF_W1 + CONVN + C,
               OFF_WM = OFF_W2 + CONVN + C, OFF_W3 = OFF_WM + 1024L * C * C, OFF_WP = OFF_W3 + CONVN + C,
               OFF_BP = OFF_WP + C, NPARAM = OFF_BP + 1;
constexpr long OFF_W[3] = {OFF_W1, OFF_W2, OFF_W3};

// input planes of layer z at (gx, gy): own charge, layers at distance 1,
// layers at distance >= 2 of the film of zf layers (times rscale); the
// film's charges at (gx, gy) are given in r[0 .. zf)
template <int ZM>
__device__ __forceinline__ void combine_planes(const float (&r)[ZM], int zl, int zf, float rscale, float (&v)[3]) {
    float p0 = 0.0f, p1 = 0.0f, p2 = 0.0f;
#pragma unroll
    for (int k = 0; k < ZM; k++) {
        if (k >= zf) break;
        const int d = k > zl ? k - zl : zl - k;
        if (d == 0) p0 = r[k];
        else if (d == 1) p1 += r[k];
        else p2 += r[k];
    }
    v[0] = p0 * rscale;
    v[1] = p1 * rscale;
    v[2] = p2 * rscale;
}
__device__ __forceinline__ void input_planes(const float *__restrict__ rho, int z, int zf, long plane, long c2,
                                             float rscale, float (&v)[3]) {
    const int f0 = (z / zf) * zf, zl = z - f0;
    float p0 = 0.0f, p1 = 0.0f, p2 = 0.0f;
    for (int k = 0; k < zf; k++) {
        const float r = rho[(long)(f0 + k) * plane + c2];
        const int d = k > zl ? k - zl : zl - k;
        if (d == 0) p0 = r;
        else if (d == 1) p1 += r;
        else p2 += r;
    }
    v[0] = p0 * rscale;
    v[1] = p1 * rscale;
    v[2] = p2 * rscale;
}

// the 18 x 18 window of the three input planes around the tile (x0, y0) of
// layer z; films of up to 8 layers issue every charge load of a thread's
// pixels before combining them (the same sums in the same order)
__device__ __forceinline__ void stage_planes(const float *__restrict__ rho, float (*in)[HS][HS], int z, int zf,
                                             int x0, int y0, int xd, int yd, float rscale) {
    const long plane = (long)xd * yd;
    constexpr int ZM = 8, NP = (HS * HS + 255) / 256;
    if (zf <= ZM) {
        const int f0 = (z / zf) * zf, zl = z - f0;
        float r[NP][ZM];
#pragma unroll
        for (int q = 0; q < NP; q++) {
            const int e = threadIdx.x + 256 * q, sy = e / HS, sx = e % HS, gx = x0 - 1 + sx, gy = y0 - 1 + sy;
            const bool ok = e < HS * HS && gx >= 0 && gx < xd && gy >= 0 && gy < yd;
            const long c2 = (long)gy * xd + gx;
#pragma unroll
            for (int k = 0; k < ZM; k++) r[q][k] = (ok && k < zf) ? rho[(long)(f0 + k) * plane + c2] : 0.0f;
        }
#pragma unroll
        for (int q = 0; q < NP; q++) {
            const int e = threadIdx.x + 256 * q;
            if (e >= HS * HS) break;
            float v[3];
            combine_planes<ZM>(r[q], zl, zf, rscale, v);
            in[0][e / HS][e % HS] = v[0];
            in[1][e / HS][e % HS] = v[1];
            in[2][e / HS][e % HS] = v[2];
        }
        return;
    }
    for (int e = threadIdx.x; e < HS * HS; e += 256) {
        const int sy = e / HS, sx = e % HS, gx = x0 - 1 + sx, gy = y0 - 1 + sy;
        float v[3] = {0.0f, 0.0f, 0.0f};
        if (gx >= 0 && gx < xd && gy >= 0 && gy < yd) input_planes(rho, z, zf, plane, (long)gy * xd + gx, rscale, v);
        in[0][sy][sx] = v[0];
        in[1][sy][sx] = v[1];
        in[2][sy][sx] = v[2];
    }
}

// --------------------------------------------------------------- lift --
// 3x3 conv of the three input planes to 64 channels: one MFMA k step of 32
// (27 taps x planes, 5 zero) per 16 x 16 output tile; wave w takes tile rows
// 4w .. 4w + 3 (64 pixels).  wl: [64 co][32 k] bf16 (k = tap * 3 + plane),
// bias [64]; zsave (or null): BF16 pre-activations.
__global__ __launch_bounds__(256) void sur_lift_kernel(const float *__restrict__ rho,
                                                       const unsigned short *__restrict__ wl,
                                                       const float *__restrict__ bias, unsigned short *__restrict__ out,
                                                       unsigned short *__restrict__ zsave, int xd, int yd, int zf,
                                                       float rscale) {
    __shared__ float in[3][HS][HS];
    const int z = blockIdx.z, x0 = blockIdx.x * TS, y0 = blockIdx.y * TS;
    stage_planes(rho, in, z, zf, x0, y0, xd, yd, rscale);
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
            const int k = 8 * g + t, kc = k < 27 ? k : 0, tap = kc / 3, pl = kc % 3;  // kc: reads stay in the window
            bv[t] = (k < 27) ? f2bf(in[pl][py + tap / 3][px + tap % 3]) : (unsigned short)0;
        }
        const bf16x8 b = __builtin_bit_cast(bf16x8, bv);
        const int gx = x0 + px, gy = y0 + py;
        const bool ok = gx < xd && gy < yd;
        const long o = (((long)z * yd + gy) * xd + gx) * C;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const f32x4 acc = mfma_bf16(a[r], b, f32x4{0.f, 0.f, 0.f, 0.f});
            // D row 4 g + i of row tile r = output channel 16 r + 4 g + i of pixel j
            float zv[4], v[4];
#pragma unroll
            for (int i = 0; i < 4; i++) {
                zv[i] = acc[i] + bias[16 * r + 4 * g + i];
                v[i] = gelu(zv[i]);
            }
            if (ok) {
                *reinterpret_cast<uint2 *>(out + o + 16 * r + 4 * g) = pack4(v);
                if (zsave) *reinterpret_cast<uint2 *>(zsave + o + 16 * r + 4 * g) = pack4(zv);
            }
        }
    }
}

// ------------------------------------------------------------- conv 3x3 --
// 64 -> 64 channels on 16 x 16 pixel tiles.  A workgroup loads the nine
// 64 x 64 weight taps into LDS once (row stride CP) and then walks
// tiles blockIdx.x, blockIdx.x + gridDim.x, ... of layer blockIdx.z: the
// next tile's 18 x 18 x 64 input window is loaded into registers while the
// current one is multiplied, then stored to LDS.  Each of the eight waves
// computes D[32 co][64 pixels] (a quarter of the tile's rows, half its output
// channels) over 18 k steps (tap, 32 input channels).  w: [9 taps][64 co]
// [64 ci] bf16.  With x2t the spectral
// branch's inverse transform is added in the same accumulators: S[co][x] =
// sum_k X2[z][y][k][co] Bx[k][x], one 32-deep MFMA per 16 x 16 output tile
// from X2^T [z][y][64 co][32 k] and Bx^T [x][32 k] (BF16, global memory).
// Epilogues:
//   EPI 0: bias + GELU, BF16 NHWC out (zsave: pre-activation)
//   EPI 1: bias + GELU, then the 1x1 projection to one channel (weights wp,
//          bias *bp), times oscale, FP32 potential out (zsave as EPI 0)
//   EPI 2: backward data pass: sum * GELU'(zin), BF16 NHWC out
struct ConvArgs {
    const unsigned short *in, *w;
    const float *bias;
    const unsigned short *x2t, *bxt;
    unsigned short *out, *zsave;
    const unsigned short *zin;
    const float *wp, *bp;
    float *phi;
    float oscale;
    int xd, yd;
};
constexpr int WCH = HS * HS * 8;                 // 16-byte chunks of a window
constexpr size_t CONV_LDS = sizeof(unsigned short) * ((size_t)HS * HS * CP + 9 * C * CP);

// Software-pipelined over the tiles a workgroup walks: the epilogue of tile
// t - 1 (held in a second set of accumulators) is computed between the
// k steps of tile t, so its VALU work (GELU, conversions) and stores issue
// while the matrix cores run; the stores go through buffer resources whose
// range check discards the lanes outside the film (and every store of the
// first, empty, held tile).  512 threads, two waves per SIMD: wave w takes
// pixel rows 4 (w & 3) .. + 3 of the tile and output channels 32 (w >> 2)
// .. + 31, so each SIMD holds two waves whose products, epilogue arithmetic
// and LDS reads interleave.  With EPI 1 the two channel
// halves' projections of a pixel meet in LDS (half 0 + half 1) and are stored
// at the start of the next tile, after its first barrier.
constexpr int NT8 = 512, WPT8 = (WCH + NT8 - 1) / NT8;
constexpr size_t CONV8_LDS = CONV_LDS + sizeof(float) * 2 * 2 * 4 * 4 * 16;
template <int EPI, bool SPEC>
__global__ __launch_bounds__(512) void sur_conv8_kernel(ConvArgs A) {
    extern __shared__ __attribute__((aligned(16))) unsigned short lds[];
    unsigned short *win = lds;                 // [HS * HS][CP]
    unsigned short *wt = lds + HS * HS * CP;   // [9 * 64][CP]
    float *pproj = reinterpret_cast<float *>(lds + HS * HS * CP + 9 * C * CP);  // [slot][half][pg][c][j]
    const int xd = A.xd, yd = A.yd;
    const unsigned short *__restrict__ in = A.in;
    const int z = blockIdx.z, ntx = (xd + TS - 1) / TS, nty = (yd + TS - 1) / TS, ntile = ntx * nty;
    const int tid = threadIdx.x;
    // window position (sy << 16 | sx) of this thread's chunk q, -1 past the
    // window: computed once (the spectral backward pass, short of registers,
    // each time)
    constexpr bool HOIST = !(EPI == 2 && SPEC);
    int wpos[WPT8];
#pragma unroll
    for (int q = 0; q < WPT8; q++) {
        const int e = tid + NT8 * q, p = e >> 3;
        wpos[q] = (HOIST && e < WCH) ? ((p / HS) << 16) | (p % HS) : -1;
    }
    // the window chunks of tile t into registers: buffer loads whose range
    // check returns zeros for chunks outside the film (no branches)
    const int nin = (int)((long)gridDim.z * yd * xd * C * 2);
    const __amdgpu_buffer_rsrc_t r_in = __builtin_amdgcn_make_buffer_rsrc(const_cast<unsigned short *>(in), (short)0, nin, 0x00020000);
    auto load_win = [&](int t, uint4 (&v)[WPT8]) {
        const int x0 = (t % ntx) * TS, y0 = (t / ntx) * TS;
        const int base = ((z * yd + y0 - 1) * xd + x0 - 1) * C * 2 + 16 * (tid & 7);  // bytes (may be negative)
#pragma unroll
        for (int q = 0; q < WPT8; q++) {
            int pos = wpos[q];
            if constexpr (!HOIST) {
                const int e = tid + NT8 * q, p = e >> 3;
                pos = e < WCH ? ((p / HS) << 16) | (p % HS) : -1;
            }
            const int sy = pos >> 16, sx = pos & 0xFFFF;
            const int gx = x0 - 1 + sx, gy = y0 - 1 + sy;
            const bool ok = pos >= 0 && gx >= 0 && gx < xd && gy >= 0 && gy < yd;
            const int off = ok ? base + (sy * xd + sx) * C * 2 : nin;
            v[q] = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(r_in, off, 0, 0));
        }
    };
    {   // weights: 9 chunks per thread, all loads in flight before the stores
        constexpr int NW = 9 * C * 8 / NT8;
        uint4 v[NW];
#pragma unroll
        for (int q = 0; q < NW; q++) {
            const int e = tid + NT8 * q, row = e >> 3, ch = e & 7;
            v[q] = *reinterpret_cast<const uint4 *>(A.w + row * C + 8 * ch);
        }
#pragma unroll
        for (int q = 0; q < NW; q++) {
            const int e = tid + NT8 * q, row = e >> 3, ch = e & 7;
            *reinterpret_cast<uint4 *>(wt + row * CP + 8 * ch) = v[q];
        }
    }
    const int lane = tid & 63, wave = tid >> 6, j = lane & 15, g = lane >> 4;
    const int pg = wave & 3, hf = wave >> 2;  // pixel rows 4 pg .., channels 32 hf ..
    // per-lane constants of the epilogue: channels 32 hf + 16 r + 4 g + i
    float bias_c[2][4], wp_c[2][4];
#pragma unroll
    for (int r = 0; r < 2; r++)
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int co = 32 * hf + 16 * r + 4 * g + i;
            bias_c[r][i] = EPI == 2 ? 0.0f : A.bias[co];
            wp_c[r][i] = EPI == 1 ? A.wp[co] : 0.0f;
        }
    const float bp = EPI == 1 ? A.bp[0] : 0.0f;
    // buffer resources over the outputs (num_records 0 when absent: every store discarded)
    const long npix = (long)gridDim.z * yd * xd;
    const int nact = (int)(npix * C * 2);
    const __amdgpu_buffer_rsrc_t r_out = __builtin_amdgcn_make_buffer_rsrc(A.out, (short)0, A.out ? nact : 0, 0x00020000);
    const __amdgpu_buffer_rsrc_t r_zs = __builtin_amdgcn_make_buffer_rsrc(A.zsave, (short)0, A.zsave ? nact : 0, 0x00020000);
    constexpr bool spec = SPEC && EPI != 1;  // the spectral operands are given (x2t, bxt)
    // the held tile: accumulators, origin, saved pre-activations; `have` false
    // for the empty tile before the first (its stores are discarded)
    f32x4 hacc[2][4];
    uint2 hz[4][2];
#pragma unroll
    for (int r = 0; r < 2; r++)
#pragma unroll
        for (int c = 0; c < 4; c++) {
            hacc[r][c] = f32x4{0.f, 0.f, 0.f, 0.f};
            hz[c][r] = make_uint2(0u, 0u);
        }
    int hx0 = 0, hy0 = 0;
    bool have = false;
    float proj = 0.0f;
    int slot = 0;  // EPI 1: projection partials of the epilogue run in this iteration
    // one 4-channel block (r, c) of the held tile's epilogue
    auto epi_block = [&](int r, int c) {
        const int gx = hx0 + j, gy = hy0 + 4 * pg + c;
        const bool ok = have && gx < xd && gy < yd;
        const long pix = ((long)z * yd + gy) * xd + gx;
        const int bo = ok ? (int)((pix * C + 32 * hf + 16 * r + 4 * g) * 2) : nact;  // out of range if not ok
        float zz[4], v[4];
        if (EPI == 2) {
            unpack4(hz[c][r], zz);
#pragma unroll
            for (int i = 0; i < 4; i++) v[i] = hacc[r][c][i] * gelu_grad(zz[i]);
        } else {
#pragma unroll
            for (int i = 0; i < 4; i++) {
                zz[i] = hacc[r][c][i] + bias_c[r][i];
                v[i] = gelu(zz[i]);
                if (EPI == 1) proj += bf2f(f2bf(v[i])) * wp_c[r][i];
            }
            const uint2 zp = pack4(zz);
            __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, zp), r_zs, bo, 0, 0);
        }
        if (EPI != 1) {
            const uint2 vp = pack4(v);
            __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, vp), r_out, bo, 0, 0);
        }
        if (EPI == 1 && r == 1) {
            // this half's 32 channels of pixel j: lanes j, j + 16, j + 32, j + 48
            proj += __shfl_xor(proj, 16);
            proj += __shfl_xor(proj, 32);
            if (g == 0) pproj[(((slot * 2 + hf) * 4 + pg) * 4 + c) * 16 + j] = proj;
            proj = 0.0f;
        }
    };
    // EPI 1: the potential of the tile at (ex0, ey0) from the partials in slot sl
    auto store_phi = [&](int sl, int ex0, int ey0) {
        if (tid < 256) {
            const int row = tid >> 4, col = tid & 15, gx = ex0 + col, gy = ey0 + row;
            const float *pp = pproj + ((sl * 2) * 4 + (row >> 2)) * 64 + (row & 3) * 16 + col;
            const float v = (pp[0] + pp[4 * 4 * 16] + bp) * A.oscale;  // half 0 + half 1
            if (gx < xd && gy < yd) A.phi[((long)z * yd + gy) * xd + gx] = v;
        }
    };
    int ex0 = 0, ey0 = 0;
    bool ehave = false;  // EPI 1: partials of a real tile wait in slot 1 - slot
    uint4 nxt[WPT8];
    int t = blockIdx.x;
    if (t < ntile) load_win(t, nxt);
    for (; t < ntile; t += gridDim.x) {
        __syncthreads();  // the previous tile's products are done with win; its partials are in LDS
        if (EPI == 1 && ehave) store_phi(1 - slot, ex0, ey0);
#pragma unroll
        for (int q = 0; q < WPT8; q++) {
            const int e = tid + NT8 * q;
            if (e < WCH) *reinterpret_cast<uint4 *>(win + (e >> 3) * CP + 8 * (e & 7)) = nxt[q];
        }
        __syncthreads();
        if (t + (int)gridDim.x < ntile) load_win(t + gridDim.x, nxt);  // in flight during the products
        const int x0 = (t % ntx) * TS, y0 = (t / ntx) * TS;
        // the saved pre-activations and the spectral operands of this tile,
        // loaded now so the products hide their latency
        uint2 pzin[4][2];
        bf16x8 sa[4][2], sb;
#pragma unroll
        for (int c = 0; c < 4; c++) {
            const int gx = x0 + j, gy = y0 + 4 * pg + c;
            const bool ok = gx < xd && gy < yd;
            const long pix = ((long)z * yd + gy) * xd + gx;
            // rows and columns beyond the film read its last ones (discarded outputs)
            const long srow = (long)z * yd + (gy < yd ? gy : yd - 1);
#pragma unroll
            for (int r = 0; r < 2; r++) {
                pzin[c][r] = make_uint2(0u, 0u);
                if (EPI == 2 && ok) pzin[c][r] = *reinterpret_cast<const uint2 *>(A.zin + pix * C + 32 * hf + 16 * r + 4 * g);
                if (spec)
                    sa[c][r] = *reinterpret_cast<const bf16x8 *>(A.x2t + (srow * C + 32 * hf + 16 * r + j) * 32 + 8 * g);
            }
        }
        if (spec) sb = *reinterpret_cast<const bf16x8 *>(A.bxt + (long)(x0 + j < xd ? x0 + j : xd - 1) * 32 + 8 * g);
        f32x4 acc[2][4];
#pragma unroll
        for (int r = 0; r < 2; r++)
#pragma unroll
            for (int c = 0; c < 4; c++) acc[r][c] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int tap = 0; tap < 9; tap++) {
            const int dy = tap / 3, dx = tap % 3;
#pragma unroll
            for (int kb = 0; kb < 2; kb++) {
                bf16x8 a[2], b[4];
#pragma unroll
                for (int r = 0; r < 2; r++)
                    a[r] = *reinterpret_cast<const bf16x8 *>(wt + (tap * C + 32 * hf + 16 * r + j) * CP + 32 * kb + 8 * g);
#pragma unroll
                for (int c = 0; c < 4; c++)
                    b[c] = *reinterpret_cast<const bf16x8 *>(win + ((4 * pg + c + dy) * HS + j + dx) * CP + 32 * kb + 8 * g);
#pragma unroll
                for (int r = 0; r < 2; r++)
#pragma unroll
                    for (int c = 0; c < 4; c++) acc[r][c] = mfma_bf16(a[r], b[c], acc[r][c]);
                const int ks = 2 * tap + kb;
                if (ks < 16 && (ks & 1) == 0) epi_block((ks >> 1) & 1, ks >> 2);  // the held tile, block by block
                if (spec && ks == 9) {
                    // the spectral term, midway: its operands have arrived and the
                    // loop does not end on these products
#pragma unroll
                    for (int r = 0; r < 2; r++)
#pragma unroll
                        for (int c = 0; c < 4; c++) acc[r][c] = mfma_bf16(sa[c][r], sb, acc[r][c]);
                }
            }
        }
#pragma unroll
        for (int r = 0; r < 2; r++)
#pragma unroll
            for (int c = 0; c < 4; c++) {
                hacc[r][c] = acc[r][c];
                hz[c][r] = pzin[c][r];
            }
        ex0 = hx0; ey0 = hy0; ehave = have;  // the tile whose partials this iteration wrote
        slot = 1 - slot;
        hx0 = x0;
        hy0 = y0;
        have = true;
        // the next iteration starts by reading the accumulators these
        // products wrote: 19 wait states after the last MFMA, which the
        // compiler does not insert across the loop's back edge
        __builtin_amdgcn_sched_barrier(0);
        asm volatile("s_nop 7\n\ts_nop 7\n\ts_nop 2" ::: "memory");
    }
    // the last tiles: pending partials, then the held tile's epilogue
    if (EPI == 1) {
        __syncthreads();
        if (ehave) store_phi(1 - slot, ex0, ey0);
    }
#pragma unroll
    for (int b = 0; b < 8; b++) epi_block(b & 1, b >> 1);
    if (EPI == 1) {
        __syncthreads();
        if (have) store_phi(slot, hx0, hy0);
    }
}

// ------------------------------------------------- weight gradients --
// The gradient tile dz (16 x 16 pixels x 64 channels, NHWC) transposed into
// LDS as dzT[co][pixel] (row stride DZS): thread (row, half, channel block)
// loads 8 pixels x 8 channels and stores 8 channel rows of 8 pixels.
constexpr int DZS = 256 + 8;
// element k of 8 packed BF16 values
__device__ __forceinline__ unsigned bf_elem(const uint4 &v, int k) {
    const unsigned w = (k >> 1) == 0 ? v.x : (k >> 1) == 1 ? v.y : (k >> 1) == 2 ? v.z : v.w;
    return (k & 1) ? (w >> 16) : (w & 0xFFFFu);
}
// element k of each of v[0..7], packed (the k-th column of an 8 x 8 BF16 block)
__device__ __forceinline__ uint4 column8(const uint4 (&v)[8], int k) {
    uint4 o;
    o.x = bf_elem(v[0], k) | bf_elem(v[1], k) << 16;
    o.y = bf_elem(v[2], k) | bf_elem(v[3], k) << 16;
    o.z = bf_elem(v[4], k) | bf_elem(v[5], k) << 16;
    o.w = bf_elem(v[6], k) | bf_elem(v[7], k) << 16;
    return o;
}
__device__ __forceinline__ void stage_dzT(const unsigned short *__restrict__ dz, unsigned short *dzT, int z, int x0,
                                          int y0, int xd, int yd) {
    const int py = threadIdx.x >> 4, q = (threadIdx.x >> 3) & 1, cb = threadIdx.x & 7;
    uint4 v[8];
#pragma unroll
    for (int i = 0; i < 8; i++) {
        const int gx = x0 + 8 * q + i, gy = y0 + py;
        v[i] = make_uint4(0u, 0u, 0u, 0u);
        if (gx < xd && gy < yd) v[i] = *reinterpret_cast<const uint4 *>(dz + (((long)z * yd + gy) * xd + gx) * C + 8 * cb);
    }
#pragma unroll
    for (int k = 0; k < 8; k++) *reinterpret_cast<uint4 *>(dzT + (8 * cb + k) * DZS + py * 16 + 8 * q) = column8(v, k);
}

// bias gradient partial of thread t: channel t / 4, pixels 64 (t % 4) .. + 63 of dzT
__device__ __forceinline__ float dzT_rowsum(const unsigned short *dzT) {
    const int co = threadIdx.x >> 2, qq = threadIdx.x & 3;
    float s = 0.0f;
#pragma unroll
    for (int u = 0; u < 8; u++) {
        const uint4 v = *reinterpret_cast<const uint4 *>(dzT + co * DZS + 64 * qq + 8 * u);
        const unsigned w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int i = 0; i < 4; i++) s += bf2f((unsigned short)(w[i] & 0xFFFF)) + bf2f((unsigned short)(w[i] >> 16));
    }
    return s;
}

// dW[tap][co][ci] = sum_p dz[p][co] a[p + off(tap)][ci] (off(tap) = (tap / 3 - 1,
// tap % 3 - 1), a zero outside the film) and db[co] = sum_p dz[p][co] as MFMA
// GEMMs over pixels: M = co, N = ci, K = 32 pixels of two tile rows per k step
// (lane group g: row 2 kk + g / 2, columns 8 (g % 2) .. + 7).  The input window
// is held transposed in three copies shifted by dx (xc[dx][ci][wy][x] =
// a(wy, x + dx)), so that every operand fragment is one 16-byte LDS read.
// Wave w owns input channels 16 w .. 16 w + 15 for all taps and output
// channels (36 accumulator tiles).  Persistent: the workgroup walks tiles
// blockIdx.x, + gridDim.x, ... over all layers and writes its partial sums
// to part[blockIdx.x][9 * 64 * 64 + 64].
constexpr int XS = HS * 16 + 8;   // xc row stride per channel (296)
constexpr size_t WGRAD_LDS = sizeof(unsigned short) * ((size_t)C * DZS + 3 * C * XS);
constexpr long WGRAD_PART = CONVN + C;

__global__ __launch_bounds__(256) void sur_wgrad_kernel(const unsigned short *__restrict__ dz,
                                                        const unsigned short *__restrict__ a, float *__restrict__ part,
                                                        int xd, int yd, int nz) {
    extern __shared__ __attribute__((aligned(16))) unsigned short lds[];
    unsigned short *dzT = lds;               // [64 co][DZS]
    unsigned short *xc = lds + C * DZS;      // [3 dx][64 ci][XS]
    const int ntx = (xd + TS - 1) / TS, nty = (yd + TS - 1) / TS, npl = ntx * nty, ntile = npl * nz;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, j = lane & 15, g = lane >> 4;
    f32x4 acc[9][4];
#pragma unroll
    for (int tap = 0; tap < 9; tap++)
#pragma unroll
        for (int ct = 0; ct < 4; ct++) acc[tap][ct] = f32x4{0.f, 0.f, 0.f, 0.f};
    float dbs = 0.0f;
    for (int t = blockIdx.x; t < ntile; t += gridDim.x) {
        const int z = t / npl, tl = t % npl, x0 = (tl % ntx) * TS, y0 = (tl / ntx) * TS;
        __syncthreads();  // the previous tile's products are done with the LDS
        stage_dzT(dz, dzT, z, x0, y0, xd, yd);
        for (int item = threadIdx.x; item < 3 * HS * 16; item += 256) {
            const int cb = item & 7, q = (item >> 3) & 1, wy = (item >> 4) % HS, dx = (item >> 4) / HS;
            const int gy = y0 - 1 + wy;
            uint4 v[8];
#pragma unroll
            for (int i = 0; i < 8; i++) {
                const int gx = x0 - 1 + 8 * q + i + dx;
                v[i] = make_uint4(0u, 0u, 0u, 0u);
                if (gx >= 0 && gx < xd && gy >= 0 && gy < yd)
                    v[i] = *reinterpret_cast<const uint4 *>(a + (((long)z * yd + gy) * xd + gx) * C + 8 * cb);
            }
#pragma unroll
            for (int k = 0; k < 8; k++)
                *reinterpret_cast<uint4 *>(xc + (dx * C + 8 * cb + k) * XS + wy * 16 + 8 * q) = column8(v, k);
        }
        __syncthreads();
        dbs += dzT_rowsum(dzT);
#pragma unroll 1
        for (int kk = 0; kk < 8; kk++) {
            const int row = 2 * kk + (g >> 1), px0 = 8 * (g & 1);
            bf16x8 A[4];
#pragma unroll
            for (int ct = 0; ct < 4; ct++)
                A[ct] = *reinterpret_cast<const bf16x8 *>(dzT + (16 * ct + j) * DZS + row * 16 + px0);
#pragma unroll
            for (int tap = 0; tap < 9; tap++) {
                const int dy = tap / 3, dx = tap % 3;
                const bf16x8 B = *reinterpret_cast<const bf16x8 *>(xc + (dx * C + 16 * wave + j) * XS + (row + dy) * 16 + px0);
#pragma unroll
                for (int ct = 0; ct < 4; ct++) acc[tap][ct] = mfma_bf16(A[ct], B, acc[tap][ct]);
            }
        }
    }
    // D row 4 g + i of tile (tap, ct) = co 16 ct + 4 g + i; column j = ci 16 wave + j
    float *pp = part + (long)blockIdx.x * WGRAD_PART;
#pragma unroll
    for (int tap = 0; tap < 9; tap++)
#pragma unroll
        for (int ct = 0; ct < 4; ct++)
#pragma unroll
            for (int i = 0; i < 4; i++) pp[(long)tap * C * C + (16 * ct + 4 * g + i) * C + 16 * wave + j] = acc[tap][ct][i];
    dbs += __shfl_xor(dbs, 1);
    dbs += __shfl_xor(dbs, 2);
    if ((threadIdx.x & 3) == 0) pp[CONVN + (threadIdx.x >> 2)] = dbs;
}

// Lift weight gradient dWl[co][k] = sum_p dz0[p][co] x[p + off(tap)][plane]
// (k = tap * 3 + plane < 27, x the BF16-rounded input planes), dbl[co] =
// sum_p dz0[p][co]; M = co (wave w: 16 w .. + 15), N = k (two tiles), K =
// pixels as in sur_wgrad_kernel.  part[blockIdx.x][64 * 32 + 64].
constexpr long LGRAD_PART = 32L * C + C;
__global__ __launch_bounds__(256) void sur_lift_wgrad_kernel(const float *__restrict__ rho,
                                                             const unsigned short *__restrict__ dz,
                                                             float *__restrict__ part, int xd, int yd, int nz, int zf,
                                                             float rscale) {
    __shared__ __attribute__((aligned(16))) unsigned short dzT[C * DZS];
    __shared__ float in[3][HS][HS];
    const int ntx = (xd + TS - 1) / TS, nty = (yd + TS - 1) / TS, npl = ntx * nty, ntile = npl * nz;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, j = lane & 15, g = lane >> 4;
    f32x4 acc[2] = {f32x4{0.f, 0.f, 0.f, 0.f}, f32x4{0.f, 0.f, 0.f, 0.f}};
    float dbs = 0.0f;
    for (int t = blockIdx.x; t < ntile; t += gridDim.x) {
        const int z = t / npl, tl = t % npl, x0 = (tl % ntx) * TS, y0 = (tl / ntx) * TS;
        __syncthreads();
        stage_dzT(dz, dzT, z, x0, y0, xd, yd);
        stage_planes(rho, in, z, zf, x0, y0, xd, yd, rscale);
        __syncthreads();
        dbs += dzT_rowsum(dzT);
#pragma unroll 1
        for (int kk = 0; kk < 8; kk++) {
            const int row = 2 * kk + (g >> 1), px0 = 8 * (g & 1);
            const bf16x8 A = *reinterpret_cast<const bf16x8 *>(dzT + (16 * wave + j) * DZS + row * 16 + px0);
#pragma unroll
            for (int nt = 0; nt < 2; nt++) {
                const int k = 16 * nt + j, kc = k < 27 ? k : 0, tap = kc / 3, pl = kc % 3, dy = tap / 3, dx = tap % 3;
                u16x8 bv;
#pragma unroll
                for (int u = 0; u < 8; u++) bv[u] = (k < 27) ? f2bf(in[pl][row + dy][px0 + u + dx]) : (unsigned short)0;
                acc[nt] = mfma_bf16(A, __builtin_bit_cast(bf16x8, bv), acc[nt]);
            }
        }
    }
    float *pp = part + (long)blockIdx.x * LGRAD_PART;
#pragma unroll
    for (int nt = 0; nt < 2; nt++)
#pragma unroll
        for (int i = 0; i < 4; i++) pp[(16 * wave + 4 * g + i) * 32 + 16 * nt + j] = acc[nt][i];
    dbs += __shfl_xor(dbs, 1);
    dbs += __shfl_xor(dbs, 2);
    if ((threadIdx.x & 3) == 0) pp[32 * C + (threadIdx.x >> 2)] = dbs;
}

// out[i] = sum_b part[b * stride + i], i < len (fixed order: deterministic)
// out[i] = sum of part[b * stride + i] over the nb partials in a fixed
// order: four interleaved sums (b mod 4, four loads in flight per thread)
// added in order.  64 outputs per workgroup; launch with sum_blocks(len).
__global__ __launch_bounds__(256) void sur_sum_kernel(const float *__restrict__ part, long stride, int nb, long len,
                                                      float *__restrict__ out) {
    __shared__ float red[4][64];
    const int o = threadIdx.x & 63, q = threadIdx.x >> 6;
    const long i = blockIdx.x * 64L + o;
    float s = 0.0f;
    if (i < len) {
        int b = q;
        for (; b + 12 < nb; b += 16) {
            const float v0 = part[(long)b * stride + i], v1 = part[(long)(b + 4) * stride + i];
            const float v2 = part[(long)(b + 8) * stride + i], v3 = part[(long)(b + 12) * stride + i];
            s += v0; s += v1; s += v2; s += v3;
        }
        for (; b < nb; b += 4) s += part[(long)b * stride + i];
    }
    red[q][o] = s;
    __syncthreads();
    if (q == 0 && i < len) out[i] = ((red[0][o] + red[1][o]) + red[2][o]) + red[3][o];
}
inline unsigned sum_blocks(long len) { return (unsigned)((len + 63) / 64); }

// ---------------------------------------------------- loss and head --
// Per film f of npf cells: loss_f = sum (phi_s - phi_t)^2 inv_norm[f]; with dz3
// non-null also dphi = gscale (phi_s - phi_t) inv_norm[f], dproj = dphi
// oscale, dz3 = dproj wp GELU'(z3) and the partials of dwp = sum dproj
// GELU(z3) (BF16-rounded as the forward projection reads it) and dbp = sum
// dproj.  Grid (HB, films); thread = (pixel slot t / 8, channel block t % 8).
// part[(f * HB + blk) * HPART + (0..63 dwp, 64 dbp, 65 loss)].
constexpr int HB = SurrogatePlan::kHeadBlocks, HPART = SurrogatePlan::kHeadPart;
static_assert(HPART == C + 2, "head partials: C projection-weight gradients, the bias gradient, the loss");
__global__ __launch_bounds__(256) void sur_head_kernel(const float *__restrict__ phs, const float *__restrict__ pht,
                                                       const float *__restrict__ inv_norm, float gscale, float oscale,
                                                       const unsigned short *__restrict__ z3,
                                                       const float *__restrict__ wp, unsigned short *__restrict__ dz3,
                                                       long npf, float *__restrict__ part) {
    __shared__ float sw[32][C + 1];
    __shared__ float sl[32], sb[32];
    const int f = blockIdx.y, ch = threadIdx.x & 7, slot = threadIdx.x >> 3;
    const long base = (long)f * npf;
    const float inv = inv_norm[f];
    float dw[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f}, wv[8];
#pragma unroll
    for (int k = 0; k < 8; k++) wv[k] = wp[8 * ch + k];
    float ls = 0.0f, db = 0.0f;
    for (long p = (long)blockIdx.x * 32 + slot; p < npf; p += (long)HB * 32) {
        const float d = phs[base + p] - pht[base + p];
        ls += d * d * inv;
        if (!dz3) continue;
        const float dproj = gscale * d * inv * oscale;
        db += dproj;
        const uint4 zv = *reinterpret_cast<const uint4 *>(z3 + (base + p) * C + 8 * ch);
        const unsigned zw[4] = {zv.x, zv.y, zv.z, zv.w};
        unsigned ow[4];
#pragma unroll
        for (int h = 0; h < 4; h++) {
            unsigned short o2[2];
#pragma unroll
            for (int e = 0; e < 2; e++) {
                const int k = 2 * h + e;
                const float zz = bf2f((unsigned short)(e ? zw[h] >> 16 : zw[h] & 0xFFFF));
                dw[k] += dproj * bf2f(f2bf(gelu(zz)));
                o2[e] = f2bf(dproj * wv[k] * gelu_grad(zz));
            }
            ow[h] = o2[0] | (unsigned)o2[1] << 16;
        }
        *reinterpret_cast<uint4 *>(dz3 + (base + p) * C + 8 * ch) = make_uint4(ow[0], ow[1], ow[2], ow[3]);
    }
#pragma unroll
    for (int k = 0; k < 8; k++) sw[slot][8 * ch + k] = dw[k];
    if (ch == 0) { sl[slot] = ls; sb[slot] = db; }
    __syncthreads();
    float *pp = part + ((long)f * HB + blockIdx.x) * HPART;
    if (threadIdx.x < C) {
        float s = 0.0f;
        for (int q = 0; q < 32; q++) s += sw[q][threadIdx.x];
        pp[threadIdx.x] = s;
    } else if (threadIdx.x == C) {
        float s = 0.0f;
        for (int q = 0; q < 32; q++) s += sb[q];
        pp[C] = s;
    } else if (threadIdx.x == C + 1) {
        float s = 0.0f;
        for (int q = 0; q < 32; q++) s += sl[q];
        pp[C + 1] = s;
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
// integer round to nearest even (the same bits as f2bf): with the hardware
// conversion the compiler waits for each staged load inside its bounds branch
template <> __device__ __forceinline__ unsigned short to_bf<float>(float v) {
    unsigned u = __float_as_uint(v);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
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

// per-mode channel mixing (one workgroup of 64 threads per (mode, layer); FP32
// in and out, BF16 weights):
//   forward     G[z][mode][co] = sum_ci H[z][mode][ci] Wm[mode][co][ci]
//   transposed  G[z][mode][ci] = sum_co H[z][mode][co] Wm[mode][co][ci]
template <bool TR>
__global__ __launch_bounds__(64) void sur_mix_kernel(const float *__restrict__ H, const unsigned short *__restrict__ wm,
                                                     float *__restrict__ G, int nmode) {
    const int mode = blockIdx.x, z = blockIdx.y, o = threadIdx.x;
    const float *h = H + ((long)z * nmode + mode) * C;
    const unsigned short *wb = wm + (long)mode * C * C;
    float s = 0.0f;
#pragma unroll 8
    for (int k = 0; k < C; k++) s += bf2f(TR ? wb[k * C + o] : wb[o * C + k]) * h[k];
    G[((long)z * nmode + mode) * C + o] = s;
}

// dWm[mode][co][ci] = sum_z dG[z][mode][co] H[z][mode][ci]; one workgroup per mode
__global__ __launch_bounds__(256) void sur_mix_wgrad_kernel(const float *__restrict__ dG, const float *__restrict__ H,
                                                            float *__restrict__ dW, int nmode, int nz) {
    __shared__ float sg[C], sh[C];
    const int mode = blockIdx.x, ci = threadIdx.x & 63, co0 = threadIdx.x >> 6;
    float acc[16];
#pragma unroll
    for (int u = 0; u < 16; u++) acc[u] = 0.0f;
    for (int z = 0; z < nz; z++) {
        __syncthreads();
        if (threadIdx.x < C) sg[threadIdx.x] = dG[((long)z * nmode + mode) * C + threadIdx.x];
        else if (threadIdx.x < 2 * C) sh[threadIdx.x - C] = H[((long)z * nmode + mode) * C + threadIdx.x - C];
        __syncthreads();
        const float hv = sh[ci];
#pragma unroll
        for (int u = 0; u < 16; u++) acc[u] += sg[co0 + 4 * u] * hv;
    }
#pragma unroll
    for (int u = 0; u < 16; u++) dW[((long)mode * C + co0 + 4 * u) * C + ci] = acc[u];
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

// --------------------------------------------- parameters and Adam --
// BF16 copy of n FP32 values
__global__ __launch_bounds__(256) void sur_to_bf_kernel(const float *__restrict__ src, unsigned short *__restrict__ dst,
                                                        long n) {
    const long i = blockIdx.x * 256L + threadIdx.x;
    if (i < n) dst[i] = f2bf(src[i]);
}
// flipped, transposed BF16 taps of a conv: wt[tap][ci][co] = w[8 - tap][co][ci]
__global__ __launch_bounds__(256) void sur_flipT_kernel(const float *__restrict__ w, unsigned short *__restrict__ wt) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= CONVN) return;
    const int tap = i / (C * C), r = i % (C * C), ci = r / C, co = r % C;
    wt[i] = f2bf(w[((8 - tap) * C + co) * C + ci]);
}
// Adam (bias-corrected; c1 = 1 / (1 - b1^t), c2 = 1 / (1 - b2^t))
__global__ __launch_bounds__(256) void sur_adam_kernel(float *__restrict__ w, float *__restrict__ m,
                                                       float *__restrict__ v, const float *__restrict__ g, long n,
                                                       float lr, float b1, float b2, float c1, float c2, float eps) {
    const long i = blockIdx.x * 256L + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
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

int launch_wgrad(SurrogatePlan &sp, const unsigned short *dz, const unsigned short *a, float *grad, hipStream_t s) {
    const Grid &g = sp.g;
    const int ntile = ((g.xd + TS - 1) / TS) * ((g.yd + TS - 1) / TS) * g.zd;
    const int nb = ntile < SurrogatePlan::kPartBlocks ? ntile : SurrogatePlan::kPartBlocks;
    hipLaunchKernelGGL(sur_wgrad_kernel, dim3(nb), dim3(256), WGRAD_LDS, s, dz, a, sp.d_wpart, g.xd, g.yd, g.zd);
    MMS_HIP(hipGetLastError());
    hipLaunchKernelGGL(sur_sum_kernel, dim3(sum_blocks(WGRAD_PART)), dim3(256), 0, s, sp.d_wpart,
                       WGRAD_PART, nb, WGRAD_PART, grad);
    MMS_HIP(hipGetLastError());
    return 0;
}

int launch_conv(int epi, const SurrogatePlan &sp, const ConvArgs &a, hipStream_t s) {
    const Grid &g = sp.g;
    // the kernel's output buffer resources take 32-bit byte sizes
    MMS_CHECK((long)g.n() * C * 2 < (1L << 31), "surrogate: grid too large for one activation buffer");
    const int ntile = ((g.xd + TS - 1) / TS) * ((g.yd + TS - 1) / TS);
    // one resident workgroup per CU per layer slice walks the tiles
    int per_layer = (256 + g.zd - 1) / g.zd;
    if (per_layer > ntile) per_layer = ntile;
    dim3 grid(per_layer, 1, g.zd);
    const bool spec = a.x2t != nullptr;
    if (epi == 0 && spec) hipLaunchKernelGGL((sur_conv8_kernel<0, true>), grid, dim3(NT8), CONV8_LDS, s, a);
    else if (epi == 0) hipLaunchKernelGGL((sur_conv8_kernel<0, false>), grid, dim3(NT8), CONV8_LDS, s, a);
    else if (epi == 1) hipLaunchKernelGGL((sur_conv8_kernel<1, false>), grid, dim3(NT8), CONV8_LDS, s, a);
    else if (spec) hipLaunchKernelGGL((sur_conv8_kernel<2, true>), grid, dim3(NT8), CONV8_LDS, s, a);
    else hipLaunchKernelGGL((sur_conv8_kernel<2, false>), grid, dim3(NT8), CONV8_LDS, s, a);
    MMS_HIP(hipGetLastError());
    return 0;
}

// X2^T [z][y][64 c][32 k] (BF16) of X2 [z][y][32 k][64 c] (FP32): one (z, y)
// row per workgroup through LDS
__global__ __launch_bounds__(256) void sur_x2t_kernel(const float *__restrict__ x2, unsigned short *__restrict__ x2t) {
    __shared__ float t[32][C + 1];
    const float *src = x2 + (long)blockIdx.x * 32 * C;
    for (int e = threadIdx.x; e < 32 * C; e += 256) t[e / C][e % C] = src[e];
    __syncthreads();
    unsigned short *dst = x2t + (long)blockIdx.x * 32 * C;
    for (int e = threadIdx.x; e < 32 * C; e += 256) dst[e] = f2bf(t[e % 32][e / 32]);
}

// Forward transform over x of every (z, y) row of a BF16 NHWC array:
// X1[row][k][c] = sum_x Bx[k][x] in[row][x][c] (FP32).  Persistent
// workgroups walk (row, 128-pixel chunk) in order.  A chunk's 128 x 64 values
// are loaded as 16-byte pieces two chunks ahead and stored transposed into
// LDS as [c][x] (four consecutive x per 8-byte store), so that the MFMA
// operands are 16-byte reads; wave w takes the chunk's 32 pixels 32w..32w+31
// (one k step) for all 64 channels and both 16-mode tiles, its Bx operand read
// from the zero-padded basis [32][xpad] in global memory.  The four waves'
// partial sums of a row are added in LDS in wave order.
constexpr int XFC = 128, XFS = XFC + 8;
__global__ __launch_bounds__(256) void sur_xfwd_kernel(const unsigned short *__restrict__ in,
                                                       const unsigned short *__restrict__ bxp, float *__restrict__ x1,
                                                       int X, int xpad, int rows) {
    __shared__ __attribute__((aligned(16))) unsigned short hT[C][XFS];
    __shared__ __attribute__((aligned(16))) float red[4][32 * C];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, j = lane & 15, g = lane >> 4;
    const int xq = threadIdx.x >> 3, cc = threadIdx.x & 7;  // this thread's pieces: x = 4 xq + u, channels 8 cc..
    const int nch = (X + XFC - 1) / XFC;
    const int nmine = rows > (int)blockIdx.x ? (rows - 1 - (int)blockIdx.x) / (int)gridDim.x + 1 : 0;
    const int total = nmine * nch;  // (row, chunk) steps of this workgroup
    auto load = [&](int s, uint4 (&v)[4]) {
        if (s >= total) return;
        const int row = blockIdx.x + (s / nch) * gridDim.x, x0 = (s % nch) * XFC;
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int x = x0 + 4 * xq + u;
            v[u] = make_uint4(0u, 0u, 0u, 0u);
            if (x < X) v[u] = *reinterpret_cast<const uint4 *>(in + ((long)row * X + x) * C + 8 * cc);
        }
    };
    // this lane's Bx operands of step s (8 pixels of the k step of wave w)
    auto loadb = [&](int s, bf16x8 (&b)[2]) {
        if (s >= total) return;
        const int xb = (s % nch) * XFC + 32 * wave + 8 * g;
#pragma unroll
        for (int nt = 0; nt < 2; nt++) b[nt] = *reinterpret_cast<const bf16x8 *>(bxp + (long)(16 * nt + j) * xpad + xb);
    };
    uint4 va[4], vb[4];
    bf16x8 bcur[2], bnxt[2];
    load(0, va);
    load(1, vb);
    loadb(0, bcur);
    f32x4 acc[4][2];
    for (int s = 0; s < total; s++) {
        const int x0 = (s % nch) * XFC;
        if (x0 == 0) {
#pragma unroll
            for (int ct = 0; ct < 4; ct++) acc[ct][0] = acc[ct][1] = f32x4{0.f, 0.f, 0.f, 0.f};
        }
        __syncthreads();  // the previous chunk's operand reads are done
#pragma unroll
        for (int t = 0; t < 8; t++) {
            uint2 w2;
            const unsigned sh = (t & 1) * 16;
            const unsigned e0 = ((t >> 1) == 0 ? va[0].x : (t >> 1) == 1 ? va[0].y : (t >> 1) == 2 ? va[0].z : va[0].w) >> sh;
            const unsigned e1 = ((t >> 1) == 0 ? va[1].x : (t >> 1) == 1 ? va[1].y : (t >> 1) == 2 ? va[1].z : va[1].w) >> sh;
            const unsigned e2 = ((t >> 1) == 0 ? va[2].x : (t >> 1) == 1 ? va[2].y : (t >> 1) == 2 ? va[2].z : va[2].w) >> sh;
            const unsigned e3 = ((t >> 1) == 0 ? va[3].x : (t >> 1) == 1 ? va[3].y : (t >> 1) == 2 ? va[3].z : va[3].w) >> sh;
            w2.x = (e0 & 0xFFFFu) | (e1 << 16);
            w2.y = (e2 & 0xFFFFu) | (e3 << 16);
            *reinterpret_cast<uint2 *>(&hT[8 * cc + t][4 * xq]) = w2;
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < 4; u++) va[u] = vb[u];
        load(s + 2, vb);  // two chunks ahead, in flight during the products
        loadb(s + 1, bnxt);
#pragma unroll
        for (int ct = 0; ct < 4; ct++) {
            const bf16x8 a = *reinterpret_cast<const bf16x8 *>(&hT[16 * ct + j][32 * wave + 8 * g]);
#pragma unroll
            for (int nt = 0; nt < 2; nt++) acc[ct][nt] = mfma_bf16(a, bcur[nt], acc[ct][nt]);
        }
        bcur[0] = bnxt[0];
        bcur[1] = bnxt[1];
        if (x0 + XFC >= X) {
            // row done: D[c][k] of lane (j, g) = channels 16 ct + 4 g + i, mode 16 nt + j
            const int row = blockIdx.x + (s / nch) * gridDim.x;
#pragma unroll
            for (int ct = 0; ct < 4; ct++)
#pragma unroll
                for (int nt = 0; nt < 2; nt++)
                    *reinterpret_cast<f32x4 *>(&red[wave][(16 * nt + j) * C + 16 * ct + 4 * g]) = acc[ct][nt];
            __syncthreads();
            float *dst = x1 + (long)row * 32 * C;
#pragma unroll
            for (int q = 0; q < 2; q++) {
                const int e = 4 * (threadIdx.x + 256 * q);
                f32x4 v = *reinterpret_cast<const f32x4 *>(&red[0][e]);
#pragma unroll
                for (int w = 1; w < 4; w++) v += *reinterpret_cast<const f32x4 *>(&red[w][e]);
                *reinterpret_cast<f32x4 *>(dst + e) = v;
            }
        }
    }
}

// The spectral branch's transform chain of the BF16 NHWC array `in`: X1 = Bx in,
// hbuf = By X1, gbuf = mix(hbuf) (transposed mixing when tr), X2 = By^T gbuf,
// then X2^T in BF16 (sp.d_x2t) for the conv kernel, which adds Bx^T X2.
int spectral(SurrogatePlan &sp, const unsigned short *in, float *hbuf, float *gbuf, bool tr, hipStream_t s) {
    const int X = sp.g.xd, Y = sp.g.yd, Z = sp.g.zd;
    {   // X1[z][y][kx][c] = sum_x Bx[kx][x] in[z][y][x][c]
        const int rows = Z * Y, nb = rows < 512 ? rows : 512;
        hipLaunchKernelGGL(sur_xfwd_kernel, dim3(nb), dim3(256), 0, s, in, sp.d_bx, sp.d_x1, X, sp.xpad, rows);
        MMS_HIP(hipGetLastError());
    }
    {
        GemmArgs a{};  // H[z][ky][(kx, c)] = sum_y By[ky][y] X1[z][y][(kx, c)]
        a.A = sp.d_by; a.sam = Y; a.sak = 1; a.sab = 0;
        a.B = sp.d_x1; a.sbk = 32L * C; a.sbn = 1; a.sbb = (long)Y * 32 * C;
        a.C = hbuf; a.scm = 32L * C; a.scn = 1; a.scb = (long)sp.nmode * C;
        a.M = 32; a.N = 32 * C; a.K = Y;
        if (gemm<unsigned short, float, float>(a, Z, s)) return 1;
    }
    if (tr) hipLaunchKernelGGL((sur_mix_kernel<true>), dim3(sp.nmode, Z), dim3(64), 0, s, hbuf, sp.d_wm, gbuf, sp.nmode);
    else hipLaunchKernelGGL((sur_mix_kernel<false>), dim3(sp.nmode, Z), dim3(64), 0, s, hbuf, sp.d_wm, gbuf, sp.nmode);
    MMS_HIP(hipGetLastError());
    {
        GemmArgs a{};  // X2[z][y][(kx, c)] = sum_ky By[ky][y] G[z][ky][(kx, c)]
        a.A = sp.d_by; a.sam = 1; a.sak = Y; a.sab = 0;
        a.B = gbuf; a.sbk = 32L * C; a.sbn = 1; a.sbb = (long)sp.nmode * C;
        a.C = sp.d_x2; a.scm = 32L * C; a.scn = 1; a.scb = (long)Y * 32 * C;
        a.M = Y; a.N = 32 * C; a.K = 32;
        if (gemm<unsigned short, float, float>(a, Z, s)) return 1;
    }
    hipLaunchKernelGGL(sur_x2t_kernel, dim3(Z * Y), dim3(256), 0, s, sp.d_x2, sp.d_x2t);
    MMS_HIP(hipGetLastError());
    return 0;
}

}  // namespace

// ------------------------------------------------------------ host plan --
// Weight vector layout of the API (FP32, in this order; surrogate_weight_count):
//   lift   [64 co][3 planes][3 ky][3 kx] + bias [64]
//   conv1  [64 co][64 ci][3][3] + bias [64]
//   conv2  [64 co][64 ci][3][3] + bias [64]
//   mode   [32 ky][32 kx][64 co][64 ci]
//   conv3  [64 co][64 ci][3][3] + bias [64]
//   proj   [64] + bias [1]
long surrogate_weight_count() {
    return (long)C * 27 + C + 3L * (C * C * 9 + C) + 32L * 32 * C * C + C + 1;
}
long surrogate_param_count() { return NPARAM; }

// API order -> device layout (the lift's k = tap * 3 + plane padded to 32,
// conv taps outermost); the other blocks keep their order
void surrogate_api_to_param(const float *api, float *prm) {
    std::fill(prm, prm + NPARAM, 0.0f);
    long o = 0;
    for (int co = 0; co < C; co++)
        for (int pl = 0; pl < 3; pl++)
            for (int tap = 0; tap < 9; tap++) prm[OFF_L + co * 32 + tap * 3 + pl] = api[o + (co * 3 + pl) * 9 + tap];
    o += C * 27;
    for (int co = 0; co < C; co++) prm[OFF_BL + co] = api[o + co];
    o += C;
    auto conv = [&](long off) {
        for (int co = 0; co < C; co++)
            for (int ci = 0; ci < C; ci++)
                for (int tap = 0; tap < 9; tap++) prm[off + ((long)tap * C + co) * C + ci] = api[o + (co * C + ci) * 9 + tap];
        o += CONVN;
        for (int co = 0; co < C; co++) prm[off + CONVN + co] = api[o + co];
        o += C;
    };
    conv(OFF_W1);
    conv(OFF_W2);
    std::copy(api + o, api + o + 1024L * C * C, prm + OFF_WM);
    o += 1024L * C * C;
    conv(OFF_W3);
    std::copy(api + o, api + o + C + 1, prm + OFF_WP);
}

void surrogate_param_to_api(const float *prm, float *api) {
    long o = 0;
    for (int co = 0; co < C; co++)
        for (int pl = 0; pl < 3; pl++)
            for (int tap = 0; tap < 9; tap++) api[o + (co * 3 + pl) * 9 + tap] = prm[OFF_L + co * 32 + tap * 3 + pl];
    o += C * 27;
    for (int co = 0; co < C; co++) api[o + co] = prm[OFF_BL + co];
    o += C;
    auto conv = [&](long off) {
        for (int co = 0; co < C; co++)
            for (int ci = 0; ci < C; ci++)
                for (int tap = 0; tap < 9; tap++) api[o + (co * C + ci) * 9 + tap] = prm[off + ((long)tap * C + co) * C + ci];
        o += CONVN;
        for (int co = 0; co < C; co++) api[o + co] = prm[off + CONVN + co];
        o += C;
    };
    conv(OFF_W1);
    conv(OFF_W2);
    std::copy(prm + OFF_WM, prm + OFF_WM + 1024L * C * C, api + o);
    o += 1024L * C * C;
    conv(OFF_W3);
    std::copy(prm + OFF_WP, prm + OFF_WP + C + 1, api + o);
}

int SurrogatePlan::init(const Grid &g_, int zf_, const float *weights, bool train_, hipStream_t s) {
    release();
    MMS_CHECK(zf_ >= 1 && g_.zd % zf_ == 0, "surrogate: layers are not a whole number of films");
    g = g_;
    zf = zf_;
    train = train_;
    const long npix = g.n();
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
    std::vector<float> prm((size_t)NPARAM);
    surrogate_api_to_param(wv.data(), prm.data());
    MMS_HIP(hipMalloc(&d_param, sizeof(float) * NPARAM));
    MMS_HIP(hipMemcpyAsync(d_param, prm.data(), sizeof(float) * NPARAM, hipMemcpyHostToDevice, s));
    MMS_HIP(hipMalloc(&d_lift, sizeof(unsigned short) * 32 * C));
    for (int l = 0; l < 3; l++) {
        MMS_HIP(hipMalloc(&d_w[l], sizeof(unsigned short) * CONVN));
        if (train) MMS_HIP(hipMalloc(&d_wt[l], sizeof(unsigned short) * CONVN));
    }
    MMS_HIP(hipMalloc(&d_wm, sizeof(unsigned short) * nmode * C * C));
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
    std::vector<unsigned short> bx0 = basis(g.xd), by = basis(g.yd), bxt(bx0.size());
    for (int k = 0; k < 32; k++)
        for (int x = 0; x < g.xd; x++) bxt[(size_t)x * 32 + k] = bx0[(size_t)k * g.xd + x];
    // Bx rows zero-padded to a multiple of the forward transform's 128-pixel chunk
    xpad = (g.xd + 127) / 128 * 128;
    std::vector<unsigned short> bx((size_t)32 * xpad, 0);
    for (int k = 0; k < 32; k++)
        for (int x = 0; x < g.xd; x++) bx[(size_t)k * xpad + x] = bx0[(size_t)k * g.xd + x];
    MMS_HIP(hipMalloc(&d_bx, bx.size() * 2));
    MMS_HIP(hipMalloc(&d_bxt, bxt.size() * 2));
    MMS_HIP(hipMemcpyAsync(d_bxt, bxt.data(), bxt.size() * 2, hipMemcpyHostToDevice, s));
    MMS_HIP(hipMalloc(&d_by, by.size() * 2));
    MMS_HIP(hipMemcpyAsync(d_bx, bx.data(), bx.size() * 2, hipMemcpyHostToDevice, s));
    MMS_HIP(hipMemcpyAsync(d_by, by.data(), by.size() * 2, hipMemcpyHostToDevice, s));
    const size_t act = sizeof(unsigned short) * npix * C;
    MMS_HIP(hipMalloc(&d_h0, act));
    MMS_HIP(hipMalloc(&d_h1, act));
    d_h2 = d_h0;  // inference: conv2's output overwrites the lift's
    MMS_HIP(hipMalloc(&d_x1, sizeof(float) * (size_t)g.zd * g.yd * 32 * C));
    MMS_HIP(hipMalloc(&d_hm, sizeof(float) * (size_t)g.zd * nmode * C));
    MMS_HIP(hipMalloc(&d_gm, sizeof(float) * (size_t)g.zd * nmode * C));
    MMS_HIP(hipMalloc(&d_x2, sizeof(float) * (size_t)g.zd * g.yd * 32 * C));
    MMS_HIP(hipMalloc(&d_x2t, sizeof(unsigned short) * (size_t)g.zd * g.yd * 32 * C));
    MMS_HIP(hipMalloc(&d_part, sizeof(double) * 2 * 1024));
    if (train) {
        MMS_HIP(hipMalloc(&d_h2, act));
        for (int l = 0; l < 4; l++) MMS_HIP(hipMalloc(&d_z[l], act));
        MMS_HIP(hipMalloc(&d_dza, act));
        MMS_HIP(hipMalloc(&d_dzb, act));
        MMS_HIP(hipMalloc(&d_dh, sizeof(float) * (size_t)g.zd * nmode * C));
        MMS_HIP(hipMalloc(&d_wpart, sizeof(float) * kPartBlocks * WGRAD_PART));
        MMS_HIP(hipMalloc(&d_hpart, sizeof(float) * (size_t)(g.zd / zf) * HB * HPART));
        MMS_HIP(hipMalloc(&d_hsum, sizeof(float) * HPART));
    }
    static bool attr = false;
    if (!attr) {
        for (const void *k : {reinterpret_cast<const void *>(&sur_conv8_kernel<0, false>),
                              reinterpret_cast<const void *>(&sur_conv8_kernel<0, true>),
                              reinterpret_cast<const void *>(&sur_conv8_kernel<1, false>),
                              reinterpret_cast<const void *>(&sur_conv8_kernel<2, false>),
                              reinterpret_cast<const void *>(&sur_conv8_kernel<2, true>)})
            MMS_HIP(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)CONV8_LDS));
        MMS_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&sur_wgrad_kernel),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)WGRAD_LDS));
        attr = true;
    }
    if (pack(s)) return 1;
    MMS_HIP(hipStreamSynchronize(s));
    return 0;
}

// BF16 device copies of the FP32 parameters
int SurrogatePlan::pack(hipStream_t s) {
    auto bf = [&](long off, unsigned short *dst, long n) -> int {
        hipLaunchKernelGGL(sur_to_bf_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, d_param + off, dst, n);
        MMS_HIP(hipGetLastError());
        return 0;
    };
    if (bf(OFF_L, d_lift, 32L * C) || bf(OFF_WM, d_wm, (long)nmode * C * C)) return 1;
    for (int l = 0; l < 3; l++) {
        if (bf(OFF_W[l], d_w[l], CONVN)) return 1;
        if (d_wt[l]) {
            hipLaunchKernelGGL(sur_flipT_kernel, dim3((unsigned)((CONVN + 255) / 256)), dim3(256), 0, s,
                               d_param + OFF_W[l], d_wt[l]);
            MMS_HIP(hipGetLastError());
        }
    }
    return 0;
}

void SurrogatePlan::release() {
    if (d_h2 == d_h0) d_h2 = nullptr;
    void *bufs[] = {d_param, d_lift, d_w[0], d_w[1], d_w[2], d_wt[0], d_wt[1], d_wt[2], d_wm, d_bx, d_by,
                    d_h0,    d_h1,   d_h2,   d_bxt,  d_x2t,  d_z[0], d_z[1],  d_z[2],  d_z[3],  d_dza, d_dzb,
                    d_x1,    d_hm,   d_gm,   d_x2,   d_dh,   d_wpart, d_hpart, d_hsum,  d_part};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    d_param = d_x1 = d_hm = d_gm = d_x2 = d_dh = d_wpart = d_hpart = d_hsum = nullptr;
    d_lift = d_wm = d_bx = d_by = d_bxt = d_x2t = d_h0 = d_h1 = d_h2 = d_dza = d_dzb = nullptr;
    for (int l = 0; l < 3; l++) d_w[l] = d_wt[l] = nullptr;
    for (int l = 0; l < 4; l++) d_z[l] = nullptr;
    d_part = nullptr;
}

int surrogate_potential(SurrogatePlan &sp, const float *rho, float *phi, float rscale, float oscale, hipStream_t s,
                        Profiler *prof) {
    const Grid &g = sp.g;
    MMS_CHECK(sp.d_h0 != nullptr, "surrogate not initialised");
    const int X = g.xd, Y = g.yd, Z = g.zd;
    const float *P = sp.d_param;
    KTimer kt(prof, K_SURROGATE);
    hipLaunchKernelGGL(sur_lift_kernel, dim3((X + TS - 1) / TS, (Y + TS - 1) / TS, Z), dim3(256), 0, s, rho, sp.d_lift,
                       P + OFF_BL, sp.d_h0, sp.d_z[0], X, Y, sp.zf, rscale);
    MMS_HIP(hipGetLastError());
    ConvArgs a{};
    a.xd = X; a.yd = Y;
    a.in = sp.d_h0; a.w = sp.d_w[0]; a.bias = P + OFF_W1 + CONVN; a.out = sp.d_h1; a.zsave = sp.d_z[1];
    if (launch_conv(0, sp, a, s)) return 1;
    // spectral branch of h1
    if (spectral(sp, sp.d_h1, sp.d_hm, sp.d_gm, false, s)) return 1;
    a.in = sp.d_h1; a.w = sp.d_w[1]; a.bias = P + OFF_W2 + CONVN; a.x2t = sp.d_x2t; a.bxt = sp.d_bxt; a.out = sp.d_h2;
    a.zsave = sp.d_z[2];
    if (launch_conv(0, sp, a, s)) return 1;
    a = ConvArgs{};
    a.xd = X; a.yd = Y;
    a.in = sp.d_h2; a.w = sp.d_w[2]; a.bias = P + OFF_W3 + CONVN; a.zsave = sp.d_z[3];
    a.wp = P + OFF_WP; a.bp = P + OFF_BP; a.phi = phi; a.oscale = oscale;
    return launch_conv(1, sp, a, s);
}

int surrogate_losses(SurrogatePlan &sp, const float *phi_s, const float *phi_t, const float *inv_norm, float gscale,
                     float oscale, bool grad, hipStream_t s) {
    const int nf = sp.g.zd / sp.zf;
    hipLaunchKernelGGL(sur_head_kernel, dim3(HB, nf), dim3(256), 0, s, phi_s, phi_t, inv_norm, gscale, oscale,
                       sp.d_z[3], sp.d_param + OFF_WP, grad ? sp.d_dza : (unsigned short *)nullptr, sp.g.n() / nf,
                       sp.d_hpart);
    MMS_HIP(hipGetLastError());
    return 0;
}

int surrogate_backward(SurrogatePlan &sp, const float *rho, const float *phi_s, const float *phi_t,
                       const float *inv_norm, float gscale, float rscale, float oscale, float *grad, hipStream_t s) {
    MMS_CHECK(sp.train, "surrogate plan without training buffers");
    const Grid &g = sp.g;
    const int nf = g.zd / sp.zf;
    // loss and projection: dz3 -> d_dza, dwp / dbp / loss partials -> d_hsum
    if (surrogate_losses(sp, phi_s, phi_t, inv_norm, gscale, oscale, true, s)) return 1;
    hipLaunchKernelGGL(sur_sum_kernel, dim3(sum_blocks(HPART)), dim3(256), 0, s, sp.d_hpart, (long)HPART, nf * HB, (long)HPART,
                       sp.d_hsum);
    MMS_HIP(hipGetLastError());
    MMS_HIP(hipMemcpyAsync(grad + OFF_WP, sp.d_hsum, sizeof(float) * (C + 1), hipMemcpyDeviceToDevice, s));
    // conv3: dW3 from (dz3, h2); dz2 = convT(dz3) GELU'(z2)
    if (launch_wgrad(sp, sp.d_dza, sp.d_h2, grad + OFF_W3, s)) return 1;
    ConvArgs a{};
    a.xd = g.xd; a.yd = g.yd;
    a.in = sp.d_dza; a.w = sp.d_wt[2]; a.zin = sp.d_z[2]; a.out = sp.d_dzb;
    if (launch_conv(2, sp, a, s)) return 1;
    // conv2: dW2 from (dz2, h1); the spectral branch: dWm and its data gradient
    if (launch_wgrad(sp, sp.d_dzb, sp.d_h1, grad + OFF_W2, s)) return 1;
    if (spectral(sp, sp.d_dzb, sp.d_gm, sp.d_dh, true, s)) return 1;
    hipLaunchKernelGGL(sur_mix_wgrad_kernel, dim3(sp.nmode), dim3(256), 0, s, sp.d_gm, sp.d_hm, grad + OFF_WM,
                       sp.nmode, g.zd);
    MMS_HIP(hipGetLastError());
    // dz1 = (convT(dz2) + spectral^T(dz2)) GELU'(z1)
    a.in = sp.d_dzb; a.w = sp.d_wt[1]; a.x2t = sp.d_x2t; a.bxt = sp.d_bxt; a.zin = sp.d_z[1]; a.out = sp.d_dza;
    if (launch_conv(2, sp, a, s)) return 1;
    // conv1: dW1 from (dz1, h0); dz0 = convT(dz1) GELU'(z0)
    if (launch_wgrad(sp, sp.d_dza, sp.d_h0, grad + OFF_W1, s)) return 1;
    a.in = sp.d_dza; a.w = sp.d_wt[0]; a.x2t = nullptr; a.bxt = nullptr; a.zin = sp.d_z[0]; a.out = sp.d_dzb;
    if (launch_conv(2, sp, a, s)) return 1;
    // lift
    const int ntile = ((g.xd + TS - 1) / TS) * ((g.yd + TS - 1) / TS) * g.zd;
    const int nb = ntile < SurrogatePlan::kPartBlocks ? ntile : SurrogatePlan::kPartBlocks;
    hipLaunchKernelGGL(sur_lift_wgrad_kernel, dim3(nb), dim3(256), 0, s, rho, sp.d_dzb, sp.d_wpart, g.xd, g.yd, g.zd,
                       sp.zf, rscale);
    MMS_HIP(hipGetLastError());
    hipLaunchKernelGGL(sur_sum_kernel, dim3(sum_blocks(LGRAD_PART)), dim3(256), 0, s, sp.d_wpart,
                       LGRAD_PART, nb, LGRAD_PART, grad + OFF_L);
    MMS_HIP(hipGetLastError());
    return 0;
}

int surrogate_adam(SurrogatePlan &sp, float *m, float *v, const float *grad, long step, float lr, hipStream_t s) {
    const double b1 = 0.9, b2 = 0.999;
    const float c1 = (float)(1.0 / (1.0 - pow(b1, (double)step))), c2 = (float)(1.0 / (1.0 - pow(b2, (double)step)));
    hipLaunchKernelGGL(sur_adam_kernel, dim3((unsigned)((NPARAM + 255) / 256)), dim3(256), 0, s, sp.d_param, m, v,
                       grad, NPARAM, lr, (float)b1, (float)b2, c1, c2, 1e-8f);
    MMS_HIP(hipGetLastError());
    return sp.pack(s);
}

int surrogate_sqnorm(SurrogatePlan &sp, const float *x, long n, double *out, hipStream_t s) {
    // the residual kernel of x against itself: its second partials are sum x^2
    hipLaunchKernelGGL(sur_residual_kernel, dim3(1024), dim3(256), 0, s, x, x, n, sp.d_part);
    MMS_HIP(hipGetLastError());
    std::vector<double> h(2048);
    MMS_HIP(hipMemcpyAsync(h.data(), sp.d_part, sizeof(double) * 2048, hipMemcpyDeviceToHost, s));
    MMS_HIP(hipStreamSynchronize(s));
    double b2 = 0;
    for (int i = 0; i < 1024; i++) b2 += h[2 * i + 1];
    *out = b2;
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
