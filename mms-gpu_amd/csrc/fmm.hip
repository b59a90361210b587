// fmm.hip -- regular-grid fast multipole potential on gfx950.
//
// Replaces the reference tree-code (fmm_calc.cu:20-408, fmm_kernel.cu:21-155),
// which recomputes multipoles from raw charges at every level and evaluates
// them directly at every target of the interaction list.  Here the same
// quadtree (per charge layer, interaction list = children of the parent's
// neighbours minus own neighbours, Box.cu:121-145) carries a full FMM:
//
//   P2M   leaf multipoles from charges              M = P2M[KP x w^2] rho_box
//   M2M   children -> parent                       M_p = sum_c M2M_c M_c
//   M2L   interaction list -> local                L_t += T(level, offset, dz) M_s   (MFMA)
//   L2L   parent local -> child                    L_c += L2L_c L_p                  (MFMA, fused)
//   L2P   local -> cells of the leaf               phi = L2P[w^2 x KP] L
//   P2P   3x3 neighbour leaf boxes, all layers     phi += sum rho/r                  (fused with L2P)
//
// On a regular grid every operator is translation invariant, so each is a
// small constant matrix built once in FP64 on the host (fmm_operators.cpp)
// and the device work is gathers plus dense contractions.  Box coefficient
// vectors are stored with KP slots (K = (P+1)(P+2)/2 rounded up to a multiple
// of 32; padding slots are zero in every operator), layout [z][by][bx][KP].
//
// M2L on the matrix cores: a workgroup owns a dense 16 x 16 block of target
// boxes of one target layer; wavefront w handles child parity (w&1, w>>1),
// i.e. 8 x 8 targets = four 16-column MFMA tiles, so every lane of a wave
// walks the same 27 interaction offsets.  The source multipoles of a 22 x 22
// window (3-box halo) of one source layer sit in LDS; for every (offset,
// source layer) the wave issues C[KP x 64] += T[KP x KP] * M_window[KP x 64].
// Operator fragments are pre-arranged on the host in lane order
// ([row tile][lane][slot group]) so each lane loads its A fragment with
// 16-byte vector loads.  Precision modes:
//   FP32    v_mfma_f32_16x16x4_f32            (FP32 operands and accumulation)
//   BF16    v_mfma_f32_16x16x32_bf16          (BF16 operands, FP32 accumulation)
//   BF16X3  hi/lo BF16 split of both operands, C += Th Mh + Th Ml + Tl Mh
//   FP16    v_mfma_f32_16x16x32_f16           (FP16 operands, FP32 accumulation)
//           with exact power-of-two scaling into FP16 range: operator row j
//           and column k carry factors a_j, b_k (host equilibration), and each
//           workgroup scales its staged window M_k / b_k by s = 2^e so that its
//           largest entry lies in [2^14, 2^15); the accumulators are
//           multiplied by 1 / (a_j s) before they are stored
// L2L is always FP32.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <type_traits>

#include "mms_internal.hpp"

namespace mms {
namespace {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ f32x4 mfma32(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f32x4 mfma32(f16x8 a, f16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}

// coefficient slots per box: 16 when K <= 16 (P <= 4), else K rounded up to 32
template <int K> struct KPad { static constexpr int KP = K <= 16 ? 16 : ((K + 31) / 32) * 32; };
template <int P> struct NC { static constexpr int K = (P + 1) * (P + 2) / 2; };

__device__ inline unsigned short f2bf(float x) {  // round to nearest even
    unsigned u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ inline float bf2f(unsigned short h) { return __uint_as_float((unsigned)h << 16); }

// ------------------------------------------------------------------ M2M --
// One wave per 64 parent boxes of one layer (4 MFMA column tiles of 16):
// C[KP x 64] = sum over the 4 child positions of M2M_c[KP x KP] * M_child[KP x 64],
// FP32 MFMA, operator fragments in lane order (see M2L below).
template <int KP>
__global__ __launch_bounds__(256) void m2m_mfma_kernel(const float *__restrict__ Mc, float *__restrict__ Mp,
                                                       const float *__restrict__ m2mfrag, int nbp, FilmSet fs) {
    constexpr int G = KP / 4, RT = KP / 16;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int j16 = lane & 15, g = lane >> 4;
    const int by = blockIdx.y, z = blockIdx.z % fs.zd;
    Mc += (long)(blockIdx.z / fs.zd) * fs.stride;
    Mp += (long)(blockIdx.z / fs.zd) * fs.stride;
    const int bx0 = (blockIdx.x * 4 + wave) * 64;
    if (bx0 >= nbp) return;
    const int nbc = 2 * nbp;
    f32x4 acc[RT][4];
#pragma unroll
    for (int r = 0; r < RT; r++)
#pragma unroll
        for (int c = 0; c < 4; c++) acc[r][c] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll 1
    for (int ci = 0; ci < 4; ci++) {
        const int cyy = 2 * by + (ci >> 1);
#pragma unroll
        for (int u4 = 0; u4 < G / 4; u4++) {
            f32x4 a[RT], b[4];
#pragma unroll
            for (int r = 0; r < RT; r++)
                a[r] = *reinterpret_cast<const f32x4 *>(m2mfrag + (((long)ci * RT + r) * 64 + lane) * G + 4 * u4);
#pragma unroll
            for (int c = 0; c < 4; c++) {
                const int bx = bx0 + 16 * c + j16;
                b[c] = f32x4{0.f, 0.f, 0.f, 0.f};
                if (bx < nbp)
                    b[c] = *reinterpret_cast<const f32x4 *>(Mc + (((long)z * nbc + cyy) * nbc + 2 * bx + (ci & 1)) * KP +
                                                            g * G + 4 * u4);
            }
#pragma unroll
            for (int u = 0; u < 4; u++)
#pragma unroll
                for (int r = 0; r < RT; r++)
#pragma unroll
                    for (int c = 0; c < 4; c++)
                        acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[r][u], b[c][u], acc[r][c], 0, 0, 0);
        }
    }
#pragma unroll
    for (int c = 0; c < 4; c++) {
        const int bx = bx0 + 16 * c + j16;
        if (bx >= nbp) continue;
        f32x4 *dst = reinterpret_cast<f32x4 *>(Mp + (((long)z * nbp + by) * nbp + bx) * KP);
#pragma unroll
        for (int r = 0; r < RT; r++) dst[4 * r + g] = acc[r][c];
    }
}

// ------------------------------------------------ upward / downward subtrees --
// Float offsets of every level in mult / loc / part (per film).
struct LevelOffsets { long off[16]; };

// Upward pass of one subtree: the workgroup owns one box of level R of one
// layer and the S x S leaves under it (S = 2^D, D = Lleaf - R <= 4).  P2M of
// the leaves as FP32 MFMA [KP x W^2] . [W^2 x leaves] with the charges staged
// in LDS, then D levels of M2M on the LDS copy; every level's multipoles are
// written to HBM.  Wave w takes MFMA column tiles w, w + 4, ... of 16 boxes.
template <int K, int KP, int W>
__global__ __launch_bounds__(256) void up_subtree_kernel(const float *__restrict__ rho, float *__restrict__ Mall,
                                                         const float *__restrict__ p2mf,
                                                         const float *__restrict__ m2mfrag, LevelOffsets lo, int R,
                                                         int D, int xd, int yd, long bufA_floats, FilmSet fs) {
    constexpr int G = KP / 4, RT = KP / 16, GP = W * W / 4, SB = KP + 4;
    extern __shared__ __attribute__((aligned(16))) float lds[];
    // S, CS, nR and sp below are powers of two: index splits are shifts and masks
    static_assert(W == 2 || W == 4 || W == 8, "leaf width must be 2, 4 or 8");
    constexpr int LW = W == 8 ? 3 : (W == 4 ? 2 : 1);
    const int S = 1 << D, CS = S * W, LCS = D + LW, nR = 1 << R, Lleaf = R + D;
    const int rbx = blockIdx.x & (nR - 1), rby = blockIdx.x >> R;
    const int f = blockIdx.y / fs.zd, z = blockIdx.y % fs.zd;
    if (fs.desc) {
        const FilmDesc d = fs.desc[f];
        xd = d.xd; yd = d.yd; rho += d.off;
    }
    float *M = Mall + (long)f * fs.stride;
    const long plane = (long)xd * yd;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, j16 = lane & 15, g = lane >> 4;
    const int cx0 = rbx * CS, cy0 = rby * CS;
    float *q = lds;  // CS x CS charges, later the leaf multipoles [leaf][SB]
#pragma unroll 1
    for (int e0 = 0; e0 < CS * CS; e0 += 256 * 8) {  // 8 loads per thread in flight
        float vb[8];
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const int e = e0 + threadIdx.x + 256 * k;
            const int x = cx0 + (e & (CS - 1)), y = cy0 + (e >> LCS);
            vb[k] = (e < CS * CS && x < xd && y < yd) ? rho[(long)z * plane + (long)y * xd + x] : 0.0f;
        }
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const int e = e0 + threadIdx.x + 256 * k;
            if (e < CS * CS) q[e] = vb[k];
        }
    }
    float a[RT][GP];
#pragma unroll
    for (int r = 0; r < RT; r++)
#pragma unroll
        for (int u = 0; u < GP; u++) a[r][u] = p2mf[((long)r * 64 + lane) * GP + u];
    __syncthreads();
    const int nleaf = S * S, ntile = (nleaf + 15) / 16;
    f32x4 acc[RT][4];
#pragma unroll
    for (int r = 0; r < RT; r++)
#pragma unroll
        for (int c = 0; c < 4; c++) acc[r][c] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ct = 0; ct < 4; ct++) {
        const int c = wave + 4 * ct;
        if (c < ntile) {
            const int n = 16 * c + j16, nx = n & (S - 1), ny = n >> D;
            const bool ok = n < nleaf;
#pragma unroll
            for (int u = 0; u < GP; u++) {
                const int cell = GP * g + u;
                const float b = ok ? q[(ny * W + cell / W) * CS + nx * W + cell % W] : 0.0f;
#pragma unroll
                for (int r = 0; r < RT; r++)
                    acc[r][ct] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[r][u], b, acc[r][ct], 0, 0, 0);
            }
        }
    }
    __syncthreads();  // charges consumed: the leaf multipoles go over them
    {
        const int nbl = 1 << Lleaf;
#pragma unroll
        for (int ct = 0; ct < 4; ct++) {
            const int c = wave + 4 * ct, n = 16 * c + j16;
            if (c < ntile && n < nleaf) {
                const int gx = rbx * S + (n & (S - 1)), gy = rby * S + (n >> D);
                float *dst = M + lo.off[Lleaf] + (((long)z * nbl + gy) * nbl + gx) * KP;
#pragma unroll
                for (int r = 0; r < RT; r++) {
                    *reinterpret_cast<f32x4 *>(dst + 16 * r + 4 * g) = acc[r][ct];
                    *reinterpret_cast<f32x4 *>(q + n * SB + 16 * r + 4 * g) = acc[r][ct];
                }
            }
        }
    }
    __syncthreads();
    float *ch = q, *pa = lds + bufA_floats;
    int s = S;
    for (int l = Lleaf - 1; l >= R; l--) {
        const int sp = s >> 1, lsp = l - R, np = sp * sp, ntp = (np + 15) / 16, nbp = 1 << l;
        for (int c = wave; c < ntp; c += 4) {
            const int p = 16 * c + j16, px = p & (sp - 1), py = p >> lsp;
            const bool ok = p < np;
            f32x4 acc2[RT];
#pragma unroll
            for (int r = 0; r < RT; r++) acc2[r] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll 1
            for (int ci = 0; ci < 4; ci++) {
                const int chn = (2 * py + (ci >> 1)) * s + 2 * px + (ci & 1);
#pragma unroll
                for (int u4 = 0; u4 < G / 4; u4++) {
                    f32x4 av[RT];
#pragma unroll
                    for (int r = 0; r < RT; r++)
                        av[r] = *reinterpret_cast<const f32x4 *>(m2mfrag + (((long)ci * RT + r) * 64 + lane) * G + 4 * u4);
                    const f32x4 b = ok ? *reinterpret_cast<const f32x4 *>(ch + chn * SB + g * G + 4 * u4)
                                       : f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
                    for (int u = 0; u < 4; u++)
#pragma unroll
                        for (int r = 0; r < RT; r++)
                            acc2[r] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[r][u], b[u], acc2[r], 0, 0, 0);
                }
            }
            if (ok) {
                const int gx = rbx * sp + px, gy = rby * sp + py;
                float *dst = M + lo.off[l] + (((long)z * nbp + gy) * nbp + gx) * KP;
#pragma unroll
                for (int r = 0; r < RT; r++) {
                    *reinterpret_cast<f32x4 *>(dst + 16 * r + 4 * g) = acc2[r];
                    *reinterpret_cast<f32x4 *>(pa + p * SB + 16 * r + 4 * g) = acc2[r];
                }
            }
        }
        __syncthreads();
        float *t = ch; ch = pa; pa = t;
        s = sp;
    }
}

// Downward pass of one subtree: the workgroup owns one box of level R of one
// layer.  Each level's local expansions are the M2L partial sums over source
// layers plus L2L from the parent (FP32 MFMA, wave w = child position w);
// the root takes its parent from HBM (level R - 1), or, with `chain`, forms
// the locals of its ancestors at levels 2 .. R - 1 itself (one column each,
// the same sums in the same order as down_coarse_kernel); the levels between
// stay in LDS.  W = 0: the leaf locals are written to HBM.  W = 4, 8: the leaf
// locals stay in LDS too and the far field is added to the potential
// (phi[cells of leaf] += scale * L2P[W*W x KP] * L, FP32 MFMA), so phi must
// already hold the near field.
template <int KP, int W>
__global__ __launch_bounds__(256) void down_subtree_kernel(const float *__restrict__ part, long part_stride,
                                                           float *__restrict__ Lall, const float *__restrict__ l2lfrag,
                                                           LevelOffsets lo, int R, int D, int zd, long film_stride,
                                                           float *__restrict__ phi, const float *__restrict__ l2pfrag,
                                                           int xd, int yd, float scale, int chain, FilmSet fs) {
    constexpr int G = KP / 4, RT = KP / 16, SB = KP + 4;
    constexpr bool FUSE = W > 0;
    extern __shared__ __attribute__((aligned(16))) float lds[];
    const int nR = 1 << R, Lleaf = R + D;
    const int rbx = blockIdx.x & (nR - 1), rby = blockIdx.x >> R;
    const int f = blockIdx.y / zd, z = blockIdx.y % zd;
    const float *Pf = part + (long)f * film_stride;
    float *L = Lall + (long)f * film_stride;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, j16 = lane & 15, g = lane >> 4;
    // two level buffers; with L2P fused the leaf level (4^D boxes) is kept too
    const int nbuf = (FUSE && D > 3) ? (1 << (2 * D)) : 64;
    float *pa = lds, *ch = lds + nbuf * SB;
    // root box (level R): wave 0, column 0
    if (wave == 0) {
        // acc = L2L[ci] . parent, the parent vector in column 0 of B
        auto l2l_col = [&](f32x4 *acc, int ci, const float *par) {
#pragma unroll
            for (int r = 0; r < RT; r++) acc[r] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int u4 = 0; u4 < G / 4; u4++) {
                f32x4 av[RT];
#pragma unroll
                for (int r = 0; r < RT; r++)
                    av[r] = *reinterpret_cast<const f32x4 *>(l2lfrag + (((long)ci * RT + r) * 64 + lane) * G + 4 * u4);
                const f32x4 b = (j16 == 0) ? *reinterpret_cast<const f32x4 *>(par + g * G + 4 * u4)
                                           : f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int u = 0; u < 4; u++)
#pragma unroll
                    for (int r = 0; r < RT; r++) acc[r] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[r][u], b[u], acc[r], 0, 0, 0);
            }
        };
        f32x4 acc[RT];
#pragma unroll
        for (int r = 0; r < RT; r++) acc[r] = f32x4{0.f, 0.f, 0.f, 0.f};
        const int ci = (rbx & 1) + 2 * (rby & 1);
        if (R > 2 && chain) {
            // ancestors at levels 2 .. R - 1; each local goes through the
            // first KP floats of ch (unused until the level loop) to become
            // the next level's B column, inside this wave
            float *cv = ch;
            for (int l = 2; l < R; l++) {
                const int ax = rbx >> (R - l), ay = rby >> (R - l), nbl = 1 << l;
                f32x4 a2[RT];
                if (l > 2) {
                    __builtin_amdgcn_wave_barrier();
                    l2l_col(a2, (ax & 1) + 2 * (ay & 1), cv);
                } else {
#pragma unroll
                    for (int r = 0; r < RT; r++) a2[r] = f32x4{0.f, 0.f, 0.f, 0.f};
                }
                __builtin_amdgcn_wave_barrier();
                if (j16 == 0) {
                    const long e = lo.off[l] + (((long)z * nbl + ay) * nbl + ax) * KP;
#pragma unroll
                    for (int r = 0; r < RT; r++) {
                        f32x4 v = a2[r];
#pragma unroll 4
                        for (int zs = 0; zs < zd; zs++)
                            v += *reinterpret_cast<const f32x4 *>(Pf + zs * part_stride + e + 16 * r + 4 * g);
                        *reinterpret_cast<f32x4 *>(cv + 16 * r + 4 * g) = v;
                    }
                }
            }
            __builtin_amdgcn_wave_barrier();
            l2l_col(acc, ci, cv);
        } else if (R > 2) {
            const int half = nR >> 1;
            l2l_col(acc, ci, L + lo.off[R - 1] + (((long)z * half + (rby >> 1)) * half + (rbx >> 1)) * KP);
        }
        if (j16 == 0) {
            const long e = lo.off[R] + (((long)z * nR + rby) * nR + rbx) * KP;
#pragma unroll
            for (int r = 0; r < RT; r++) {
                f32x4 v = acc[r];
#pragma unroll 4
                for (int s = 0; s < zd; s++) v += *reinterpret_cast<const f32x4 *>(Pf + s * part_stride + e + 16 * r + 4 * g);
                if (D == 0 && !FUSE) *reinterpret_cast<f32x4 *>(L + e + 16 * r + 4 * g) = v;
                else *reinterpret_cast<f32x4 *>(pa + 16 * r + 4 * g) = v;
            }
        }
    }
    __syncthreads();
    const int cxp = wave & 1, cyp = wave >> 1;
    int sp = 1;
    for (int l = R + 1; l <= Lleaf; l++) {
        const int s = 2 * sp, lsp = l - 1 - R, np = sp * sp, ntp = (np + 15) / 16, nb = 1 << l;
        for (int c = 0; c < ntp; c++) {
            const int p = 16 * c + j16, px = p & (sp - 1), py = p >> lsp;
            const bool ok = p < np;
            f32x4 acc[RT];
#pragma unroll
            for (int r = 0; r < RT; r++) acc[r] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int u4 = 0; u4 < G / 4; u4++) {
                f32x4 av[RT];
#pragma unroll
                for (int r = 0; r < RT; r++)
                    av[r] = *reinterpret_cast<const f32x4 *>(l2lfrag + (((long)wave * RT + r) * 64 + lane) * G + 4 * u4);
                const f32x4 b = ok ? *reinterpret_cast<const f32x4 *>(pa + p * SB + g * G + 4 * u4)
                                   : f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int u = 0; u < 4; u++)
#pragma unroll
                    for (int r = 0; r < RT; r++) acc[r] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[r][u], b[u], acc[r], 0, 0, 0);
            }
            if (ok) {
                const int cx = 2 * px + cxp, cy = 2 * py + cyp;
                const long e = lo.off[l] + (((long)z * nb + rby * s + cy) * nb + rbx * s + cx) * KP;
#pragma unroll
                for (int r = 0; r < RT; r++) {
                    f32x4 v = acc[r];
#pragma unroll 4
                    for (int zs = 0; zs < zd; zs++)
                        v += *reinterpret_cast<const f32x4 *>(Pf + zs * part_stride + e + 16 * r + 4 * g);
                    if (l == Lleaf && !FUSE) *reinterpret_cast<f32x4 *>(L + e + 16 * r + 4 * g) = v;
                    else *reinterpret_cast<f32x4 *>(ch + (cy * s + cx) * SB + 16 * r + 4 * g) = v;
                }
            }
        }
        __syncthreads();
        float *t = pa; pa = ch; ch = t;
        sp = s;
    }
    if constexpr (FUSE) {
        // pa = the sp x sp leaf locals of this subtree; wave w takes column
        // tiles of 16 leaves w, w + 4, ...
        constexpr int RTW = W * W / 16;
        float *ph = phi;
        if (fs.desc) {
            const FilmDesc d = fs.desc[f];
            xd = d.xd; yd = d.yd; ph += d.off;
        }
        const long plane = (long)xd * yd;
        const int nl = sp * sp;
        for (int c = wave; c < (nl + 15) / 16; c += 4) {
            const int p = 16 * c + j16, px = p & (sp - 1), py = p >> D;
            const bool ok = p < nl;
            f32x4 acc[RTW];
#pragma unroll
            for (int r = 0; r < RTW; r++) acc[r] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int u4 = 0; u4 < G / 4; u4++) {
                const f32x4 b = ok ? *reinterpret_cast<const f32x4 *>(pa + p * SB + g * G + 4 * u4)
                                   : f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int r = 0; r < RTW; r++) {
                    const f32x4 a = *reinterpret_cast<const f32x4 *>(l2pfrag + ((long)r * 64 + lane) * G + 4 * u4);
#pragma unroll
                    for (int u = 0; u < 4; u++) acc[r] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[u], b[u], acc[r], 0, 0, 0);
                }
            }
            const int bx = rbx * sp + px, by = rby * sp + py;
            // D row 4g + q of tile r = cell 16 r + 4g + q of leaf j16: 4 consecutive x cells
#pragma unroll
            for (int r = 0; r < RTW; r++) {
                const int cc = 16 * r + 4 * g;
                const int y = by * W + cc / W, xb = bx * W + cc % W;
                if (!ok || y >= yd) continue;
                float *dst = ph + (long)z * plane + (long)y * xd + xb;
                if (xb + 3 < xd && (((long)z * plane + (long)y * xd + xb) & 3) == 0) {
                    const f32x4 v = *reinterpret_cast<const f32x4 *>(dst);
                    *reinterpret_cast<f32x4 *>(dst) = v + acc[r] * scale;
                } else {
#pragma unroll
                    for (int q = 0; q < 4; q++)
                        if (xb + q < xd) dst[q] += acc[r][q] * scale;
                }
            }
        }
    }
}

// Coarse levels of the upward pass in one workgroup per (film, layer):
// levels Ltop - 1 .. 2 by M2M, the children of the first level read from HBM
// (level Ltop <= 5), the coarser ones from LDS; every level is written.
template <int KP>
__global__ __launch_bounds__(256) void up_coarse_kernel(float *__restrict__ Mall, const float *__restrict__ m2mfrag,
                                                        LevelOffsets lo, int Ltop, int zd, long film_stride) {
    constexpr int G = KP / 4, RT = KP / 16, SB = KP + 4;
    extern __shared__ __attribute__((aligned(16))) float lds[];
    const int f = blockIdx.x / zd, z = blockIdx.x % zd;
    float *M = Mall + (long)f * film_stride;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, j16 = lane & 15, g = lane >> 4;
    float *const b0 = lds, *const b1 = lds + 256 * SB;  // level l goes to b1 if l is odd, else b0
    // the M2M operator fragments of the four child positions, held in
    // registers for every level (at most 32 slots: 64 VGPRs); larger blocks
    // read them inside the product loop
    constexpr bool HOLD = KP <= 32;
    f32x4 avh[HOLD ? 4 : 1][HOLD ? G / 4 : 1][RT];
    if constexpr (HOLD) {
#pragma unroll
        for (int ci = 0; ci < 4; ci++)
#pragma unroll
            for (int u4 = 0; u4 < G / 4; u4++)
#pragma unroll
                for (int r = 0; r < RT; r++)
                    avh[ci][u4][r] = *reinterpret_cast<const f32x4 *>(m2mfrag + (((long)ci * RT + r) * 64 + lane) * G + 4 * u4);
    }
    int s = 1 << Ltop;
    for (int l = Ltop - 1; l >= 2; l--) {
        const int sp = s >> 1, np = sp * sp, ntp = (np + 15) / 16;
        const float *ch = ((l + 1) & 1) ? b1 : b0;
        float *pa = (l & 1) ? b1 : b0;
        const bool from_hbm = (l == Ltop - 1);
        for (int c = wave; c < ntp; c += 4) {
            const int p = 16 * c + j16, px = p % sp, py = p / sp;
            const bool ok = p < np;
            f32x4 acc[RT];
#pragma unroll
            for (int r = 0; r < RT; r++) acc[r] = f32x4{0.f, 0.f, 0.f, 0.f};
            // the four children's vectors, every load issued before the products
            f32x4 bv[4][G / 4];
#pragma unroll
            for (int ci = 0; ci < 4; ci++) {
                const int cx = 2 * px + (ci & 1), cy = 2 * py + (ci >> 1);
                const float *bsrc = from_hbm ? M + lo.off[l + 1] + (((long)z * s + cy) * s + cx) * KP
                                             : ch + (cy * s + cx) * SB;
#pragma unroll
                for (int u4 = 0; u4 < G / 4; u4++)
                    bv[ci][u4] = ok ? *reinterpret_cast<const f32x4 *>(bsrc + g * G + 4 * u4) : f32x4{0.f, 0.f, 0.f, 0.f};
            }
#pragma unroll
            for (int ci = 0; ci < 4; ci++) {
#pragma unroll
                for (int u4 = 0; u4 < G / 4; u4++) {
                    f32x4 av[RT];
#pragma unroll
                    for (int r = 0; r < RT; r++)
                        av[r] = HOLD ? avh[HOLD ? ci : 0][HOLD ? u4 : 0][r]
                                     : *reinterpret_cast<const f32x4 *>(m2mfrag + (((long)ci * RT + r) * 64 + lane) * G + 4 * u4);
                    const f32x4 b = bv[ci][u4];
#pragma unroll
                    for (int u = 0; u < 4; u++)
#pragma unroll
                        for (int r = 0; r < RT; r++) acc[r] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[r][u], b[u], acc[r], 0, 0, 0);
                }
            }
            if (ok) {
                float *dst = M + lo.off[l] + (((long)z * sp + py) * sp + px) * KP;
#pragma unroll
                for (int r = 0; r < RT; r++) {
                    *reinterpret_cast<f32x4 *>(dst + 16 * r + 4 * g) = acc[r];
                    *reinterpret_cast<f32x4 *>(pa + p * SB + 16 * r + 4 * g) = acc[r];
                }
            }
        }
        __syncthreads();
        s = sp;
    }
}

// Coarse levels of the downward pass in one workgroup per (film, layer):
// levels 2 .. Lbot (Lbot <= 4): the M2L partial sums plus L2L from the
// parent (LDS), every level written to HBM.
template <int KP>
__global__ __launch_bounds__(256) void down_coarse_kernel(const float *__restrict__ part, long part_stride,
                                                          float *__restrict__ Lall, const float *__restrict__ l2lfrag,
                                                          LevelOffsets lo, int Lbot, int zd, long film_stride) {
    constexpr int G = KP / 4, RT = KP / 16, SB = KP + 4;
    extern __shared__ __attribute__((aligned(16))) float lds[];
    const int f = blockIdx.x / zd, z = blockIdx.x % zd;
    const float *Pf = part + (long)f * film_stride;
    float *L = Lall + (long)f * film_stride;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, j16 = lane & 15, g = lane >> 4;
    float *const b0 = lds, *const b1 = lds + 256 * SB;  // level l goes to b1 if l is odd, else b0
    // level 2: the partial sums only (16 boxes, one per lane of wave 0's column tile)
    if (wave == 0) {
        const long e = lo.off[2] + ((long)z * 16 + j16) * KP;
#pragma unroll
        for (int r = 0; r < RT; r++) {
            f32x4 v = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll 4
            for (int zs = 0; zs < zd; zs++) v += *reinterpret_cast<const f32x4 *>(Pf + zs * part_stride + e + 16 * r + 4 * g);
            *reinterpret_cast<f32x4 *>(L + e + 16 * r + 4 * g) = v;
            *reinterpret_cast<f32x4 *>(b0 + j16 * SB + 16 * r + 4 * g) = v;
        }
    }
    __syncthreads();
    const int cxp = wave & 1, cyp = wave >> 1;
    int sp = 4;
    for (int l = 3; l <= Lbot; l++) {
        const int s = 2 * sp, np = sp * sp, ntp = (np + 15) / 16;
        const float *pa = ((l - 1) & 1) ? b1 : b0;
        float *ch = (l & 1) ? b1 : b0;
        for (int c = 0; c < ntp; c++) {
            const int p = 16 * c + j16, px = p % sp, py = p / sp;
            const bool ok = p < np;
            f32x4 acc[RT];
#pragma unroll
            for (int r = 0; r < RT; r++) acc[r] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int u4 = 0; u4 < G / 4; u4++) {
                f32x4 av[RT];
#pragma unroll
                for (int r = 0; r < RT; r++)
                    av[r] = *reinterpret_cast<const f32x4 *>(l2lfrag + (((long)wave * RT + r) * 64 + lane) * G + 4 * u4);
                const f32x4 b = ok ? *reinterpret_cast<const f32x4 *>(pa + p * SB + g * G + 4 * u4) : f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int u = 0; u < 4; u++)
#pragma unroll
                    for (int r = 0; r < RT; r++) acc[r] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[r][u], b[u], acc[r], 0, 0, 0);
            }
            if (ok) {
                const int cx = 2 * px + cxp, cy = 2 * py + cyp;
                const long e = lo.off[l] + (((long)z * s + cy) * s + cx) * KP;
#pragma unroll
                for (int r = 0; r < RT; r++) {
                    f32x4 v = acc[r];
#pragma unroll 4
                    for (int zs = 0; zs < zd; zs++)
                        v += *reinterpret_cast<const f32x4 *>(Pf + zs * part_stride + e + 16 * r + 4 * g);
                    *reinterpret_cast<f32x4 *>(L + e + 16 * r + 4 * g) = v;
                    *reinterpret_cast<f32x4 *>(ch + (cy * s + cx) * SB + 16 * r + 4 * g) = v;
                }
            }
        }
        __syncthreads();
        sp = s;
    }
}

// -------------------------------------------------------- M2L + L2L --
// A workgroup owns a dense TX x 16 block of target boxes (TX = 16 or 32) of one
// target layer; wave w handles child parity (w & 1, w >> 1): an SX x 8
// sub-lattice (SX = TX / 2) of targets in NC = SX / 2 MFMA column tiles of 16.
// The source window holds the tile plus a 3-box halo: (TX + 6) x 22 boxes.
// Interaction-list offsets (ox + 3) | (oy + 3) << 4 of child parity p =
// px + 2 py: the 6 x 6 block [-2-px, 3-px] x [-2-py, 3-py] of source boxes
// minus the 3 x 3 neighbourhood (children of the parent's neighbours minus
// own neighbours, Box.cu:121-145).
// The 28th entry repeats the 27th: the second offset of the last pair in the
// 16-slot form, whose operator block is zero.
#define MMS_M2L_OFFSETS                                                                                           \
    {{17, 18, 19, 20, 21, 22, 33, 37, 38, 49, 53, 54, 65, 69, 70, 81, 82, 83, 84, 85, 86, 97, 98, 99, 100, 101, 102, 102}, \
     {16, 17, 18, 19, 20, 21, 32, 33, 37, 48, 49, 53, 64, 65, 69, 80, 81, 82, 83, 84, 85, 96, 97, 98, 99, 100, 101, 101}, \
     {1, 2, 3, 4, 5, 6, 17, 18, 19, 20, 21, 22, 33, 37, 38, 49, 53, 54, 65, 69, 70, 81, 82, 83, 84, 85, 86, 86},           \
     {0, 1, 2, 3, 4, 5, 16, 17, 18, 19, 20, 21, 32, 33, 37, 48, 49, 53, 64, 65, 69, 80, 81, 82, 83, 84, 85, 85}}
__constant__ int c_m2l_offs[4][28] = MMS_M2L_OFFSETS;  // 32-bit entries: wave-uniform scalar loads
const int h_m2l_offs[4][28] = MMS_M2L_OFFSETS;          // the same table for the host's operator layout

template <int KP, int MODE, int NC> struct M2LCfg {
    static constexpr int G = KP / 4;    // slots per lane group (lanes 16g..16g+15)
    static constexpr int RT = KP / 16;  // 16-row output tiles
    static constexpr int SX = 2 * NC;   // sub-lattice targets per row
    static constexpr int TX = 2 * SX;   // dense tile width in boxes
    static constexpr int WX = TX + 6, WY = 22;
    static constexpr int SB = (MODE == MMS_M2L_FP32) ? KP + 4 : KP + 8;  // LDS stride per box (elements)
    static constexpr int NW = (MODE == MMS_M2L_BF16X3) ? 2 : 1;          // bf16 planes (hi, lo)
    static constexpr size_t lds_bytes =
        (MODE == MMS_M2L_FP32) ? sizeof(float) * WX * WY * SB : sizeof(unsigned short) * NW * WX * WY * SB;
};

// Levels of one M2L launch: the flat tile index blockIdx.x runs over the
// levels' tiles back to back.
struct M2LLevels {
    static constexpr int kMax = 16;
    int n;                   // levels in the launch
    int tile0[kMax + 1];     // first flat tile of level i; tile0[n] = total
    int ntx[kMax];           // tiles along x
    int nb[kMax];            // boxes per side
    long off[kMax];          // float offset of the level in mult / loc / partial sums (per film)
    long toff[kMax];         // byte offset of the level's operator fragments
};

// One workgroup = one dense tile of target boxes of one level, one target
// layer zt and one source layer zs; the result is the zs partial of the
// local expansions (summed over zs, with L2L from the parent, by
// l2l_sum_kernel).
template <int KP, int MODE, int NC, int ZP>
__global__ __launch_bounds__(256 * ZP) void m2l_mfma_kernel(const float *__restrict__ Mall, float *__restrict__ Part,
                                                            const char *__restrict__ Tall, M2LLevels lv, int zd,
                                                            long film_stride, long part_stride,
                                                            const float *__restrict__ f16scale) {
    // ZP target layers share the staged source window: waves 4 zl .. 4 zl + 3
    // take target layer zt = ZP * zg + zl
    typedef M2LCfg<KP, MODE, NC> C;
    int li = 0;
    while (li + 1 < lv.n && (int)blockIdx.x >= lv.tile0[li + 1]) li++;
    const int tl = blockIdx.x - lv.tile0[li], nb = lv.nb[li];
    const int nzg = (zd + ZP - 1) / ZP, zz = blockIdx.z;
    const int zs = zz % zd, film = zz / (zd * nzg), zt = ((zz / zd) % nzg) * ZP + (int)(threadIdx.x >> 8);
    const int ztc = zt < zd ? zt : zd - 1;  // operator loads of a padding wave stay in bounds
    const float *M = Mall + film * film_stride + lv.off[li];
    float *Lz = Part + zs * part_stride + film * film_stride + lv.off[li];
    const char *Tfrag = Tall + lv.toff[li];
    constexpr int G = C::G, RT = C::RT, SB = C::SB, SX = C::SX, WX = C::WX, WY = C::WY;
    extern __shared__ __attribute__((aligned(16))) char smem[];
    const int lane = threadIdx.x & 63, wave = (threadIdx.x >> 6) & 3;
    const int px = wave & 1, py = wave >> 1;
    const int x0 = (tl % lv.ntx[li]) * C::TX, y0 = (tl / lv.ntx[li]) * 16;
    const int j16 = lane & 15, g = lane >> 4;
    // window offset of this lane's target in column tile c: target t = 16c + j16
    int tbox[NC];
#pragma unroll
    for (int c = 0; c < NC; c++) {
        const int t = 16 * c + j16, ti = t % SX, tj = t / SX;
        tbox[c] = (2 * tj + py + 3) * WX + (2 * ti + px + 3);
    }

    f32x4 acc[RT][NC];
#pragma unroll
    for (int r = 0; r < RT; r++)
#pragma unroll
        for (int c = 0; c < NC; c++) acc[r][c] = f32x4{0.f, 0.f, 0.f, 0.f};

    // only n + m even coefficients are kept, for which the M2L operator is even
    // in dz: the blocks of +dz and -dz are equal, so only the dz >= 0 half is
    // read (half the operator data through L2)
    const int dzi = (ztc > zs ? ztc - zs : zs - ztc) + zd - 1;
    const int *offs = c_m2l_offs[__builtin_amdgcn_readfirstlane(wave)];
    // operator fragments of the first three offsets, in flight while the window is staged
    constexpr bool F16M = (MODE == MMS_M2L_FP16 || MODE == MMS_M2L_FP16X2);  // scaled FP16 source window
    constexpr int NP = (MODE == MMS_M2L_BF16X3 || MODE == MMS_M2L_FP16X2) ? 2 : 1;  // operator planes (hi, lo)
    // 16-slot blocks with 16-bit operands: the 32-deep k step of one MFMA takes
    // two offsets (lane groups 0-1 the first, 2-3 the second): 14 offset pairs
    // (the last with a zero second block) in place of 27 offsets
    constexpr bool PAIR = (KP == 16 && MODE != MMS_M2L_FP32);
    constexpr int NU = PAIR ? 14 : 27, RG = PAIR ? 2 : 3;  // product units; operator blocks in flight
    constexpr int NK = (MODE == MMS_M2L_FP32) ? G / 4 : (PAIR ? 1 : G / 8);  // 16-byte A loads per row tile and plane
    typedef typename std::conditional<F16M, f16x8, bf16x8>::type BV;  // 8 x 16-bit operand
    typedef typename std::conditional<MODE == MMS_M2L_FP32, f32x4, BV>::type AV;
    AV abuf[RG][NP][RT][NK];
    // operator lo plane offset (elements); the pair form's fragments are
    // [dz][child parity][pair][64 lanes][8]
    const long lo_base = PAIR ? (long)(2 * zd - 1) * 4 * 14 * 64 * 8 : (long)RT * 64 * G * 49 * (2 * zd - 1);
    auto loadA = [&](AV (&a)[NP][RT][NK], int u) {
        if constexpr (PAIR) {
            const long fb = (((long)dzi * 4 + wave) * 14 + u) * 64 * 8;
#pragma unroll
            for (int q = 0; q < NP; q++)
                a[q][0][0] = *reinterpret_cast<const AV *>(reinterpret_cast<const unsigned short *>(Tfrag) + q * lo_base +
                                                            fb + lane * 8);
            return;
        }
        const int code = offs[u];
        const long fbase = ((long)dzi * 49 + (code >> 4) * 7 + (code & 15)) * RT;
#pragma unroll
        for (int q = 0; q < NP; q++)
#pragma unroll
            for (int r = 0; r < RT; r++)
#pragma unroll
                for (int sk = 0; sk < NK; sk++) {
                    if constexpr (MODE == MMS_M2L_FP32)
                        a[q][r][sk] = *reinterpret_cast<const f32x4 *>(
                            reinterpret_cast<const float *>(Tfrag) + ((fbase + r) * 64 + lane) * G + 4 * sk);
                    else
                        a[q][r][sk] = *reinterpret_cast<const AV *>(
                            reinterpret_cast<const unsigned short *>(Tfrag) + q * lo_base + ((fbase + r) * 64 + lane) * G + 8 * sk);
                }
    };
#pragma unroll
    for (int k = 0; k < RG; k++) loadA(abuf[k], k);

    // stage the source window of layer zs (zero outside the grid), in
    // batches of 8 loads per thread issued before any of them is used
    float inv_s = 1.0f;  // FP16: 1 / (window scale)
    if constexpr (F16M) {
        // every thread stages the same 4 slots 4q .. 4q + 3 of its boxes
        const f32x4 *M4 = reinterpret_cast<const f32x4 *>(M);
        constexpr int NT = 256 * ZP, NQ = KP / 4, NE = WX * WY * NQ, NB = (NE + NT - 1) / NT;
        static_assert(NT % NQ == 0, "staging slot must be fixed per thread");
        const int q = threadIdx.x % NQ;
        const f32x4 cs = *reinterpret_cast<const f32x4 *>(f16scale + KP + 4 * q);
        auto ld = [&](int k) {  // k >= NB reads nothing and returns 0
            const int e = threadIdx.x + NT * k, box = e / NQ;
            const int gx = x0 - 3 + box % WX, gy = y0 - 3 + box / WX;
            f32x4 v = f32x4{0.f, 0.f, 0.f, 0.f};
            if (e < NE && gx >= 0 && gx < nb && gy >= 0 && gy < nb)
                v = M4[(((long)zs * nb + gy) * nb + gx) * NQ + q] * cs;
            return v;
        };
        auto st = [&](int k, f32x4 v, float sc) {
            const int e = threadIdx.x + NT * k;
            if (e >= NE) return;
            const int box = e / NQ;
            unsigned short h[4];
#pragma unroll
            for (int t = 0; t < 4; t++) h[t] = __builtin_bit_cast(unsigned short, (_Float16)(v[t] * sc));
            uint2 hv;
            hv.x = h[0] | ((unsigned)h[1] << 16);
            hv.y = h[2] | ((unsigned)h[3] << 16);
            *reinterpret_cast<uint2 *>(reinterpret_cast<unsigned short *>(smem) + box * SB + 4 * q) = hv;
        };
        // the whole window is held in registers when it is at most 8 loads
        // per thread; otherwise it is read twice (maximum, then conversion)
        constexpr bool HOLD = NB <= 8;
        f32x4 vb[HOLD ? NB : 8];
        float mx = 0.0f;
        if constexpr (HOLD) {
#pragma unroll
            for (int k = 0; k < NB; k++) vb[k] = ld(k);
#pragma unroll
            for (int k = 0; k < NB; k++)
#pragma unroll
                for (int t = 0; t < 4; t++) mx = fmaxf(mx, fabsf(vb[k][t]));
        } else {
#pragma unroll 1
            for (int k0 = 0; k0 < NB; k0 += 8) {
#pragma unroll
                for (int k = 0; k < 8; k++) vb[k] = ld(k0 + k);
#pragma unroll
                for (int k = 0; k < 8; k++)
#pragma unroll
                    for (int t = 0; t < 4; t++) mx = fmaxf(mx, fabsf(vb[k][t]));
            }
        }
        // workgroup maximum: wave reduction, then one value per wave in LDS
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
        __shared__ float wmax[NT / 64];
        if (lane == 0) wmax[threadIdx.x >> 6] = mx;
        __syncthreads();
        mx = wmax[0];
#pragma unroll
        for (int w = 1; w < NT / 64; w++) mx = fmaxf(mx, wmax[w]);
        // s = 2^(14 - e) for 2^e <= max < 2^(e+1), clamped so s and 1/s stay normal
        int es = 14 - ((int)((__float_as_uint(mx) >> 23) & 255u) - 127);
        es = es > 100 ? 100 : (es < -100 ? -100 : es);
        const float sc = __uint_as_float((unsigned)(es + 127) << 23);
        inv_s = __uint_as_float((unsigned)(127 - es) << 23);
        if constexpr (HOLD) {
#pragma unroll
            for (int k = 0; k < NB; k++) st(k, vb[k], sc);
        } else {
#pragma unroll 1
            for (int k0 = 0; k0 < NB; k0 += 8) {
#pragma unroll
                for (int k = 0; k < 8; k++) vb[k] = ld(k0 + k);
#pragma unroll
                for (int k = 0; k < 8; k++) st(k0 + k, vb[k], sc);
            }
        }
    } else {
        const f32x4 *M4 = reinterpret_cast<const f32x4 *>(M);
        constexpr int NT = 256 * ZP, NE = WX * WY * (KP / 4), BATCH = 8, NIT = (NE + NT * BATCH - 1) / (NT * BATCH);
#pragma unroll 1
        for (int it = 0; it < NIT; it++) {
        f32x4 vb[BATCH];
#pragma unroll
        for (int k = 0; k < BATCH; k++) {
            const int e = threadIdx.x + NT * (it * BATCH + k);
            const int box = e / (KP / 4), q = e % (KP / 4);
            const int gx = x0 - 3 + box % WX, gy = y0 - 3 + box / WX;
            vb[k] = f32x4{0.f, 0.f, 0.f, 0.f};
            if (e < NE && gx >= 0 && gx < nb && gy >= 0 && gy < nb)
                vb[k] = M4[(((long)zs * nb + gy) * nb + gx) * (KP / 4) + q];
        }
#pragma unroll
        for (int k = 0; k < BATCH; k++) {
            const int e = threadIdx.x + NT * (it * BATCH + k);
            if (e >= NE) break;
            const int box = e / (KP / 4), q = e % (KP / 4);
            const f32x4 v = vb[k];
            if constexpr (MODE == MMS_M2L_FP32) {
                *reinterpret_cast<f32x4 *>(reinterpret_cast<float *>(smem) + box * SB + 4 * q) = v;
            } else {
                unsigned short *hi = reinterpret_cast<unsigned short *>(smem);
                unsigned short h[4], l[4];
#pragma unroll
                for (int t = 0; t < 4; t++) {
                    h[t] = f2bf(v[t]);
                    l[t] = f2bf(v[t] - bf2f(h[t]));
                }
                uint2 hv, lv2;
                hv.x = h[0] | ((unsigned)h[1] << 16); hv.y = h[2] | ((unsigned)h[3] << 16);
                lv2.x = l[0] | ((unsigned)l[1] << 16); lv2.y = l[2] | ((unsigned)l[3] << 16);
                *reinterpret_cast<uint2 *>(hi + box * SB + 4 * q) = hv;
                if constexpr (MODE == MMS_M2L_BF16X3)
                    *reinterpret_cast<uint2 *>(hi + WX * WY * SB + box * SB + 4 * q) = lv2;
            }
        }
        }
    }
    __syncthreads();
    if (zt >= zd) return;  // padding waves of the last layer group (after the only barrier)

    // 27 offsets (14 offset pairs), RG operator blocks in flight (ring of registers)
#pragma unroll 1
    for (int ob = 0; ob < NU; ob += RG) {
#pragma unroll
        for (int k = 0; k < RG; k++) {
            // window offset of the source box; the pair form: lane groups 0-1
            // read the first offset's box, 2-3 the second's, at slots 8 (g % 2)
            int woff, slot;
            if constexpr (PAIR) {
                const int ca = offs[2 * (ob + k)], cb = offs[2 * (ob + k) + 1], code = (g < 2) ? ca : cb;
                woff = ((code >> 4) - 3) * WX + ((code & 15) - 3);
                slot = 8 * (g & 1);
            } else {
                const int code = offs[ob + k];
                woff = ((code >> 4) - 3) * WX + ((code & 15) - 3);
                slot = g * G;
            }
            if constexpr (MODE == MMS_M2L_FP32) {
                const float *lds = reinterpret_cast<const float *>(smem);
#pragma unroll
                for (int u4 = 0; u4 < NK; u4++) {
                    f32x4 b[NC];
#pragma unroll
                    for (int c = 0; c < NC; c++)
                        b[c] = *reinterpret_cast<const f32x4 *>(lds + (tbox[c] + woff) * SB + slot + 4 * u4);
#pragma unroll
                    for (int u = 0; u < 4; u++)
#pragma unroll
                        for (int r = 0; r < RT; r++)
#pragma unroll
                            for (int c = 0; c < NC; c++)
                                acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x4f32(abuf[k][0][r][u4][u], b[c][u], acc[r][c], 0, 0, 0);
                }
            } else {
                const unsigned short *hi = reinterpret_cast<const unsigned short *>(smem);
#pragma unroll
                for (int sk = 0; sk < NK; sk++) {
                    BV bh[NC];
#pragma unroll
                    for (int c = 0; c < NC; c++)
                        bh[c] = *reinterpret_cast<const BV *>(hi + (tbox[c] + woff) * SB + slot + 8 * sk);
#pragma unroll
                    for (int r = 0; r < RT; r++)
#pragma unroll
                        for (int c = 0; c < NC; c++)
                            acc[r][c] = mfma32(abuf[k][0][r][sk], bh[c], acc[r][c]);
                    if constexpr (MODE == MMS_M2L_FP16X2) {
#pragma unroll
                        for (int r = 0; r < RT; r++)
#pragma unroll
                            for (int c = 0; c < NC; c++)
                                acc[r][c] = mfma32(abuf[k][1][r][sk], bh[c], acc[r][c]);
                    }
                    if constexpr (MODE == MMS_M2L_BF16X3) {
                        bf16x8 bl[NC];
#pragma unroll
                        for (int c = 0; c < NC; c++)
                            bl[c] = *reinterpret_cast<const bf16x8 *>(hi + WX * WY * SB + (tbox[c] + woff) * SB + slot + 8 * sk);
#pragma unroll
                        for (int r = 0; r < RT; r++)
#pragma unroll
                            for (int c = 0; c < NC; c++)
                                acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(abuf[k][0][r][sk], bl[c], acc[r][c], 0, 0, 0);
#pragma unroll
                        for (int r = 0; r < RT; r++)
#pragma unroll
                            for (int c = 0; c < NC; c++)
                                acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(abuf[k][NP - 1][r][sk], bh[c], acc[r][c], 0, 0, 0);
                    }
                }
            }
            if (ob + RG < NU) loadA(abuf[k], ob + RG + k);
        }
    }
    // D tile row 4g + q (q = 0..3), column j16: slots 16r + 4g .. +3 of target j16
    if constexpr (F16M) {
#pragma unroll
        for (int r = 0; r < RT; r++) {
            const f32x4 rs = *reinterpret_cast<const f32x4 *>(f16scale + 16 * r + 4 * g) * inv_s;
#pragma unroll
            for (int c = 0; c < NC; c++) acc[r][c] *= rs;
        }
    }
#pragma unroll
    for (int c = 0; c < NC; c++) {
        const int t = 16 * c + j16;
        const int bx = x0 + 2 * (t % SX) + px, by = y0 + 2 * (t / SX) + py;
        if (bx >= nb || by >= nb) continue;
        f32x4 *dst = reinterpret_cast<f32x4 *>(Lz + (((long)zt * nb + by) * nb + bx) * KP);
#pragma unroll
        for (int r = 0; r < RT; r++) dst[4 * r + g] = acc[r][c];
    }
}

// Local expansions of one level: the source-layer partial sums of M2L plus
// L2L from the parent level (same layer), FP32 MFMA.  Wave w = child parity;
// 64 consecutive parents of one row per workgroup, 16 per column tile.
template <int KP>
__global__ __launch_bounds__(256) void l2l_sum_kernel(const float *__restrict__ Part, long part_stride,
                                                      float *__restrict__ Lall, long off, long off_par,
                                                      const float *__restrict__ l2lfrag, int nb, int zd,
                                                      long film_stride) {
    constexpr int G = KP / 4, RT = KP / 16, NC = 4;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int px = wave & 1, py = wave >> 1, j16 = lane & 15, g = lane >> 4;
    const int film = blockIdx.z / zd, z = blockIdx.z % zd, half = nb >> 1;
    const int pby = blockIdx.y, pbx0 = blockIdx.x * 64;
    float *L = Lall + film * film_stride + off;
    const float *Pz = Part + film * film_stride + off;
    f32x4 acc[RT][NC];
#pragma unroll
    for (int r = 0; r < RT; r++)
#pragma unroll
        for (int c = 0; c < NC; c++) acc[r][c] = f32x4{0.f, 0.f, 0.f, 0.f};
    if (off_par >= 0) {
        const float *Lp = Lall + film * film_stride + off_par;
#pragma unroll
        for (int u4 = 0; u4 < G / 4; u4++) {
            f32x4 a[RT], b[NC];
#pragma unroll
            for (int r = 0; r < RT; r++)
                a[r] = *reinterpret_cast<const f32x4 *>(l2lfrag + (((long)wave * RT + r) * 64 + lane) * G + 4 * u4);
#pragma unroll
            for (int c = 0; c < NC; c++) {
                const int pbx = pbx0 + 16 * c + j16;
                b[c] = f32x4{0.f, 0.f, 0.f, 0.f};
                if (pbx < half)
                    b[c] = *reinterpret_cast<const f32x4 *>(Lp + (((long)z * half + pby) * half + pbx) * KP + g * G + 4 * u4);
            }
#pragma unroll
            for (int u = 0; u < 4; u++)
#pragma unroll
                for (int r = 0; r < RT; r++)
#pragma unroll
                    for (int c = 0; c < NC; c++)
                        acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[r][u], b[c][u], acc[r][c], 0, 0, 0);
        }
    }
#pragma unroll
    for (int c = 0; c < NC; c++) {
        const int pbx = pbx0 + 16 * c + j16;
        if (pbx >= half) continue;
        const long e = (((long)z * nb + 2 * pby + py) * nb + 2 * pbx + px) * KP;
#pragma unroll
        for (int r = 0; r < RT; r++) {
            f32x4 v = acc[r][c];
#pragma unroll 4
            for (int s = 0; s < zd; s++) v += *reinterpret_cast<const f32x4 *>(Pz + s * part_stride + e + 4 * (4 * r + g));
            *reinterpret_cast<f32x4 *>(L + e + 4 * (4 * r + g)) = v;
        }
    }
}

// -------------------------------------------------------- P2P + L2P --
// Block = TW x TW target cells of one layer (TW = 16); the charges of the
// tile plus a W-cell halo in every layer are staged in LDS together with the
// 1/r table over all (dx, dy, dz) offsets of the near field.
template <int K, int KP, int W>
__global__ __launch_bounds__(256) void p2p_l2p_kernel(const float *__restrict__ rho, float *__restrict__ phi,
                                                      const float *__restrict__ Lleaf, const float *__restrict__ l2p,
                                                      const float *__restrict__ table, int xd, int yd, int zd,
                                                      int nb, float scale, FilmSet fs) {
    constexpr int TW = 16, SW = TW + 2 * W, TS = 4 * W - 1;
    extern __shared__ float lds[];
    float *src = lds;                      // [zd][SW][SW]
    float *tab = lds + zd * SW * SW;       // [2zd-1][TS][TS]
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const int x0 = blockIdx.x * TW, y0 = blockIdx.y * TW, zt = blockIdx.z % zd;
    {
        const int f = blockIdx.z / zd;
        if (fs.desc) {
            const FilmDesc d = fs.desc[f];
            xd = d.xd; yd = d.yd; rho += d.off; phi += d.off;
            if (x0 >= xd || y0 >= yd) return;  // block-uniform
        }
        Lleaf += (long)f * fs.stride;
    }
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
            const int sy = ly0 + syi;                   // LDS row of the source
            const int dyi = sy - (ty + W) + 2 * W - 1;  // table row
            const float *srow = s + sy * SW + lx0;
            const float *trow = tb + dyi * TS + (lx0 - (tx + W) + 2 * W - 1);
#pragma unroll
            for (int sxi = 0; sxi < 3 * W; sxi++) near = fmaf(srow[sxi], trow[sxi], near);
        }
    }
    if (x >= xd || y >= yd) return;
    const int bx = x / W, by = y / W, c = (y % W) * W + (x % W);
    const float *Lv = Lleaf + (((long)zt * nb + by) * nb + bx) * KP;
    const float *row = l2p + (long)c * K;
    float far = 0.0f;
#pragma unroll
    for (int k = 0; k < K; k++) far = fmaf(row[k], Lv[k], far);
    phi[(long)zt * plane + (long)y * xd + x] = (near + far) * scale;
}

// ---------------------------------------------- P2P on the matrix cores --
// Leaf width W (4 or 8): a leaf box has W*W cells = RT 16-row MFMA tiles.  A
// block owns 4 x 4 leaf boxes (4W x 4W cells, one 16-column tile) and all
// layers; wave zt (looping when zd > 4) computes for its target layer
//   phi[cell i of box j] = sum_k G[i][k] q_j[k]
// where k runs over the 3W x 3W x zd source window of box j (its 3 x 3
// neighbour leaves in every layer) and G[i][k] = 1/r is read from the LDS
// offset table (self entry 0).  FP32 MFMA 16x16x4 with k = (zs, sy, sx) and
// sx in groups of 4 matching the 4 lane groups.  The downward pass adds the
// far field afterwards.
template <int W>
__global__ __launch_bounds__(256) void p2p_mfma_kernel(const float *__restrict__ rho, float *__restrict__ phi,
                                                       const float *__restrict__ table, int xd, int yd, int zd,
                                                       float scale, FilmSet fs) {
    constexpr int TW = 4 * W, SW = TW + 2 * W, TS = 4 * W - 1;
    if (fs.desc) {
        const FilmDesc d = fs.desc[blockIdx.z];
        xd = d.xd; yd = d.yd; rho += d.off; phi += d.off;
        if ((int)blockIdx.x * TW >= xd || (int)blockIdx.y * TW >= yd) return;  // block-uniform
    }
    constexpr int RT = W * W / 16, NSG = 3 * W / 4;
    extern __shared__ float lds[];
    float *src = lds;                      // [zd][SW][SW]
    float *tab = lds + zd * SW * SW;       // [2zd-1][TS][TS]
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int j16 = lane & 15, g = lane >> 4;
    const int x0 = blockIdx.x * TW, y0 = blockIdx.y * TW;
    const long plane = (long)xd * yd;
    {   // charges of the window and the 1/r table, 8 loads per thread in flight per batch
        const int ne = zd * SW * SW, nt = (2 * zd - 1) * TS * TS;
#pragma unroll 1
        for (int e0 = 0; e0 < ne; e0 += 256 * 8) {
            float vb[8];
#pragma unroll
            for (int k = 0; k < 8; k++) {
                const int e = e0 + threadIdx.x + 256 * k;
                const int zs = e / (SW * SW), r = e % (SW * SW), sy = r / SW, sx = r % SW;
                const int gx = x0 - W + sx, gy = y0 - W + sy;
                vb[k] = (e < ne && gx >= 0 && gx < xd && gy >= 0 && gy < yd) ? rho[(long)zs * plane + (long)gy * xd + gx]
                                                                             : 0.0f;
            }
#pragma unroll
            for (int k = 0; k < 8; k++) {
                const int e = e0 + threadIdx.x + 256 * k;
                if (e < ne) src[e] = vb[k];
            }
        }
#pragma unroll 1
        for (int e0 = 0; e0 < nt; e0 += 256 * 8) {
            float vb[8];
#pragma unroll
            for (int k = 0; k < 8; k++) {
                const int e = e0 + threadIdx.x + 256 * k;
                vb[k] = e < nt ? table[e] : 0.0f;
            }
#pragma unroll
            for (int k = 0; k < 8; k++) {
                const int e = e0 + threadIdx.x + 256 * k;
                if (e < nt) tab[e] = vb[k];
            }
        }
    }
    __syncthreads();
    // B column j16: box (jx, jy) of the block; A row j16 of tile r: cell 16 r + j16
    const int jx = j16 & 3, jy = j16 >> 2;
    int toff[RT];
#pragma unroll
    for (int r = 0; r < RT; r++) {
        const int c = 16 * r + j16, iy = c / W, ix = c % W;
        toff[r] = (W - 1 - iy) * TS + (W - 1 - ix);
    }
    for (int zt = wave; zt < zd; zt += 4) {
        f32x4 acc[RT];
#pragma unroll
        for (int r = 0; r < RT; r++) acc[r] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll 1
        for (int zs = 0; zs < zd; zs++) {
            const float *tb = tab + (zs - zt + zd - 1) * TS * TS;
            const float *sb = src + zs * SW * SW + (W * jy) * SW + (W * jx);
#pragma unroll 2
            for (int sy = 0; sy < 3 * W; sy++) {
#pragma unroll
                for (int sg = 0; sg < NSG; sg++) {
                    const int sx = 4 * sg + g;  // this lane group's source column
                    const float b = sb[sy * SW + sx];
#pragma unroll
                    for (int r = 0; r < RT; r++)
                        acc[r] = __builtin_amdgcn_mfma_f32_16x16x4f32(tb[toff[r] + sy * TS + sx], b, acc[r], 0, 0, 0);
                }
            }
        }
        // D row 4g + q of tile r = cell 16 r + 4g + q of box j16: 4 consecutive x cells
#pragma unroll
        for (int r = 0; r < RT; r++) {
            const int c = 16 * r + 4 * g;
            const int y = y0 + W * jy + c / W, xb = x0 + W * jx + c % W;
            if (y >= yd) continue;
            float *dst = phi + (long)zt * plane + (long)y * xd + xb;
            if (xb + 3 < xd && (((long)zt * plane + (long)y * xd + xb) & 3) == 0) {
                *reinterpret_cast<f32x4 *>(dst) = acc[r] * scale;
            } else {
#pragma unroll
                for (int q = 0; q < 4; q++)
                    if (xb + q < xd) dst[q] = acc[r][q] * scale;
            }
        }
    }
}

// Near field in the BF16 split form, leaf width 8.  For a target row pair
// (iy = 2p, 2p + 1) and a source row sy of the 24 x 24 window of each box,
// the 1/r sum over the row is one product with a constant operator
//   A_{dz,dy}[8 r + ix][sx] = G(dz, dy - r, sx - ix),  dy = sy - 2p + 7,
// (16 target cells x 32 sources, sources 24..31 zero), stored per (layer
// offset, dy) as BF16 hi / lo MFMA fragments in lane order.  The column tile
// is the 16 boxes of the block (B = the boxes' source row, BF16 hi / lo from
// LDS), so per (dz, sy) a wave issues 4 pairs x 3 products of
// v_mfma_f32_16x16x32_bf16 (Ah Bh + Ah Bl + Al Bh).  Walking sy with a
// stride of 2, three of the four operators of the next row are the previous
// row's: a four-deep register window, one new operator per row.
// F16: one v_mfma_f32_16x16x32_f16 product per step instead of three; the
// window is scaled by s = 2^e (its largest |charge| in [2^14, 2^15)) before
// the FP16 conversion and the sums by 1 / s afterwards (exact), the 1/r
// operator (values in (0, 1]) is stored as plain FP16.
// The downward pass adds the far field afterwards.
// NF: 0 = BF16 hi/lo split of operator and charges (3 products), 1 = scaled
// FP16 operator and charges (1 product), 2 = FP16 hi/lo split of the operator
// against the scaled FP16 charges (2 products).
template <int NCB, int NF>
__global__ __launch_bounds__(256) void p2p_16bit_kernel(const float *__restrict__ rho, float *__restrict__ phi,
                                                        const unsigned short *__restrict__ afrag, int xd, int yd,
                                                        int zd, float scale, FilmSet fs) {
    constexpr bool F16 = NF != 0;          // scaled FP16 charge window (one plane)
    constexpr bool OP_LO = NF != 1;        // operator lo plane read
    // block = (4 NCB) x 4 boxes; NCB column tiles of 16 boxes share each operator load
    constexpr int W = 8, TWX = 32 * NCB, TWY = 32, SWX = TWX + 16, SWY = 48, SP = SWX + 8;
    if (fs.desc) {
        const FilmDesc d = fs.desc[blockIdx.z];
        xd = d.xd; yd = d.yd; rho += d.off; phi += d.off;
        if ((int)blockIdx.x * TWX >= xd || (int)blockIdx.y * TWY >= yd) return;  // block-uniform
    }
    constexpr int RT = 4;
    extern __shared__ __attribute__((aligned(16))) unsigned short sh[];
    unsigned short *qh = sh, *ql = sh + zd * SWY * SP;  // [zd][SWY][SP] hi, lo (F16: hi only)
    typedef typename std::conditional<F16, f16x8, bf16x8>::type BV;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int j16 = lane & 15, g = lane >> 4;
    const int x0 = blockIdx.x * TWX, y0 = blockIdx.y * TWY;
    const long plane = (long)xd * yd;
    float inv_s = 1.0f;
    if constexpr (F16) {
        // The window's largest |charge| fixes the scale: up to 4 layers
        // (64 charges per thread) the window is loaded once into registers,
        // deeper films read it twice (maximum, then the scaled values).
        constexpr int BATCH = 16, HOLD = 64;
        const int ne = zd * SWY * SWX;
        auto ld = [&](int e) {  // e >= ne reads nothing and returns 0
            const int zs = e / (SWY * SWX), r = e % (SWY * SWX), sy = r / SWX, sx = r % SWX;
            const int gx = x0 - W + sx, gy = y0 - W + sy;
            return (e < ne && gx >= 0 && gx < xd && gy >= 0 && gy < yd) ? rho[(long)zs * plane + (long)gy * xd + gx]
                                                                      : 0.0f;
        };
        auto st = [&](int e, float v, float sc) {
            if (e >= ne) return;
            const int zs = e / (SWY * SWX), r = e % (SWY * SWX), sy = r / SWX, sx = r % SWX;
            qh[(zs * SWY + sy) * SP + sx] = __builtin_bit_cast(unsigned short, (_Float16)(v * sc));
        };
        const bool hold = ne <= 256 * HOLD;  // uniform
        float vh[HOLD];
        float mx = 0.0f;
        if (hold) {
#pragma unroll
            for (int k = 0; k < HOLD; k++) vh[k] = ld(threadIdx.x + 256 * k);
#pragma unroll
            for (int k = 0; k < HOLD; k++) mx = fmaxf(mx, fabsf(vh[k]));
        } else {
#pragma unroll 1
            for (int e0 = 0; e0 < ne; e0 += 256 * BATCH) {
                float vb[BATCH];
#pragma unroll
                for (int k = 0; k < BATCH; k++) vb[k] = ld(e0 + threadIdx.x + 256 * k);
#pragma unroll
                for (int k = 0; k < BATCH; k++) mx = fmaxf(mx, fabsf(vb[k]));
            }
        }
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
        __shared__ float wmax[4];
        if (lane == 0) wmax[wave] = mx;
        __syncthreads();
        mx = fmaxf(fmaxf(wmax[0], wmax[1]), fmaxf(wmax[2], wmax[3]));
        // s = 2^(14 - e) for 2^e <= max < 2^(e+1), clamped so s and 1/s stay normal
        int es = 14 - ((int)((__float_as_uint(mx) >> 23) & 255u) - 127);
        es = es > 100 ? 100 : (es < -100 ? -100 : es);
        const float sc = __uint_as_float((unsigned)(es + 127) << 23);
        inv_s = __uint_as_float((unsigned)(127 - es) << 23);
        if (hold) {
#pragma unroll
            for (int k = 0; k < HOLD; k++) st(threadIdx.x + 256 * k, vh[k], sc);
        } else {
#pragma unroll 1
            for (int e0 = 0; e0 < ne; e0 += 256 * BATCH) {
                float vb[BATCH];
#pragma unroll
                for (int k = 0; k < BATCH; k++) vb[k] = ld(e0 + threadIdx.x + 256 * k);
#pragma unroll
                for (int k = 0; k < BATCH; k++) st(e0 + threadIdx.x + 256 * k, vb[k], sc);
            }
        }
    } else {   // charges of the window, 16 loads per thread in flight per batch
        constexpr int BATCH = 16;
        const int ne = zd * SWY * SWX;
#pragma unroll 1
        for (int e0 = 0; e0 < ne; e0 += 256 * BATCH) {
            float vb[BATCH];
#pragma unroll
            for (int k = 0; k < BATCH; k++) {
                const int e = e0 + threadIdx.x + 256 * k;
                const int zs = e / (SWY * SWX), r = e % (SWY * SWX), sy = r / SWX, sx = r % SWX;
                const int gx = x0 - W + sx, gy = y0 - W + sy;
                vb[k] = (e < ne && gx >= 0 && gx < xd && gy >= 0 && gy < yd) ? rho[(long)zs * plane + (long)gy * xd + gx]
                                                                             : 0.0f;
            }
#pragma unroll
            for (int k = 0; k < BATCH; k++) {
                const int e = e0 + threadIdx.x + 256 * k;
                if (e >= ne) break;
                const int zs = e / (SWY * SWX), r = e % (SWY * SWX), sy = r / SWX, sx = r % SWX;
                const unsigned short h = f2bf(vb[k]);
                qh[(zs * SWY + sy) * SP + sx] = h;
                ql[(zs * SWY + sy) * SP + sx] = f2bf(vb[k] - bf2f(h));
            }
        }
    }
    __syncthreads();
    const int jy = j16 >> 2;
    for (int zt = wave; zt < zd; zt += 4) {
        f32x4 acc[RT][NCB];
#pragma unroll
        for (int r = 0; r < RT; r++)
#pragma unroll
            for (int c = 0; c < NCB; c++) acc[r][c] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll 1
        for (int zs = 0; zs < zd; zs++) {
            // products of this source layer go to fresh accumulators that are
            // added to acc at the end of the iteration: no MFMA accumulator is
            // carried across the loop's back edge (see hazard_lint.py)
            f32x4 accz[RT][NCB];
#pragma unroll
            for (int r = 0; r < RT; r++)
#pragma unroll
                for (int c = 0; c < NCB; c++) accz[r][c] = f32x4{0.f, 0.f, 0.f, 0.f};
            const int dzi = (zs > zt ? zs - zt : zt - zs) + zd - 1;  // 1/r is even in dz
            const unsigned short *Ab = afrag + (long)dzi * 42 * 2 * 64 * 8 + lane * 8;
            auto loadA = [&](BV &h, BV &l, int op) {
                h = *reinterpret_cast<const BV *>(Ab + (long)(op * 2) * 64 * 8);
                if constexpr (OP_LO) l = *reinterpret_cast<const BV *>(Ab + (long)(op * 2 + 1) * 64 * 8);
            };
            const int bbase = (zs * SWY + W * jy) * SP;  // row 0 of this lane's box window in layer zs
#pragma unroll
            for (int ph = 0; ph < 3; ph++) {
                // k step (a, ph) = flat window sources 96 a + 32 ph .. + 31 (rows of 24 packed
                // back to back): lane group g reads 8 sources of row 4 a + srow at column scol
                const int fl = 32 * ph + 8 * g, srow = fl / 24, scol = fl % 24;
                // window: A[p] = operator (ph, d = 4 a - 2 p), index 14 ph + 2 a - p + 3;
                // An0 / An1 = pairs 0 / 1 of step a + 1
                BV Ah[4], Al[4], An0h, An0l, An1h, An1l;
#pragma unroll
                for (int p = 0; p < 4; p++) loadA(Ah[p], Al[p], 14 * ph + 3 - p);
                loadA(An0h, An0l, 14 * ph + 5);
                loadA(An1h, An1l, 14 * ph + 4);
#pragma unroll
                for (int aa = 0; aa < 6; aa++) {
                    const int ro = bbase + (4 * aa + srow) * SP + scol;
                    BV bh[NCB], bl[NCB];
#pragma unroll
                    for (int c = 0; c < NCB; c++) {
                        const int jx = 4 * c + (j16 & 3);
                        bh[c] = *reinterpret_cast<const BV *>(qh + ro + W * jx);
                        if constexpr (!F16) bl[c] = *reinterpret_cast<const BV *>(ql + ro + W * jx);
                    }
#pragma unroll
                    for (int p = 0; p < 4; p++)
#pragma unroll
                        for (int c = 0; c < NCB; c++) {
                            accz[p][c] = mfma32(Ah[p], bh[c], accz[p][c]);
                            if constexpr (!F16) {
                                accz[p][c] = mfma32(Ah[p], bl[c], accz[p][c]);
                                accz[p][c] = mfma32(Al[p], bh[c], accz[p][c]);
                            } else if constexpr (NF == 2) {
                                accz[p][c] = mfma32(Al[p], bh[c], accz[p][c]);
                            }
                        }
                    if (aa < 5) {  // slide: pairs 2, 3 of step a + 1 use pairs 0, 1 of step a
                        Ah[3] = Ah[1]; Al[3] = Al[1];
                        Ah[2] = Ah[0]; Al[2] = Al[0];
                        Ah[0] = An0h; Al[0] = An0l;
                        Ah[1] = An1h; Al[1] = An1l;
                        if (aa < 4) {
                            loadA(An0h, An0l, 14 * ph + 2 * aa + 7);
                            loadA(An1h, An1l, 14 * ph + 2 * aa + 6);
                        }
                    }
                }
            }
#pragma unroll
            for (int r = 0; r < RT; r++)
#pragma unroll
                for (int c = 0; c < NCB; c++) acc[r][c] += accz[r][c];
        }
        if constexpr (F16) {
#pragma unroll
            for (int r = 0; r < RT; r++)
#pragma unroll
                for (int c = 0; c < NCB; c++) acc[r][c] *= inv_s;
        }
#pragma unroll
        for (int c = 0; c < NCB; c++) {
            const int jx = 4 * c + (j16 & 3);
            // D row 4g + q of tile r = cell 16 r + 4g + q of box j16: 4 consecutive x cells
#pragma unroll
            for (int r = 0; r < RT; r++) {
                const int cc = 16 * r + 4 * g;
                const int y = y0 + W * jy + cc / W, xb = x0 + W * jx + cc % W;
                if (y >= yd) continue;
                float *dst = phi + (long)zt * plane + (long)y * xd + xb;
                if (xb + 3 < xd && (((long)zt * plane + (long)y * xd + xb) & 3) == 0) {
                    *reinterpret_cast<f32x4 *>(dst) = acc[r][c] * scale;
                } else {
#pragma unroll
                    for (int q = 0; q < 4; q++)
                        if (xb + q < xd) dst[q] = acc[r][c][q] * scale;
                }
            }
        }
    }
}

// -------------------------------------------------------------- dispatch --
template <int KP, int MODE, int NC>
int launch_m2l(const FmmPlan &pl, hipStream_t s) {
    typedef M2LCfg<KP, MODE, NC> C;
    // target layers per staged source window (two for 16- and 32-slot blocks;
    // 64-slot blocks need the registers of a 256-thread workgroup)
    constexpr int ZP = (KP <= 32) ? 2 : 1;
    static int attr_done = 0;
    if (!attr_done) {
        MMS_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&m2l_mfma_kernel<KP, MODE, NC, ZP>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)C::lds_bytes));
        attr_done = 1;
    }
    M2LLevels lv;
    lv.n = 0;
    int tiles = 0;
    // finest level first: its tiles are the bulk of the launch
    for (int L = pl.Lleaf; L >= 2; L--) {
        MMS_CHECK(lv.n < M2LLevels::kMax, "too many FMM levels");
        const int nb = 1 << L, i = lv.n++;
        lv.tile0[i] = tiles;
        lv.ntx[i] = (nb + C::TX - 1) / C::TX;
        lv.nb[i] = nb;
        lv.off[i] = pl.lvl_off[L];
        lv.toff[i] = (long)(L - 2) * (long)pl.m2l_level_bytes;
        tiles += lv.ntx[i] * ((nb + 15) / 16);
    }
    lv.tile0[lv.n] = tiles;
    const long zz = (long)pl.nfilm * ((pl.zd + ZP - 1) / ZP) * pl.zd;
    MMS_CHECK(zz <= 65535, "FMM batch too large for one M2L launch (films x layers^2 > 65535)");
    hipLaunchKernelGGL((m2l_mfma_kernel<KP, MODE, NC, ZP>), dim3(tiles, 1, (unsigned)zz), dim3(256 * ZP), (C::lds_bytes), s,
                       pl.mult, pl.part, reinterpret_cast<const char *>(pl.d_m2l), lv, pl.zd, pl.film_stride,
                       pl.film_stride * pl.nfilm, pl.d_m2l_scale);
    MMS_HIP(hipGetLastError());
    return 0;
}

template <int KP, int MODE>
int launch_m2l_nc(const FmmPlan &pl, hipStream_t s) {
    // wide tiles (8 column tiles per wave) when the source window fits in LDS
    if constexpr (M2LCfg<KP, MODE, 8>::lds_bytes <= 160 * 1024) {
        if (pl.m2l_cols == 8) return launch_m2l<KP, MODE, 8>(pl, s);
    }
    static_assert(M2LCfg<KP, MODE, 4>::lds_bytes <= 160 * 1024, "M2L source window exceeds LDS");
    return launch_m2l<KP, MODE, 4>(pl, s);
}

template <int KP>
int launch_l2l_sum(const FmmPlan &pl, int L, hipStream_t s) {
    const int nb = 1 << L, half = nb >> 1;
    dim3 grid((half + 63) / 64, half, pl.zd * pl.nfilm);
    hipLaunchKernelGGL((l2l_sum_kernel<KP>), grid, dim3(256), 0, s, pl.part, pl.film_stride * pl.nfilm, pl.loc,
                       pl.lvl_off[L], L > 2 ? pl.lvl_off[L - 1] : -1L, pl.d_l2l, nb, pl.zd, pl.film_stride);
    MMS_HIP(hipGetLastError());
    return 0;
}

// Dynamic LDS asked by a near-field launch that runs beside the multipole
// passes.  BF16 split and FP32 near field: at least 82 KB, so one near-field
// workgroup and one BF16 split M2L workgroup (77 KB) share a CU instead of
// two near-field workgroups filling it.  FP16 near field: at least 40 KB
// (its window needs 34 KB), beside one 77 KB or three 39 KB M2L workgroups
// (measured on the 1024 x 1024 x 4 film against 82 KB: 1210 vs 1162 steps/s
// with the BF16 split M2L, 1313 vs 1265 with the FP16 M2L; with the FP16x2
// M2L 34, 40, 56 and 80 KB gave 1451, 1452, 1380 and 1281 steps/s).
size_t near_lds_shared(const FmmPlan &pl) { return (pl.near_f16 ? 40 : 82) * 1024; }

template <int NCB, int NF>
int launch_p2p_16bit(const Grid &g, const FmmPlan &pl, const float *rho, float *phi, float scale, bool shared,
                     hipStream_t s) {
    constexpr bool F16 = NF != 0;
    constexpr int SP = 32 * NCB + 16 + 8;
    size_t lds = sizeof(unsigned short) * (F16 ? 1 : 2) * (size_t)g.zd * 48 * SP;
    if (shared && lds < near_lds_shared(pl)) lds = near_lds_shared(pl);
    constexpr int kMaxDyn = 160 * 1024 - 256;  // the FP16 form also has a few bytes of static LDS
    MMS_CHECK(lds <= (size_t)kMaxDyn, "FMM near field: too many layers for the LDS tile");
    static int attr = 0;
    if (!attr) {
        MMS_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&p2p_16bit_kernel<NCB, NF>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, kMaxDyn));
        attr = 1;
    }
    dim3 grid((g.xd + 32 * NCB - 1) / (32 * NCB), (g.yd + 31) / 32, pl.nfilm);
    hipLaunchKernelGGL((p2p_16bit_kernel<NCB, NF>), grid, dim3(256), lds, s, rho, phi, pl.d_p2pA, g.xd, g.yd,
                       g.zd, scale, pl.films());
    MMS_HIP(hipGetLastError());
    return 0;
}

// Near field: leaf widths 4 and 8 write the near field alone (the downward
// pass adds the far field); leaf width 2 adds L2P in the same kernel.
template <int K, int KP, int W>
int launch_near(const Grid &g, const FmmPlan &pl, const float *rho, float *phi, float scale, bool shared,
                hipStream_t s) {
    const int zd = g.zd;
    if constexpr (W == 8) {
        if (pl.near_f16) {
            if (pl.near_f16x2) return launch_p2p_16bit<2, 2>(g, pl, rho, phi, scale, shared, s);
            return launch_p2p_16bit<2, 1>(g, pl, rho, phi, scale, shared, s);
        }
        if (pl.precision != MMS_M2L_FP32) return launch_p2p_16bit<2, 0>(g, pl, rho, phi, scale, shared, s);
    }
    if constexpr (W == 4 || W == 8) {
        constexpr int TW = 4 * W, SW = TW + 2 * W, TS = 4 * W - 1;
        size_t lds = sizeof(float) * ((size_t)zd * SW * SW + (size_t)(2 * zd - 1) * TS * TS);
        if (shared && lds < near_lds_shared(pl)) lds = near_lds_shared(pl);
        static int attr = 0;
        if (!attr) {
            MMS_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&p2p_mfma_kernel<W>),
                                        hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
            attr = 1;
        }
        MMS_CHECK(lds <= 160 * 1024, "FMM near field: too many layers for the LDS tile");
        dim3 grid((g.xd + TW - 1) / TW, (g.yd + TW - 1) / TW, pl.nfilm);
        hipLaunchKernelGGL((p2p_mfma_kernel<W>), grid, dim3(256), lds, s, rho, phi, pl.d_near, g.xd, g.yd, zd,
                           scale, pl.films());
    } else {
        constexpr int TW = 16, SW = TW + 2 * W, TS = 4 * W - 1;
        const size_t lds = sizeof(float) * ((size_t)zd * SW * SW + (size_t)(2 * zd - 1) * TS * TS);
        dim3 grid((g.xd + TW - 1) / TW, (g.yd + TW - 1) / TW, zd * pl.nfilm);
        hipLaunchKernelGGL((p2p_l2p_kernel<K, KP, W>), grid, dim3(256), lds, s, rho, phi,
                           pl.loc + pl.lvl_off[pl.Lleaf], pl.d_l2p, pl.d_near, g.xd, g.yd, zd, 1 << pl.Lleaf, scale,
                           pl.films());
    }
    MMS_HIP(hipGetLastError());
    return 0;
}

template <int K, int W>
int run_fmm(const Grid &g, FmmPlan &pl, const float *rho, float *phi, hipStream_t s, Profiler *prof) {
    constexpr int KP = KPad<K>::KP, SB = KP + 4;
    const int zd = g.zd;
    LevelOffsets lo{};
    for (int L = 2; L <= pl.Lleaf; L++) lo.off[L] = pl.lvl_off[L];
    // subtrees rooted at level R hold up to 8 x 8 leaves (16 x 16 when the
    // grid is so deep that the coarse levels would pass level 4)
    int D = (pl.Lleaf - 3 <= 5) ? 3 : 4;
    if (pl.Lleaf - D < 2) D = pl.Lleaf - 2;
    const int R = pl.Lleaf - D, nR = 1 << R;
    const int Lc = R - 1 < 4 ? R - 1 : 4;  // coarse levels 2..Lc in one workgroup per layer
    const float scale = (float)((double)(g.dx * g.dy * g.dz) / (double)g.dx);
    // Leaf widths 4 and 8: the near field (no L2P) writes phi first, on the
    // plan's second stream concurrently with the multipole passes unless a
    // profiling window wants one kernel at a time; L2P adds the far field
    // after the join.  Leaf width 2 keeps the fused near field + L2P at the end.
    constexpr bool split = (W == 4 || W == 8);
    const bool conc = split && pl.s_near && !(prof && prof->on);
    if constexpr (split) {
        if (conc) {
            MMS_HIP(hipEventRecord(pl.ev_fork, s));
        } else {
            KTimer kt(prof, K_P2P_L2P);
            if (launch_near<K, KP, W>(g, pl, rho, phi, scale, false, s)) return 1;
        }
    }
    {
        KTimer kt(prof, K_P2M);
        const int S = 1 << D, CS = S * W;
        const long bufA = std::max((long)CS * CS, (long)S * S * SB);
        const long bufB = (long)std::max(1, (S / 2) * (S / 2)) * SB;
        const size_t lds = sizeof(float) * (size_t)(bufA + bufB);
        MMS_CHECK(lds <= 160 * 1024, "FMM upward pass: subtree exceeds LDS");
        static int attr = 0;
        if (!attr) {
            MMS_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&up_subtree_kernel<K, KP, W>),
                                        hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
            attr = 1;
        }
        hipLaunchKernelGGL((up_subtree_kernel<K, KP, W>), dim3(nR * nR, zd * pl.nfilm), dim3(256), lds, s, rho,
                           pl.mult, pl.d_p2mf, pl.d_m2m, lo, R, D, g.xd, g.yd, bufA, pl.films());
        MMS_HIP(hipGetLastError());
    }
    if constexpr (split) {
        // the near field depends on the fork point only (recorded before the
        // upward pass); it is issued after the upward pass so that a captured
        // step graph keeps the multipole chain, the critical path, on the
        // caller's queue and puts the near field on the second one
        if (conc) {
            MMS_HIP(hipStreamWaitEvent(pl.s_near, pl.ev_fork, 0));
            if (launch_near<K, KP, W>(g, pl, rho, phi, scale, true, pl.s_near)) return 1;
            MMS_HIP(hipEventRecord(pl.ev_join, pl.s_near));
        }
    }
    {
        KTimer kt(prof, K_M2M);
        for (int L = R - 1; L > Lc; L--) {
            const int nbp = 1 << L;
            dim3 grid((nbp + 255) / 256, nbp, zd * pl.nfilm);
            hipLaunchKernelGGL((m2m_mfma_kernel<KP>), grid, dim3(256), 0, s, pl.mult + pl.lvl_off[L + 1],
                               pl.mult + pl.lvl_off[L], pl.d_m2m, nbp, pl.films());
            MMS_HIP(hipGetLastError());
        }
        if (Lc >= 2) {
            static int attr = 0;
            if (!attr) {
                MMS_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&up_coarse_kernel<KP>),
                                            hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
                attr = 1;
            }
            hipLaunchKernelGGL((up_coarse_kernel<KP>), dim3(zd * pl.nfilm), dim3(256),
                               sizeof(float) * (256 + 64) * SB, s, pl.mult, pl.d_m2m, lo, Lc + 1, zd, pl.film_stride);
            MMS_HIP(hipGetLastError());
        }
    }
    {
        KTimer kt(prof, K_M2L);
        int st;
        switch (pl.m2l_mode) {
        case MMS_M2L_BF16: st = launch_m2l_nc<KP, MMS_M2L_BF16>(pl, s); break;
        case MMS_M2L_BF16X3: st = launch_m2l_nc<KP, MMS_M2L_BF16X3>(pl, s); break;
        case MMS_M2L_FP16: st = launch_m2l_nc<KP, MMS_M2L_FP16>(pl, s); break;
        case MMS_M2L_FP16X2: st = launch_m2l_nc<KP, MMS_M2L_FP16X2>(pl, s); break;
        default: st = launch_m2l_nc<KP, MMS_M2L_FP32>(pl, s);
        }
        if (st) return 1;
    }
    {
        KTimer kt(prof, K_L2L);
        // subtree roots at levels 3..5 form their ancestors' locals themselves
        const int chain = (R > 2 && R - 1 == Lc) ? 1 : 0;
        if (Lc >= 2 && !chain) {
            static int attr = 0;
            if (!attr) {
                MMS_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&down_coarse_kernel<KP>),
                                            hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
                attr = 1;
            }
            hipLaunchKernelGGL((down_coarse_kernel<KP>), dim3(zd * pl.nfilm), dim3(256),
                               sizeof(float) * (256 + 256) * SB, s, pl.part, pl.film_stride * pl.nfilm, pl.loc,
                               pl.d_l2l, lo, Lc, zd, pl.film_stride);
            MMS_HIP(hipGetLastError());
        }
        for (int L = Lc + 1; L < R; L++)
            if (launch_l2l_sum<KP>(pl, L, s)) return 1;
        // leaf widths 4 and 8: L2P fused into the leaf level, after the near field wrote phi
        constexpr int WF = split ? W : 0;
        const int nbuf = (split && D > 3) ? (1 << (2 * D)) : 64;
        const size_t lds = sizeof(float) * 2 * (size_t)nbuf * SB;
        MMS_CHECK(lds <= 160 * 1024, "FMM downward pass: subtree exceeds LDS");
        static int attr = 0;
        if (!attr) {
            MMS_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&down_subtree_kernel<KP, WF>),
                                        hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
            attr = 1;
        }
        if (conc) MMS_HIP(hipStreamWaitEvent(s, pl.ev_join, 0));
        hipLaunchKernelGGL((down_subtree_kernel<KP, WF>), dim3(nR * nR, zd * pl.nfilm), dim3(256), lds, s, pl.part,
                           pl.film_stride * pl.nfilm, pl.loc, pl.d_l2l, lo, R, D, zd, pl.film_stride, phi, pl.d_l2pf,
                           g.xd, g.yd, scale, chain, pl.films());
        MMS_HIP(hipGetLastError());
    }
    if constexpr (!split) {
        KTimer kt(prof, K_P2P_L2P);
        if (launch_near<K, KP, W>(g, pl, rho, phi, scale, false, s)) return 1;
    }
    return 0;
}

template <int W>
int run_fmm_w(const Grid &g, FmmPlan &pl, const float *rho, float *phi, hipStream_t s, Profiler *prof) {
    switch (pl.P) {
    case 1: return run_fmm<NC<1>::K, W>(g, pl, rho, phi, s, prof);
    case 2: return run_fmm<NC<2>::K, W>(g, pl, rho, phi, s, prof);
    case 3: return run_fmm<NC<3>::K, W>(g, pl, rho, phi, s, prof);
    case 4: return run_fmm<NC<4>::K, W>(g, pl, rho, phi, s, prof);
    case 5: return run_fmm<NC<5>::K, W>(g, pl, rho, phi, s, prof);
    case 6: return run_fmm<NC<6>::K, W>(g, pl, rho, phi, s, prof);
    case 7: return run_fmm<NC<7>::K, W>(g, pl, rho, phi, s, prof);
    case 8: return run_fmm<NC<8>::K, W>(g, pl, rho, phi, s, prof);
    case 9: return run_fmm<NC<9>::K, W>(g, pl, rho, phi, s, prof);
    default: set_error("FMM expansion order must be 1..9"); return 1;
    }
}

unsigned short host_bf16(float x) {
    unsigned u;
    memcpy(&u, &x, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
float host_bf2f(unsigned short h) {
    unsigned u = (unsigned)h << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
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

int FmmPlan::init_batch(const Grid &g, int P_, int w_, int precision_, int cols_, const std::vector<FilmDesc> &fl) {
    MMS_CHECK(!fl.empty(), "empty film batch");
    if (init(g, P_, w_, precision_, cols_, (int)fl.size())) return 1;
    MMS_HIP(hipMalloc(&d_films, sizeof(FilmDesc) * fl.size()));
    MMS_HIP(hipMemcpy(d_films, fl.data(), sizeof(FilmDesc) * fl.size(), hipMemcpyHostToDevice));
    return 0;
}

int FmmPlan::init(const Grid &g, int P_, int w_, int precision_, int cols_, int nfilm_) {
    release();
    // second stream (default priority) and fork/join events for the near field;
    // MMS_GPU_SERIAL=1 keeps every kernel on the caller's stream, in order
    const char *serial = std::getenv("MMS_GPU_SERIAL");
    if ((w_ == 4 || w_ == 8) && !(serial && serial[0] == '1')) {
        MMS_HIP(hipStreamCreateWithFlags(&s_near, hipStreamNonBlocking));
        MMS_HIP(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
        MMS_HIP(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming));
    }
    nfilm = nfilm_ > 0 ? nfilm_ : 1;
    MMS_CHECK(P_ >= 1 && P_ <= 9, "FMM expansion order must be 1..9");
    MMS_CHECK(w_ == 2 || w_ == 4 || w_ == 8, "FMM leaf width must be 2, 4 or 8");
    MMS_CHECK(g.zd >= 1 && g.zd <= 16, "FMM supports 1..16 layers");
    MMS_CHECK(precision_ >= MMS_M2L_FP32 && precision_ <= MMS_M2L_FP16X2_FP16X2NEAR, "unknown M2L precision");
    P = P_; w = w_; precision = precision_; zd = g.zd;
    m2l_mode = (precision == MMS_M2L_BF16X3_FP16NEAR) ? MMS_M2L_BF16X3
               : (precision == MMS_M2L_FP16X2_BF16X3NEAR || precision == MMS_M2L_FP16X2_FP16X2NEAR) ? MMS_M2L_FP16X2
                                                                                                     : precision;
    near_f16x2 = (precision == MMS_M2L_FP16X2_FP16X2NEAR);
    near_f16 = (precision == MMS_M2L_FP16 || precision == MMS_M2L_BF16X3_FP16NEAR || precision == MMS_M2L_FP16X2 ||
                near_f16x2);
    const bool f16_ops = (m2l_mode == MMS_M2L_FP16 || m2l_mode == MMS_M2L_FP16X2);  // equilibrated FP16 operator
    m2l_cols = (cols_ == 8) ? 8 : 4;
    K = fmm_ncoef(P);
    KP = K <= 16 ? 16 : ((K + 31) / 32) * 32;
    const int side = g.xd > g.yd ? g.xd : g.yd;
    Lmax = 0;
    while ((1 << Lmax) < side) Lmax++;
    int lw = 0;
    while ((1 << lw) < w) lw++;
    Lleaf = Lmax - lw;
    MMS_CHECK(Lleaf >= 2, "grid too small for the FMM at this leaf width (needs >= 4 leaf boxes per side)");
    lvl_off.assign(Lleaf + 1, 0);
    long tot = 0;
    for (int L = 2; L <= Lleaf; L++) {
        lvl_off[L] = tot;
        tot += (long)zd * (1L << (2 * L)) * KP;
    }
    film_stride = tot;
    MMS_HIP(hipMalloc(&mult, sizeof(float) * tot * nfilm));
    MMS_HIP(hipMalloc(&loc, sizeof(float) * tot * nfilm));
    MMS_HIP(hipMalloc(&part, sizeof(float) * tot * nfilm * zd));
    MMS_HIP(hipMemset(mult, 0, sizeof(float) * tot * nfilm));
    MMS_HIP(hipMemset(loc, 0, sizeof(float) * tot * nfilm));
    FmmHostOps ops;
    fmm_build_ops(P, w, (double)g.ay, ops);
    auto upload = [&](void **dst, const void *src, size_t bytes) -> int {
        MMS_HIP(hipMalloc(dst, bytes));
        MMS_HIP(hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice));
        return 0;
    };
    {   // P2M as MFMA fragments [KP/16][64][w*w/4]: P2M[16 r + lane % 16][(w*w/4) (lane / 16) + u]
        const int gp = w * w / 4, rtk = KP / 16;
        std::vector<float> pf((size_t)rtk * 64 * gp, 0.0f);
        for (int r = 0; r < rtk; r++)
            for (int lane = 0; lane < 64; lane++)
                for (int u = 0; u < gp; u++) {
                    const int row = 16 * r + lane % 16, col = gp * (lane / 16) + u;
                    if (row < K) pf[((size_t)r * 64 + lane) * gp + u] = (float)ops.p2m[(size_t)row * w * w + col];
                }
        if (upload((void **)&d_p2mf, pf.data(), sizeof(float) * pf.size())) return 1;
    }
    std::vector<float> f(ops.l2p.begin(), ops.l2p.end());                      // [w*w][K]
    if (upload((void **)&d_l2p, f.data(), sizeof(float) * f.size())) return 1;
    // MFMA A-fragments: frag[blk][rt][lane][u] = Op[16 rt + lane % 16][G (lane / 16) + u]
    const int G = KP / 4, RT = KP / 16;
    auto frag = [&](const double *op, std::vector<float> &out, size_t off) {
        for (int rt = 0; rt < RT; rt++)
            for (int lane = 0; lane < 64; lane++)
                for (int u = 0; u < G; u++) {
                    const int row = 16 * rt + lane % 16, col = G * (lane / 16) + u;
                    out[off + ((size_t)rt * 64 + lane) * G + u] =
                        (row < K && col < K) ? (float)op[(size_t)row * K + col] : 0.0f;
                }
    };
    const size_t blk = (size_t)RT * 64 * G;  // floats per KP x KP fragment block (= KP*KP)
    std::vector<float> l2lf(4 * blk), m2mf(4 * blk);
    for (int ci = 0; ci < 4; ci++) {
        frag(&ops.l2l[(size_t)ci * K * K], l2lf, ci * blk);
        frag(&ops.m2m[(size_t)ci * K * K], m2mf, ci * blk);
    }
    if (upload((void **)&d_l2l, l2lf.data(), sizeof(float) * l2lf.size())) return 1;
    if (upload((void **)&d_m2m, m2mf.data(), sizeof(float) * m2mf.size())) return 1;
    if (w == 4 || w == 8) {  // L2P fragments [w*w/16 row tiles][64 lanes][G]: L2P[16 r + lane % 16][G (lane / 16) + u]
        const int rtw = w * w / 16;
        std::vector<float> l2pf((size_t)rtw * 64 * G, 0.0f);
        for (int r = 0; r < rtw; r++)
            for (int lane = 0; lane < 64; lane++)
                for (int u = 0; u < G; u++) {
                    const int row = 16 * r + lane % 16, col = G * (lane / 16) + u;
                    if (col < K) l2pf[((size_t)r * 64 + lane) * G + u] = (float)ops.l2p[(size_t)row * K + col];
                }
        if (upload((void **)&d_l2pf, l2pf.data(), sizeof(float) * l2pf.size())) return 1;
    }
    // M2L fragments per level: [dz][49][blk]; 16-bit operands of 16-slot
    // blocks: offset pairs [dz][child parity][14 pairs][64 lanes][8] (lane
    // (m, g) holds row m, columns 8 (g % 2) .. + 7 of the pair's first block
    // for g < 2, of its second for g >= 2)
    const size_t nblk = (size_t)(2 * zd - 1) * 49;
    const bool pair = (KP == 16 && m2l_mode != MMS_M2L_FP32);
    const size_t pair_elems = (size_t)(2 * zd - 1) * 4 * 14 * 64 * 8;
    const size_t nplane = (m2l_mode == MMS_M2L_BF16X3 || m2l_mode == MMS_M2L_FP16X2) ? 2 : 1;
    m2l_level_bytes = (m2l_mode == MMS_M2L_FP32) ? nblk * blk * sizeof(float)
                                                 : (pair ? pair_elems : nblk * blk) * sizeof(unsigned short) * nplane;
    std::vector<std::vector<double>> pair_T;  // pair form: every level's operator blocks [dz][49][K][K]
    std::vector<char> all(m2l_level_bytes * (Lleaf - 1));
    std::vector<float> lv(nblk * blk);
    // fragment element i = (blk, rt, lane, u) holds Op[16 rt + lane % 16][G (lane / 16) + u]
    auto frow = [&](size_t i) { return (int)(16 * ((i / (64 * G)) % RT) + (i / G) % 16); };
    auto fcol = [&](size_t i) { return (int)(G * ((i / G) % 64 / 16) + i % G); };
    std::vector<float> fp16_levels;  // FP16: every level's FP32 fragments, scaled once all are known
    std::vector<double> fp16_a, fp16_b;  // FP16: the operator's row and column factors
    for (int L = 2; L <= Lleaf; L++) {
        std::vector<double> T = ops.m2l_level(L, Lmax, zd, (double)g.ay, (double)g.az);
        for (size_t b = 0; b < nblk; b++) frag(&T[b * K * K], lv, b * blk);
        char *dst = all.data() + (size_t)(L - 2) * m2l_level_bytes;
        if (pair) pair_T.push_back(T);
        if (m2l_mode == MMS_M2L_FP32) {
            memcpy(dst, lv.data(), m2l_level_bytes);
        } else if (f16_ops) {
            fp16_levels.insert(fp16_levels.end(), lv.begin(), lv.end());
        } else if (!pair) {
            unsigned short *h = reinterpret_cast<unsigned short *>(dst);
            for (size_t i = 0; i < lv.size(); i++) {
                h[i] = host_bf16(lv[i]);
                if (m2l_mode == MMS_M2L_BF16X3) h[lv.size() + i] = host_bf16(lv[i] - host_bf2f(h[i]));
            }
        }
    }
    if (f16_ops) {
        // Row and column factors a_j, b_k (powers of two) that bring the
        // largest |a_j T_jk b_k| of every row and column over all levels,
        // offsets and layer distances near 1 (alternating square-root
        // equilibration), so the operator sits well inside FP16 range.
        std::vector<double> a(KP, 1.0), b(KP, 1.0), mx(KP);
        for (int it = 0; it < 8; it++) {
            std::fill(mx.begin(), mx.end(), 0.0);
            for (size_t i = 0; i < fp16_levels.size(); i++)
                mx[frow(i)] = std::max(mx[frow(i)], std::fabs(fp16_levels[i]) * a[frow(i)] * b[fcol(i)]);
            for (int j = 0; j < KP; j++)
                if (mx[j] > 0) a[j] /= std::sqrt(mx[j]);
            std::fill(mx.begin(), mx.end(), 0.0);
            for (size_t i = 0; i < fp16_levels.size(); i++)
                mx[fcol(i)] = std::max(mx[fcol(i)], std::fabs(fp16_levels[i]) * a[frow(i)] * b[fcol(i)]);
            for (int k = 0; k < KP; k++)
                if (mx[k] > 0) b[k] /= std::sqrt(mx[k]);
        }
        std::vector<float> sc(2 * KP);
        for (int j = 0; j < KP; j++) {
            a[j] = std::ldexp(1.0, (int)std::lround(std::log2(a[j])));
            b[j] = std::ldexp(1.0, (int)std::lround(std::log2(b[j])));
            sc[j] = (float)(1.0 / a[j]);
            sc[KP + j] = (float)(1.0 / b[j]);
        }
        // FP16X2: each level's lo plane (v - hi, FP16) follows its hi plane
        unsigned short *h = reinterpret_cast<unsigned short *>(all.data());
        const size_t per = lv.size(), np = (m2l_mode == MMS_M2L_FP16X2) ? 2 : 1;
        for (size_t i = 0; i < (pair ? 0 : fp16_levels.size()); i++) {
            const float v = (float)(fp16_levels[i] * a[frow(i)] * b[fcol(i)]);
            MMS_CHECK(std::fabs(v) < 65504.0f, "FP16 M2L operator out of range");
            const _Float16 hv = (_Float16)v;
            const size_t o = (i / per) * np * per + i % per;
            memcpy(&h[o], &hv, 2);
            if (np == 2) {
                const _Float16 lv16 = (_Float16)(v - (float)hv);
                memcpy(&h[o + per], &lv16, 2);
            }
        }
        if (upload((void **)&d_m2l_scale, sc.data(), sizeof(float) * sc.size())) return 1;
        fp16_a.assign(a.begin(), a.end());
        fp16_b.assign(b.begin(), b.end());
    }
    if (pair) {
        for (int L = 2; L <= Lleaf; L++) {
            const std::vector<double> &T = pair_T[L - 2];
            unsigned short *h = reinterpret_cast<unsigned short *>(all.data() + (size_t)(L - 2) * m2l_level_bytes);
            for (int dzi = 0; dzi < 2 * zd - 1; dzi++)
                for (int p = 0; p < 4; p++)
                    for (int u = 0; u < 14; u++)
                        for (int lane = 0; lane < 64; lane++)
                            for (int t = 0; t < 8; t++) {
                                const int m = lane % 16, gg = lane / 16, kk = 8 * (gg & 1) + t, idx = 2 * u + (gg >> 1);
                                double v = 0.0;
                                if (idx < 27 && m < K && kk < K) {
                                    const int code = h_m2l_offs[p][idx], oi = (code >> 4) * 7 + (code & 15);
                                    v = T[(((size_t)dzi * 49 + oi) * K + m) * K + kk];
                                    if (f16_ops) v *= fp16_a[m] * fp16_b[kk];
                                }
                                const size_t o = ((((size_t)dzi * 4 + p) * 14 + u) * 64 + lane) * 8 + t;
                                if (f16_ops) {
                                    MMS_CHECK(std::fabs(v) < 65504.0, "FP16 M2L operator out of range");
                                    const _Float16 hv = (_Float16)(float)v;
                                    memcpy(&h[o], &hv, 2);
                                    if (nplane == 2) {
                                        const _Float16 lv16 = (_Float16)((float)v - (float)hv);
                                        memcpy(&h[pair_elems + o], &lv16, 2);
                                    }
                                } else {
                                    h[o] = host_bf16((float)v);
                                    if (nplane == 2) h[pair_elems + o] = host_bf16((float)v - host_bf2f(h[o]));
                                }
                            }
        }
    }
    if (upload((void **)&d_m2l, all.data(), all.size())) return 1;
    // near-field 1/r table over offsets dx, dy in [-(2w-1), 2w-1], dz in [-(zd-1), zd-1]
    near_side = 4 * w - 1;
    std::vector<float> tab((size_t)(2 * zd - 1) * near_side * near_side);
    for (int dz = -(zd - 1); dz <= zd - 1; dz++)
        for (int dy = -(2 * w - 1); dy <= 2 * w - 1; dy++)
            for (int dx = -(2 * w - 1); dx <= 2 * w - 1; dx++) {
                const double ry = dy * (double)g.ay, rz = dz * (double)g.az;
                const double r2 = (double)dx * dx + ry * ry + rz * rz;
                tab[((size_t)(dz + zd - 1) * near_side + (dy + 2 * w - 1)) * near_side + (dx + 2 * w - 1)] =
                    r2 > 0 ? (float)(1.0 / std::sqrt(r2)) : 0.0f;
            }
    if (upload((void **)&d_near, tab.data(), sizeof(float) * tab.size())) return 1;
    if (w == 8) {
        // BF16 split (FP16: plain FP16 in the hi plane) near-field operators
        // [2zd-1][phase 0..2][d index 0..13][hi, lo][64 lanes][8].
        // The 24 x 24 source window of a leaf is read as rows packed back to
        // back, 32 sources per k step: step (a, ph) holds flat sources
        // 96 a + 32 ph + k, i.e. row 4 a + (32 ph + k) / 24, column (32 ph + k) % 24.
        // For target row pair p (rows 2 p + r, r = 0, 1) the operator depends
        // on d = 4 a - 2 p only: lane (m = 8 r + ix, k group gg) holds, for
        // k = 8 gg + t, G(dz, (32 ph + k) / 24 + d - r + 7, (32 ph + k) % 24 - ix + 7).
        const int nd = 2 * zd - 1, ts = near_side;
        std::vector<unsigned short> af((size_t)nd * 42 * 2 * 64 * 8, 0);
        for (int dzi = 0; dzi < nd; dzi++)
            for (int ph = 0; ph < 3; ph++)
                for (int di = 0; di < 14; di++)
                    for (int lane = 0; lane < 64; lane++) {
                        const int m = lane % 16, gg = lane / 16, r = m / 8, ix = m % 8, d = 2 * di - 6;
                        for (int t = 0; t < 8; t++) {
                            const int fl = 32 * ph + 8 * gg + t, row = fl / 24 + d - r + 7, col = fl % 24 - ix + 7;
                            const float v = tab[((size_t)dzi * ts + row) * ts + col];
                            const size_t o = (((size_t)((dzi * 3 + ph) * 14 + di) * 2) * 64 + lane) * 8 + t;
                            if (near_f16) {  // FP16 (lo plane: v - hi in FP16 for the hi/lo split, else unused)
                                const _Float16 hv = (_Float16)v;
                                memcpy(&af[o], &hv, 2);
                                if (near_f16x2) {
                                    const _Float16 lv = (_Float16)(v - (float)hv);
                                    memcpy(&af[o + 64 * 8], &lv, 2);
                                }
                            } else {
                                const unsigned short h = host_bf16(v);
                                af[o] = h;
                                af[o + 64 * 8] = host_bf16(v - host_bf2f(h));
                            }
                        }
                    }
        if (upload((void **)&d_p2pA, af.data(), sizeof(unsigned short) * af.size())) return 1;
    }
    return 0;
}

void FmmPlan::release() {
    void *bufs[] = {mult, loc, part, d_p2pA, d_p2mf, d_l2p, d_l2pf, d_m2m, d_l2l, d_m2l, d_near, d_m2l_scale};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    mult = loc = part = d_p2mf = d_l2p = d_l2pf = d_m2m = d_l2l = d_near = nullptr;
    d_m2l = nullptr;
    d_m2l_scale = nullptr;
    d_p2pA = nullptr;
    m2l_level_bytes = 0;
    if (d_films) (void)hipFree(d_films);
    d_films = nullptr;
    nfilm = 1;
    if (s_near) (void)hipStreamSynchronize(s_near);
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    if (ev_join) (void)hipEventDestroy(ev_join);
    if (s_near) (void)hipStreamDestroy(s_near);
    s_near = nullptr;
    ev_fork = ev_join = nullptr;
}

}  // namespace mms
