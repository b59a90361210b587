// direct_sum.hip -- O(N^2) scalar potential on gfx950.
//
// phi_i = (dV/dx) * sum_{j != i} rho_j / |r_i - r_j|_u   with distances in units
// of dx (r_u = (dxi, dyi*ay, dzi*az)), the computation of the reference's
// calc_potential_exact (fmm_calc.cu:412-461) and calc_potential_exact_kernel
// (potential_calc.cu:18-116).
//
// Work split: a 2D grid of (target block of 256) x (source chunk).  A block
// stages its chunk through LDS in tiles of 256 sources as float4 (x, y, z,
// rho) built from the linear index once per tile; every lane then walks the
// tile with LDS broadcast reads (one ds_read_b128 per source per wave).  Each
// lane accumulates a tile in FP32 and folds the tile sum into an FP64
// accumulator, so summation error grows with the tile, not with N.  Chunk
// partials go to an FP64 scratch [nchunk][n] and a second kernel adds them in
// fixed order (bitwise reproducible run to run).  The self pair is excluded by
// r^2 > 0.  1/r is the hardware reciprocal square root (v_rsq_f32): r^2 of
// two distinct cells is at least min(1, ay^2, az^2) in cell units, never
// denormal, so the denormal scaling of rsqrtf is not needed.
#include <mutex>

#include "mms_internal.hpp"

namespace mms {
namespace {

constexpr int TB = 256;    // targets per block (one per lane)
constexpr int TILE = 256;  // sources per LDS tile

// Chunk partial sums of the entry point that allocates nothing
// (mms_potential_exact_noalloc): 16 MiB per GPU, held until the process ends.
constexpr long kScratchDoubles = 2L * 1024 * 1024;
static __device__ __align__(256) double scratch[kScratchDoubles];

__global__ __launch_bounds__(TB) void direct_partial_kernel(const float *__restrict__ rho, double *__restrict__ part,
                                                            int xd, int yd, int zd, float ay, float az,
                                                            long n, long chunk) {
    __shared__ float4 src[TILE];
    const long t = (long)blockIdx.x * TB + threadIdx.x;
    const long c0 = (long)blockIdx.y * chunk;
    long c1 = c0 + chunk;
    if (c1 > n) c1 = n;
    const long plane = (long)xd * yd;
    float tx = 0.f, ty = 0.f, tz = 0.f;
    if (t < n) {
        tx = (float)(t % xd);
        ty = (float)((t / xd) % yd) * ay;
        tz = (float)(t / plane) * az;
    }
    double acc = 0.0;
    for (long base = c0; base < c1; base += TILE) {
        const long j = base + threadIdx.x;
        float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
        if (j < c1) {
            v.x = (float)(j % xd);
            v.y = (float)((j / xd) % yd) * ay;
            v.z = (float)(j / plane) * az;
            v.w = rho[j];
        }
        __syncthreads();
        src[threadIdx.x] = v;
        __syncthreads();
        const int cnt = (int)((c1 - base) < TILE ? (c1 - base) : TILE);
        float s = 0.f;
#pragma unroll 8
        for (int k = 0; k < cnt; k++) {
            const float4 q = src[k];
            const float ddx = tx - q.x, ddy = ty - q.y, ddz = tz - q.z;
            const float r2 = ddx * ddx + ddy * ddy + ddz * ddz;
            const float inv = (r2 > 0.f) ? __builtin_amdgcn_rsqf(r2) : 0.f;  // r2 >= min(1, ay^2, az^2): never denormal
            s = fmaf(q.w, inv, s);
        }
        acc += (double)s;
    }
    if (t < n) part[(long)blockIdx.y * n + t] = acc;
}

__global__ __launch_bounds__(256) void direct_reduce_kernel(const double *__restrict__ part, float *__restrict__ phi,
                                                            long n, int nchunk, double scale) {
    const long t = (long)blockIdx.x * 256 + threadIdx.x;
    if (t >= n) return;
    double s = 0.0;
    for (int c = 0; c < nchunk; c++) s += part[(long)c * n + t];
    phi[t] = (float)(s * scale);
}

// FP64 probe oracle on the device: a 2D grid of (block of 256 probes) x
// (source chunk); sources are staged through LDS as (x, y, z, rho) once per
// tile and every pair is evaluated in FP64.  Chunk partials [nchunk][np] are
// added in fixed order by probes_f64_reduce_kernel.
__global__ __launch_bounds__(256) void probes_f64_kernel(const float *__restrict__ rho, const int64_t *__restrict__ probes,
                                                         double *__restrict__ part, int np, int xd, int yd, int zd,
                                                         double dx, double dy, double dz, long chunk) {
    __shared__ double4 src[TILE];
    const long n = (long)xd * yd * zd, plane = (long)xd * yd;
    const int p = blockIdx.x * 256 + threadIdx.x;
    double tx = 0.0, ty = 0.0, tz = 0.0;
    if (p < np) {
        const long i = probes[p];
        tx = (double)(i % xd) * dx;
        ty = (double)((i / xd) % yd) * dy;
        tz = (double)(i / plane) * dz;
    }
    const long c0 = (long)blockIdx.y * chunk;
    const long c1 = (c0 + chunk < n) ? c0 + chunk : n;
    double acc = 0.0;
    for (long base = c0; base < c1; base += TILE) {
        const long j = base + threadIdx.x;
        double4 v = make_double4(0.0, 0.0, 0.0, 0.0);
        if (j < c1) {
            v.x = (double)(j % xd) * dx;
            v.y = (double)((j / xd) % yd) * dy;
            v.z = (double)(j / plane) * dz;
            v.w = (double)rho[j];
        }
        __syncthreads();
        src[threadIdx.x] = v;
        __syncthreads();
        const int cnt = (int)((c1 - base) < TILE ? (c1 - base) : TILE);
        for (int k = 0; k < cnt; k++) {
            const double4 q = src[k];
            const double rx = tx - q.x, ry = ty - q.y, rz = tz - q.z;
            const double r2 = rx * rx + ry * ry + rz * rz;
            if (r2 > 0.0) acc += q.w / sqrt(r2);
        }
    }
    if (p < np) part[(long)blockIdx.y * np + p] = acc;
}

__global__ __launch_bounds__(256) void probes_f64_reduce_kernel(const double *__restrict__ part, double *__restrict__ out,
                                                                int np, int nchunk, double dv) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= np) return;
    double s = 0.0;
    for (int c = 0; c < nchunk; c++) s += part[(long)c * np + p];
    out[p] = s * dv;
}

// Batched small films: block b takes work item items[b] = (film, target block)
// and sums every source of its film (LDS tiles, FP32 per tile, FP64 across
// tiles), writing phi directly.
__global__ __launch_bounds__(TB) void direct_batch_kernel(const float *__restrict__ rho, float *__restrict__ phi,
                                                          const int2 *__restrict__ items,
                                                          const BatchFilm *__restrict__ films, float ay, float az,
                                                          double scale) {
    __shared__ float4 src[TILE];
    const int2 it = items[blockIdx.x];
    const BatchFilm f = films[it.x];
    const int xd = f.xd, yd = f.yd;
    const long plane = (long)xd * yd, n = plane * f.zd;
    rho += f.off;
    phi += f.off;
    const long t = (long)it.y * TB + threadIdx.x;
    float tx = 0.f, ty = 0.f, tz = 0.f;
    if (t < n) {
        tx = (float)(t % xd);
        ty = (float)((t / xd) % yd) * ay;
        tz = (float)(t / plane) * az;
    }
    double acc = 0.0;
    for (long base = 0; base < n; base += TILE) {
        const long j = base + threadIdx.x;
        float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
        if (j < n) {
            v.x = (float)(j % xd);
            v.y = (float)((j / xd) % yd) * ay;
            v.z = (float)(j / plane) * az;
            v.w = rho[j];
        }
        __syncthreads();
        src[threadIdx.x] = v;
        __syncthreads();
        const int cnt = (int)((n - base) < TILE ? (n - base) : TILE);
        float s = 0.f;
#pragma unroll 8
        for (int k = 0; k < cnt; k++) {
            const float4 q = src[k];
            const float ddx = tx - q.x, ddy = ty - q.y, ddz = tz - q.z;
            const float r2 = ddx * ddx + ddy * ddy + ddz * ddz;
            const float inv = (r2 > 0.f) ? __builtin_amdgcn_rsqf(r2) : 0.f;  // r2 >= min(1, ay^2, az^2): never denormal
            s = fmaf(q.w, inv, s);
        }
        acc += (double)s;
    }
    if (t < n) phi[t] = (float)(acc * scale);
}

// Batched small films on a 1/r table.  On the grid, 1/r depends only on the
// absolute cell offsets, so the film's table G[|dz|][|dy|][|dx|] (xd*yd*zd
// floats, built on the host in FP64 and rounded once; G[0] = 0 excludes the
// cell itself) and its charges (rows padded to a multiple of 4 with zeros)
// sit in LDS.  Per pair: one absolute difference plus row base (the table
// byte offset, v_sad_u32), one LDS read, one FMA; FP32 over a source row (at
// most 256 sources, four interleaved chains), FP64 across rows.  Two threads
// per target (the even and the odd source rows of every layer), added at the end.  The
// padded sources read table entries past the row (or past the table, into
// its zeroed padding and the charges), all finite, times a zero charge.
constexpr int TT = 2 * TB;  // threads of the table kernel
__global__ __launch_bounds__(TT) void direct_table_kernel(const float *__restrict__ rho, float *__restrict__ phi,
                                                          const float *__restrict__ gtab,
                                                          const int2 *__restrict__ items,
                                                          const BatchFilm *__restrict__ films, double scale) {
    extern __shared__ __attribute__((aligned(16))) float dlds[];
    __shared__ double half1[TB];
    const int2 it = items[blockIdx.x];
    const BatchFilm f = films[it.x];
    const int xd = f.xd, yd = f.yd, zd = f.zd, xdp = (xd + 3) & ~3;
    const int plane = xd * yd, n = plane * zd, n4 = (n + 3) & ~3;
    float *tab = dlds;       // [zd][yd][xd]
    float *q = dlds + n4;    // [zd][yd][xdp]
    rho += f.off;
    phi += f.off;
    gtab += f.off;
    for (int e = threadIdx.x; e < n4; e += TT) tab[e] = e < n ? gtab[e] : 0.0f;  // padding read by padded sources
    for (int e = threadIdx.x; e < zd * yd * xdp; e += TT) {
        const int xs = e % xdp, r = e / xdp;
        q[e] = xs < xd ? rho[(long)r * xd + xs] : 0.0f;
    }
    __syncthreads();
    const int h = threadIdx.x / TB, t = it.y * TB + (threadIdx.x % TB);
    double acc = 0.0;
    if (t < n) {
        const int xt = t % xd, yt = (t / xd) % yd, zt = t / plane;
        const char *tb = reinterpret_cast<const char *>(tab);
        const unsigned x4t = 4u * (unsigned)xt;
        const int ng = xdp >> 2;  // float4 groups of a row
        for (int zs = 0; zs < zd; zs++) {
            const unsigned bz = 4u * (unsigned)((zt > zs ? zt - zs : zs - zt) * plane);
            for (int ys = h; ys < yd; ys += 2) {
                // byte offset of the table row; a source's entry is |4 xt - 4 xs| further
                const unsigned rb = __umul24((unsigned)(yt > ys ? yt - ys : ys - yt), 4u * (unsigned)xd) + bz;
                const float4 *qr = reinterpret_cast<const float4 *>(q + (zs * yd + ys) * xdp);
                for (int g0 = 0; g0 < ng; g0 += 64) {  // FP32 over at most 256 sources
                    const int g1 = (g0 + 64 < ng) ? g0 + 64 : ng;
                    float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f, s3 = 0.0f;  // one chain per column mod 4
                    for (int g = g0; g < g1; g++) {
                        const float4 c = qr[g];
                        const unsigned x4s = 16u * (unsigned)g;
                        s0 = fmaf(c.x, *reinterpret_cast<const float *>(tb + (max(x4t, x4s) - min(x4t, x4s) + rb)), s0);
                        s1 = fmaf(c.y, *reinterpret_cast<const float *>(tb + (max(x4t, x4s + 4) - min(x4t, x4s + 4) + rb)), s1);
                        s2 = fmaf(c.z, *reinterpret_cast<const float *>(tb + (max(x4t, x4s + 8) - min(x4t, x4s + 8) + rb)), s2);
                        s3 = fmaf(c.w, *reinterpret_cast<const float *>(tb + (max(x4t, x4s + 12) - min(x4t, x4s + 12) + rb)), s3);
                    }
                    acc += (double)((s0 + s1) + (s2 + s3));
                }
            }
        }
    }
    if (h == 1) half1[threadIdx.x - TB] = acc;
    __syncthreads();
    if (h == 0 && t < n) phi[t] = (float)((acc + half1[threadIdx.x]) * scale);
}

// The summation order and float arithmetic of the reference's CUDA kernel
// calc_potential_exact_kernel (potential_calc.cu:18-116), for callers that
// want its results bit for bit:
//   - 1024 partial sums per target: partial t adds the sources
//     [t * tstride, (t + 1) * tstride), tstride = ceilf(N / 1024.f), in
//     index order, each term (charge_i * dV) / sqrtf(Rx^2 + Ry^2 + Rz^2)
//     with R = ((x - x_) dx, (y - y_) dy, (z - z_) dz), dV = dx dy dz, the
//     self term skipped (potential_calc.cu:57-75);
//   - the partials added by the halving tree s[t] += s[t + h], h = 512 .. 1
//     (potential_calc.cu:82-111).
// Every operation is rounded on its own (no fused multiply-add), with
// IEEE division and square root.  A workgroup is the reference's 1024
// virtual threads applied to RT targets at once: each source is loaded and
// its coordinates decoded once for the RT targets, and the RT reduction
// trees run together in LDS.
constexpr int RT = 8;

__global__ __launch_bounds__(1024) void direct_refgpu_kernel(const float *__restrict__ rho, float *__restrict__ phi,
                                                             int xd, int yd, int n, float dx, float dy, float dz,
                                                             int tstride) {
#pragma clang fp contract(off)
    __shared__ float sdata[RT][1024];
    const int t = threadIdx.x;
    const int b0 = blockIdx.x * RT;
    const int plane = xd * yd;
    int tx[RT], ty[RT], tz[RT];
#pragma unroll
    for (int k = 0; k < RT; k++) {
        const int bi = (b0 + k < n) ? b0 + k : 0;
        tx[k] = bi % xd;
        ty[k] = ((bi - tx[k]) / xd) % yd;
        tz[k] = (bi - tx[k] - ty[k] * xd) / plane;
    }
    float pot[RT];
#pragma unroll
    for (int k = 0; k < RT; k++) pot[k] = 0.0f;
    const int ti1 = t * tstride;
    if (ti1 < n) {
        const float dV = dx * dy * dz;
        const int i1 = (ti1 + tstride < n) ? ti1 + tstride : n;
        for (int i = ti1; i < i1; i++) {
            const float q = rho[i] * dV;
            const int x_ = i % xd;
            const int y_ = ((i - x_) / xd) % yd;
            const int z_ = (i - x_ - y_ * xd) / plane;
#pragma unroll
            for (int k = 0; k < RT; k++) {
                if (b0 + k == i) continue;
                const float Rx = (float)(tx[k] - x_) * dx, Ry = (float)(ty[k] - y_) * dy, Rz = (float)(tz[k] - z_) * dz;
                const float r = sqrtf(Rx * Rx + Ry * Ry + Rz * Rz);
                pot[k] += q / r;
            }
        }
    }
#pragma unroll
    for (int k = 0; k < RT; k++) sdata[k][t] = pot[k];
    __syncthreads();
    for (int h = 512; h > 0; h >>= 1) {
        if (t < h) {
#pragma unroll
            for (int k = 0; k < RT; k++) sdata[k][t] += sdata[k][t + h];
        }
        __syncthreads();
    }
    if (t < RT && b0 + t < n) phi[b0 + t] = sdata[t][0];
}

}  // namespace

int direct_refgpu_potential(const Grid &g, const float *rho, float *phi, hipStream_t s, Profiler *prof) {
    KTimer kt(prof, K_DIRECT);
    const long n = g.n();
    MMS_CHECK(n < (1L << 30), "reference-order direct sum: grid too large");
    // tstride as the reference computes it: ceilf(N / (float)blockDim.x)
    const int tstride = (int)ceilf((float)n / 1024.0f);
    hipLaunchKernelGGL(direct_refgpu_kernel, dim3((unsigned)((n + RT - 1) / RT)), dim3(1024), 0, s, rho, phi, g.xd,
                       g.yd, (int)n, g.dx, g.dy, g.dz, tstride);
    MMS_HIP(hipGetLastError());
    return 0;
}

int direct_batch_potential(const float *rho, float *phi, const int2 *items, int nitems, const BatchFilm *films,
                           float ay, float az, double scale, const float *tab, size_t tab_lds, hipStream_t s,
                           Profiler *prof) {
    KTimer kt(prof, K_DIRECT);
    if (nitems <= 0) return 0;
    if (tab) {
        MMS_CHECK(tab_lds <= kDirectTableLds, "direct sum: film table exceeds LDS");
        static bool attr = false;
        if (!attr) {
            MMS_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&direct_table_kernel),
                                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)kDirectTableLds));
            attr = true;
        }
        hipLaunchKernelGGL(direct_table_kernel, dim3(nitems), dim3(TT), tab_lds, s, rho, phi, tab, items, films, scale);
    } else {
        hipLaunchKernelGGL(direct_batch_kernel, dim3(nitems), dim3(TB), 0, s, rho, phi, items, films, ay, az, scale);
    }
    MMS_HIP(hipGetLastError());
    return 0;
}

size_t direct_table_lds(int xd, int yd, int zd) {
    const long n = (long)xd * yd * zd;
    return sizeof(float) * (size_t)(((n + 3) & ~3L) + (long)zd * yd * ((xd + 3) & ~3));
}

// Source chunks of the direct sum for n cells: aim for >= 1024 blocks in
// total (4 per CU) without chunks below one tile.
static void direct_chunks(long n, int split, long *chunk, int *nchunk) {
    const long tblocks = (n + TB - 1) / TB;
    long want = (1024L * split + tblocks - 1) / tblocks;
    long maxchunks = (n + TILE - 1) / TILE;
    if (want > maxchunks) want = maxchunks;
    if (want < 1) want = 1;
    long c = (n + want - 1) / want;
    c = ((c + TILE - 1) / TILE) * TILE;
    *chunk = c;
    *nchunk = (int)((n + c - 1) / c);
}

int DirectPlan::init(const Grid &g, int split_) {
    release();
    split = (split_ == 2 || split_ == 4) ? split_ : 1;
    const long n = g.n();
    direct_chunks(n, split, &chunk, &nchunk);
    MMS_HIP(hipMalloc(&part, sizeof(double) * (size_t)n * nchunk));
    return 0;
}

void DirectPlan::release() {
    if (part) (void)hipFree(part);
    part = nullptr;
    nchunk = 0;
}

// the two kernels of the direct sum with the chunk partials in `part`
static int direct_launch(const Grid &g, long chunk, int nchunk, double *part, const float *rho, float *phi,
                         hipStream_t s, Profiler *prof) {
    const long n = g.n();
    {
        KTimer kt(prof, K_DIRECT);
        dim3 grid((unsigned)((n + TB - 1) / TB), (unsigned)nchunk);
        hipLaunchKernelGGL(direct_partial_kernel, grid, dim3(TB), 0, s, rho, part, g.xd, g.yd, g.zd, g.ay, g.az, n,
                           chunk);
        MMS_HIP(hipGetLastError());
    }
    {
        KTimer kt(prof, K_DIRECT_REDUCE);
        const double scale = (double)(g.dx * g.dy * g.dz) / (double)g.dx;
        hipLaunchKernelGGL(direct_reduce_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, part, phi, n,
                           nchunk, scale);
        MMS_HIP(hipGetLastError());
    }
    return 0;
}

int direct_potential(const Grid &g, DirectPlan &plan, const float *rho, float *phi, hipStream_t s,
                     Profiler *prof) {
    MMS_CHECK(plan.part != nullptr, "direct solver plan not initialised");
    return direct_launch(g, plan.chunk, plan.nchunk, plan.part, rho, phi, s, prof);
}

// Device address of this GPU's `scratch`, looked up once per device with
// that device current (the host-side name of a __device__ array is not a
// device address).
static int scratch_ptr(double **out) {
    static std::mutex mu;
    static double *ptr[64] = {nullptr};
    int dev = 0;
    MMS_HIP(hipGetDevice(&dev));
    MMS_CHECK(dev >= 0 && dev < 64, "device index out of range");
    std::lock_guard<std::mutex> lk(mu);
    if (!ptr[dev]) {
        void *p = nullptr;
        MMS_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        MMS_CHECK(p != nullptr, "no device address for the direct-sum scratch");
        ptr[dev] = static_cast<double *>(p);
    }
    *out = ptr[dev];
    return 0;
}

int direct_probes_f64(const Grid &g, const float *rho, const int64_t *probes, int np, double *out,
                      hipStream_t s, Profiler *prof) {
    if (np <= 0) return 0;
    const long n = g.n();
    const int nbp = (np + 255) / 256;
    int nchunk = 2048 / nbp;
    if (nchunk < 1) nchunk = 1;
    if (nchunk > 64) nchunk = 64;
    long chunk = (n + nchunk - 1) / nchunk;
    chunk = ((chunk + TILE - 1) / TILE) * TILE;
    nchunk = (int)((n + chunk - 1) / chunk);
    double *part = nullptr;
    MMS_HIP(hipMalloc(&part, sizeof(double) * (size_t)nchunk * np));
    int st = 0;
    {
        KTimer kt(prof, K_PROBES);
        hipLaunchKernelGGL(probes_f64_kernel, dim3(nbp, nchunk), dim3(256), 0, s, rho, probes, part, np, g.xd, g.yd,
                           g.zd, (double)g.dx, (double)g.dy, (double)g.dz, chunk);
        if (hipGetLastError() != hipSuccess) st = 1;
        if (!st) {
            hipLaunchKernelGGL(probes_f64_reduce_kernel, dim3(nbp), dim3(256), 0, s, part, out, np, nchunk,
                               (double)g.dx * (double)g.dy * (double)g.dz);
            if (hipGetLastError() != hipSuccess) st = 1;
        }
    }
    if (hipStreamSynchronize(s) != hipSuccess) st = 1;
    (void)hipFree(part);
    MMS_CHECK(st == 0, "probe kernel launch failed");
    return 0;
}

}  // namespace mms

// Direct-sum potential from device arrays without any allocation: the chunk
// partial sums live in the file-scope `scratch` array of each GPU; the
// kernels and the chunking are those of calc_potential_exact, so the results
// are the same bit for bit.  Calls that overlap on one GPU share `scratch`
// and race.  Every partial is written before it is read, so nothing has to
// be cleared between calls.
extern "C" int mms_potential_exact_noalloc(const float *charge_dev, float *potential_dev, int xdim, int ydim,
                                           int zdim, float dx, float dy, float dz, void *stream) {
    using namespace mms;
    MMS_CHECK(charge_dev && potential_dev, "null device array");
    MMS_CHECK(xdim > 0 && ydim > 0 && zdim > 0, "grid dimensions must be positive");
    MMS_CHECK(dx > 0 && dy > 0 && dz > 0, "cell sizes must be positive");
    Grid g;
    g.xd = xdim; g.yd = ydim; g.zd = zdim;
    g.dx = dx; g.dy = dy; g.dz = dz;
    g.ay = dy / dx; g.az = dz / dx;
    const long n = g.n();
    long chunk = 0;
    int nchunk = 0;
    direct_chunks(n, 1, &chunk, &nchunk);
    if (n * nchunk > kScratchDoubles) {
        set_error("mms_potential_exact_noalloc: grid of " + std::to_string(n) + " cells needs " +
                  std::to_string(n * nchunk) + " partial sums; the scratch holds " + std::to_string(kScratchDoubles) +
                  " (at most 2,097,152 cells)");
        return 1;
    }
    double *part = nullptr;
    if (scratch_ptr(&part)) return 1;
    return direct_launch(g, chunk, nchunk, part, charge_dev, potential_dev, static_cast<hipStream_t>(stream),
                         nullptr);
}
