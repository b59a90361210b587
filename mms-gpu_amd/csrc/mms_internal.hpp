// mms_internal.hpp -- shared declarations of the gfx950 library (not installed).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <vector>
#include <map>

#include "../../include/mms_gpu.h"

namespace mms {

// ---------------------------------------------------------------- errors --
void set_error(const std::string &msg);
#define MMS_HIP(call)                                                            \
    do {                                                                         \
        hipError_t e_ = (call);                                                  \
        if (e_ != hipSuccess) {                                                  \
            ::mms::set_error(std::string(#call) + ": " + hipGetErrorString(e_)); \
            return 1;                                                            \
        }                                                                        \
    } while (0)
#define MMS_CHECK(cond, msg)           \
    do {                                \
        if (!(cond)) {                  \
            ::mms::set_error(msg);      \
            return 1;                   \
        }                               \
    } while (0)

// ------------------------------------------------------------- geometry --
// Cell (x,y,z) at linear index z*ydim*xdim + y*xdim + x.  Distances are
// measured in units of dx; ay = dy/dx, az = dz/dx.
struct Grid {
    int xd, yd, zd;
    float dx, dy, dz;
    float ay, az;
    long n() const { return (long)xd * yd * zd; }
    long plane() const { return (long)xd * yd; }
};

// ------------------------------------------------------------- profiling --
// Kernel classes timed with hipEvents when profiling is on.
enum KClass {
    K_DIV = 0, K_STAGE, K_DIRECT, K_DIRECT_REDUCE, K_P2M, K_M2M, K_M2L, K_L2L, K_P2P_L2P, K_L2P,
    K_HFIELD, K_ENERGY, K_PROBES, K_MISC, K_SURROGATE, K_COUNT
};
const char *kclass_name(int k);

struct Profiler {
    bool on = false;
    hipStream_t stream = nullptr;
    struct Pending { int k; hipEvent_t a, b; };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> pool;
    double ms[K_COUNT] = {0};
    long count[K_COUNT] = {0};
    int begin(int k, hipEvent_t *a, hipEvent_t *b);
    int end_record(int k, hipEvent_t a, hipEvent_t b);
    int flush();
    void reset();
    ~Profiler();
};

// Helper that brackets a launch with profiler events (no-op when off).
struct KTimer {
    Profiler *p; int k; hipEvent_t a = nullptr, b = nullptr;
    KTimer(Profiler *p_, int k_) : p(p_), k(k_) { if (p && p->on) p->begin(k, &a, &b); }
    ~KTimer() { if (p && p->on && a) p->end_record(k, a, b); }
};

// --------------------------------------------------------- direct solver --
struct DirectPlan {
    int nchunk = 0;      // source chunks
    long chunk = 0;      // sources per chunk
    double *part = nullptr;  // [nchunk][n] partial sums
    int split = 1;           // multiplier on the number of source chunks
    int init(const Grid &g, int split = 1);
    void release();
};
int direct_potential(const Grid &g, DirectPlan &plan, const float *rho, float *phi,
                     hipStream_t s, Profiler *prof);
// the float arithmetic and summation order of the reference's CUDA direct-sum kernel
int direct_refgpu_potential(const Grid &g, const float *rho, float *phi, hipStream_t s, Profiler *prof);
int direct_probes_f64(const Grid &g, const float *rho, const int64_t *probes, int np,
                      double *out, hipStream_t s, Profiler *prof);

// ------------------------------------------------------------ FMM solver --
// Regular-grid FMM: per-layer quadtrees over a zero-padded 2^Lmax square,
// leaf boxes of w x w cells at level Lleaf; expansions are truncated at order
// P and keep the (n+m even) coefficients only (in-plane sources and targets),
// stored as K = (P+1)(P+2)/2 reals per box and layer.
// A batch of films sharing one padded FMM size class: film f has xd x yd cells
// (<= the plan's grid) starting at cell `off` of the packed rho/phi arrays.
struct FilmDesc { int xd, yd; long off; };
struct FilmSet { const FilmDesc *desc; long stride; int zd; };  // desc null: single film
struct FmmPlan {
    int P = 0, K = 0, KP = 0, w = 4, Lmax = 0, Lleaf = 0, zd = 0;
    int precision = MMS_M2L_FP32;  // requested form (mms_m2l_precision)
    int m2l_mode = MMS_M2L_FP32;   // M2L operands: FP32, BF16, BF16X3 or FP16
    bool near_f16 = false;         // leaf-8 near field in scaled FP16 (else BF16 split / FP32)
    bool near_f16x2 = false;       // ... with the FP16 hi/lo split of its operator (two products)
    int m2l_cols = 4;        // MFMA column tiles (16 targets each) per wave in M2L: 4 or 8
    // device buffers; box vectors have KP slots (K rounded up to 32, zero padded)
    float *mult = nullptr;   // all levels, [level][z][by][bx][KP]
    float *loc = nullptr;    // all levels, same layout
    float *part = nullptr;   // M2L partial sums per source layer: [zs][film][level layout of loc]
    std::vector<long> lvl_off;  // float offset of each level in mult/loc
    float *d_p2mf = nullptr; // P2M as MFMA fragments [KP/16][64][w*w/4]
    float *d_l2p = nullptr;  // [w*w][K]
    float *d_l2pf = nullptr; // leaf width 4: L2P as MFMA fragments [64][KP/4]
    float *d_m2m = nullptr;  // MFMA fragments [4 child positions][KP/16][64][KP/4]
    float *d_l2l = nullptr;  // MFMA fragments [4 child positions][KP/16][64][KP/4]
    float *d_m2l_scale = nullptr;  // FP16 M2L: 1/row scale [KP], 1/column scale [KP] (powers of two)
    void *d_m2l = nullptr;   // MFMA fragments per level [2zd-1][49 offsets][KP/16][64][KP/4]
                             // (float, or bf16 hi [+ bf16 lo] planes)
    size_t m2l_level_bytes = 0;
    float *d_near = nullptr; // near-field kernel table 1/r over offsets
    int near_side = 0;       // 4w-1
    unsigned short *d_p2pA = nullptr;  // leaf width 8: BF16 hi/lo near-field operators [2zd-1][3][14][2][64][8]
    int nfilm = 1;           // films in the batch (1 = single film)
    long film_stride = 0;    // floats per film in mult/loc
    FilmDesc *d_films = nullptr;
    hipStream_t s_near = nullptr;         // near field runs here, concurrently with the multipole passes
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    FilmSet films() const { return FilmSet{d_films, film_stride, zd}; }
    int init(const Grid &g, int P, int w, int precision, int cols = 8, int nfilm = 1);
    int init_batch(const Grid &g, int P, int w, int precision, int cols, const std::vector<FilmDesc> &films);
    void release();
};
int fmm_potential(const Grid &g, FmmPlan &plan, const float *rho, float *phi,
                  hipStream_t s, Profiler *prof);

// Host-side operator construction (FP64), real bases as described above.
struct FmmHostOps {
    int P, K, w;
    std::vector<double> p2m;   // [K][w*w]
    std::vector<double> l2p;   // [w*w][K]
    std::vector<double> m2m;   // [4][K][K]
    std::vector<double> l2l;   // [4][K][K]
    // M2L for one level: [2zd-1][49][K][K] (offset index (oy+3)*7 + (ox+3))
    std::vector<double> m2l_level(int level, int Lmax, int zd, double ay, double az) const;
};
int fmm_build_ops(int P, int w, double ay, FmmHostOps &ops);
int fmm_ncoef(int P);

// ------------------------------------------------- learned surrogate --
// Plan of the charge -> potential network (surrogate.hip): FP32 parameters in
// device layout with their BF16 copies, Fourier bases, activation and
// transform buffers; with `train` also the saved pre-activations, gradient
// buffers and the flipped, transposed conv taps of the backward pass.  The
// grid's zd layers are zd / zf films of zf layers each.
struct SurrogatePlan {
    static constexpr int kPartBlocks = 256;  // workgroups of the weight-gradient kernels
    // loss / projection partials: per film kHeadBlocks blocks of kHeadPart
    // floats (64 projection-weight gradients, the bias gradient, the loss)
    static constexpr int kHeadBlocks = 64, kHeadPart = 66;
    Grid g{};
    int zf = 1;
    int nmode = 0;
    int xpad = 0;  // row length of the zero-padded Bx (d_bx [32][xpad])
    bool train = false;
    float *d_param = nullptr;
    unsigned short *d_lift = nullptr, *d_wm = nullptr, *d_bx = nullptr, *d_by = nullptr, *d_bxt = nullptr;
    unsigned short *d_w[3] = {nullptr, nullptr, nullptr}, *d_wt[3] = {nullptr, nullptr, nullptr};
    unsigned short *d_h0 = nullptr, *d_h1 = nullptr, *d_h2 = nullptr;  // NHWC BF16
    unsigned short *d_x2t = nullptr;  // spectral X2^T [z][y][64][32] BF16
    unsigned short *d_z[4] = {nullptr, nullptr, nullptr, nullptr};                       // pre-activations
    unsigned short *d_dza = nullptr, *d_dzb = nullptr;                                     // gradients
    float *d_x1 = nullptr, *d_hm = nullptr, *d_gm = nullptr, *d_x2 = nullptr, *d_dh = nullptr;
    float *d_wpart = nullptr, *d_hpart = nullptr, *d_hsum = nullptr;
    double *d_part = nullptr;
    int init(const Grid &g, int zf, const float *weights, bool train, hipStream_t s);
    int pack(hipStream_t s);
    void release();
};
long surrogate_weight_count();
long surrogate_param_count();
void surrogate_api_to_param(const float *api, float *param);
void surrogate_param_to_api(const float *param, float *api);
// forward pass (with sp.train the pre-activations are kept for surrogate_backward)
int surrogate_potential(SurrogatePlan &sp, const float *rho, float *phi, float rscale, float oscale, hipStream_t s,
                        Profiler *prof);
// per-film losses sum (phi_s - phi_t)^2 inv_norm[f] into the head partials
// (sp.d_hpart[f][kHeadBlocks][kHeadPart], last entry); with grad also the projection's gradient
int surrogate_losses(SurrogatePlan &sp, const float *phi_s, const float *phi_t, const float *inv_norm, float gscale,
                     float oscale, bool grad, hipStream_t s);
// gradient (device layout) of gscale / 2 * sum_f loss_f after surrogate_potential on the same rho;
// sp.d_hsum[kHeadPart - 1] receives sum_f loss_f
int surrogate_backward(SurrogatePlan &sp, const float *rho, const float *phi_s, const float *phi_t,
                       const float *inv_norm, float gscale, float rscale, float oscale, float *grad, hipStream_t s);
// Adam step `step` (from 1) at rate lr on sp.d_param with moments m, v; refreshes the BF16 copies
int surrogate_adam(SurrogatePlan &sp, float *m, float *v, const float *grad, long step, float lr, hipStream_t s);
// sum of x^2 over n values (synchronous)
int surrogate_sqnorm(SurrogatePlan &sp, const float *x, long n, double *out, hipStream_t s);
// relative L2 residual ||a - b|| / ||b|| (r null: partials only, left on the device)
int surrogate_residual(SurrogatePlan &sp, const float *a, const float *b, long n, double *r, hipStream_t s);

// -------------------------------------------------------- stencil / LLG --
// Per-film record of a batch for the stencil/stage kernels (16-byte aligned).
struct BatchFilm { int xd, yd, zd, pad; long off; float hx, hy, hz, pad2; };
struct Material {
    float mu_0, Ms, Aexch, alfa, gamma;
    int demag, exchange, external;
    float hext[3];
};
int launch_div(const Grid &g, const float *m, float *rho, hipStream_t s, Profiler *prof);
// stage in 0..3; m_in/m_out FP32 SoA [3][n]; mn FP64 SoA [3][n]; acc (slope) FP32 SoA [3][n];
// t_dev: device double
int launch_stage(const Grid &g, const Material &mat, int stage, double dt, const float *phi,
                 const float *m_in, float *m_out, double *mn, const float *m_base, float *acc,
                 const uint8_t *matl, const double *t_dev, hipStream_t s, Profiler *prof);
// H (AoS float) at state m (SoA float); phi may be null if !demag
int launch_hfield(const Grid &g, const Material &mat, const float *phi, const float *m,
                  const double *t_dev, float *H_aos, hipStream_t s, Profiler *prof);
// Block partials of sum(M.H) and max |M x H| over material cells; H (AoS)
// must have been computed by launch_hfield.  partial: [nblk][2] doubles.
int launch_energy_h(const float *m, const float *h_aos, const uint8_t *matl, long n,
                    double *partial, int nblk, hipStream_t s, Profiler *prof);
// Stage kernel with the rk4_step drop-in's extra inputs/outputs (AoS):
// h_given replaces the computed field; h_out/k_out/slope_out receive the field,
// dM/dt and the float slope sum.  Null pointers disable each.
// m_base: FP32 copy of M_n that stays intact through stages 0-2 (the stage-0
// input); null makes every stage read the FP64 base.  m_keep: FP32 SoA values
// written to the non-material cells of the stage output (null: the input's);
// slope_out is 0 on non-material cells, k_out is not written there.
struct StageExtra { const float *h_given; float *h_out; float *k_out; float *slope_out; int normalize; double *mn_out;
                    const float *m_base; const float *m_keep; };
int launch_stage_ex(const Grid &g, const Material &mat, int stage, double dt, const float *phi,
                    const float *m_in, float *m_out, double *mn, float *acc, const uint8_t *matl,
                    const double *t_dev, const StageExtra &ex, hipStream_t s, Profiler *prof);
// Block partials [nblk][5] of (sum Mx, My, Mz, |M|, count) over material cells
int launch_mavg(const float *m, const uint8_t *matl, long n, double *partial, int nblk, hipStream_t s);
int launch_soa_from_aos(const float *aos, float *soa32, double *soa64, long n, hipStream_t s);
// Batched films (tile list of (film, first column): 256 consecutive (x, y)
// columns of a film in row-major order; component planes pstride cells apart;
// per-film constant external field)
constexpr int STENCIL_TILE_COLUMNS = 256;
int launch_div_batch(const float *m, float *rho, const int4 *tiles, int ntiles, const BatchFilm *films, long pstride,
                     float dx, float dy, float dz, int zd, hipStream_t s, Profiler *prof);
int launch_stage_batch(const Grid &g, const Material &mat, int stage, double dt, const float *phi, const float *m_in,
                       float *m_out, double *mn, const float *m_base, float *acc, const uint8_t *matl,
                       const int4 *tiles, int ntiles, const BatchFilm *films, long pstride, hipStream_t s,
                       Profiler *prof);
// tab (or null: per-pair 1/r): every direct film's 1/r table G[|dz|][|dy|][|dx|] at the film's cell
// offset; tab_lds: the largest direct film's table + padded charges (direct_table_lds) in bytes
int direct_batch_potential(const float *rho, float *phi, const int2 *items, int nitems, const BatchFilm *films,
                           float ay, float az, double scale, const float *tab, size_t tab_lds, hipStream_t s,
                           Profiler *prof);
constexpr size_t kDirectTableLds = 96 * 1024;  // films whose table and charges fit take the table kernel
size_t direct_table_lds(int xd, int yd, int zd);
int launch_aos_from_soa(const float *soa32, float *aos, long n, hipStream_t s);
int launch_add_time(double *t_dev, double dt, hipStream_t s);
// zero the FP32 SoA planes [3][n] of m on non-material cells
int launch_zero_nonmagnetic(float *m, const uint8_t *matl, long n, hipStream_t s);
int launch_fill_f32(float *p, float v, long n, hipStream_t s);

}  // namespace mms

// --------------------------------------------------------------- context --
struct mms_ctx {
    mms_params p;
    mms::Grid g;
    mms::Material mat;
    hipStream_t stream = nullptr;
    int device = 0;
    // device state (SoA planes [3][n])
    float *m32[3] = {nullptr, nullptr, nullptr};  // FP32 stage magnetisation slots
    double *mn64 = nullptr;    // FP64 base M_n
    double *mnew64 = nullptr;  // FP64 proposal M_{n+1} (adaptive time loop)
    float *acc32 = nullptr;    // FP32 slope accumulator (k1 + 2 k2 + 2 k3 + k4)
    float *rho = nullptr, *phi = nullptr;
    uint8_t *matl = nullptr;
    double *t_dev = nullptr;
    float *h_aos = nullptr;    // field output / staging, AoS [n][3]
    float *aux_aos = nullptr;  // staging, 3 AoS arrays
    double *red = nullptr;     // reduction scratch [red_blocks][5]
    int red_blocks = 0;
    int cur = 0;               // slot of m32 holding the current state
    double t_host = 0.0;
    std::vector<unsigned char> host_matl;
    bool all_magnetic = true;       // every cell magnetic: stage kernels skip the material bytes
    mms::DirectPlan direct;
    mms::FmmPlan fmm;
    mms::Profiler prof;
    // learned surrogate: 0 off, 1 at RK4 stages 2-3 always, 2 at stages 2-3
    // when its residual against the solver at stage 1 is within sur_tol
    mms::SurrogatePlan sur;
    int sur_mode = 0;
    float sur_tol = 1e-3f, sur_rscale = 1.0f, sur_oscale = 1.0f;
    float *phi_s = nullptr;       // surrogate potential of the stage-1 judge
    long sur_judged = 0, sur_accepted = 0;
    double sur_last = -1.0;
    // graph of one RK4 step
    hipGraphExec_t graph = nullptr;
    float graph_dt = 0.0f;
    int potential(const float *rho_d, float *phi_d);
    int one_step(float dt);
};
