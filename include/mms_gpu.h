/*
 * mms_gpu.h -- C ABI of the MI355X (gfx950) field-evaluation + LLG RK4 library
 * (libmmsgpu.so).
 *
 * Part 1 declares drop-in replacements for the reference's own entry points,
 * with the same argument meaning and 0/1 (EXIT_SUCCESS/EXIT_FAILURE) return
 * convention.  Each cites the reference declaration it replaces.  Differences a
 * caller can observe:
 *   - plain C linkage; Vector3 arrays are passed as `float *` holding x,y,z
 *     per cell (the 12-byte AoS layout of the reference's struct Vector3);
 *   - all computation runs on the GPU whatever `use_gpu` says (the argument is
 *     accepted for signature compatibility); there is no CPU path;
 *   - errors (HIP failures, bad shapes) return EXIT_FAILURE with a message in
 *     mms_last_error() instead of calling exit();
 *   - device buffers are cached per problem shape inside the library instead of
 *     function-level statics sized by the first call, so any sequence of shapes
 *     is accepted;
 *   - fmm_calc accepts any xdim, ydim (zero-padded to a power-of-two square)
 *     and expansion orders P = 1..9 (the reference asserts P <= 4).
 *
 * Part 2 is the resident API: an explicit context that owns the state on the
 * device, runs fixed-step RK4 without host round trips, and selects the
 * potential solver and the M2L precision.
 *
 * Thread safety: calls on one context must not overlap; distinct contexts may
 * be used from distinct host threads.  The Part-1 entry points share an
 * internal cache and are serialised by a mutex.
 */
#ifndef MMS_GPU_H
#define MMS_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* Part 1: the reference's entry points                                     */
/* ------------------------------------------------------------------------ */

/* O(N^2) scalar potential phi_i = sum_{j != i} charge_j dx dy dz / |r_i - r_j|
 * over the whole xdim*ydim*zdim grid (cell index z*ydim*xdim + y*xdim + x).
 * `potential` is fully overwritten.
 * Replaces: int calc_potential_exact(...)  helper_functions.hpp:31-35,
 *           defined fmm_calc.cu:412-461. */
int calc_potential_exact(const float *charge, int xdim, int ydim, int zdim,
                         float dx, float dy, float dz, float *potential,
                         int use_gpu, int verbosity);

/* Same physics as calc_potential_exact, in the float arithmetic and
 * summation order of the reference's CUDA kernel calc_potential_exact_kernel
 * (potential_calc.cu:18-116: per target, 1024 partial sums over contiguous
 * source chunks, then a halving tree), so results match that kernel bit for
 * bit wherever its compiler rounded every operation on its own.
 * Replaces: int calc_potential_exact_gpu(...)  helper_functions.hpp:37-40,
 *           defined potential_calc.cu:120-188. */
int calc_potential_exact_gpu(const float *charge, int xdim, int ydim, int zdim,
                             float dx, float dy, float dz, float *potential);

/* Fast-multipole scalar potential of order P (same physics as
 * calc_potential_exact, truncation error falling geometrically with P).
 * `potential` is fully overwritten.
 * Replaces: int fmm_calc(...)  helper_functions.hpp:24-29,
 *           defined fmm_calc.cu:295-408. */
int fmm_calc(const float *charge, float *potential, int xdim, int ydim, int zdim,
             float dx, float dy, float dz, int P, int use_gpu, int verbosity);

/* Effective field H (A/m) of magnetisation M (A/m): demag (1/4pi grad phi of
 * charge = div M), exchange (2A/(mu0 Ms^2) lap M) and the external ramp
 * (1/mu0) max(0, 0.3 - 0.3 t/1ns) x.  charge/potential receive the
 * intermediate fields as in the reference (when demag != 0).
 * Replaces: int Hfield(...)  ode_functions.hpp:8-15, ode_functions.cpp:21-70. */
int Hfield(const float *M, float *H, float *charge, float *potential, float t,
           int xdim, int ydim, int zdim, float dx, float dy, float dz, int P,
           float mu_0, float Ms, float Aexch, int demag, int exchange, int external,
           int use_fmm, int use_gpu, int verbosity);

/* One classic RK4 step of the Landau-Lifshitz equation from M (with H = field
 * at M on entry) to M2, updating material cells only; renormalises |M2| = Ms
 * when `normalize`.  On return H holds the field at the last stage, slope =
 * k1 + 2 k2 + 2 k3 + k4, this_slope = k4, *t2 = t + dt.  Stage combination and
 * renormalisation are carried out in double precision.
 * Replaces: int rk4_step(...)  ode_functions.cpp:76-148 (no header). */
int rk4_step(float t, float dt, float *t2, const unsigned char *material,
             const float *M, float *M2, float *H, float *slope, float *this_slope,
             float *charge, float *potential, int xdim, int ydim, int zdim,
             float dx, float dy, float dz, int P, float mu_0, float Ms, float Aexch,
             float alfa, float gamma, int demag, int exchange, int external,
             int use_fmm, int normalize, int use_gpu, int verbosity);

/* Adaptive time marching with the reference policy: energy-increase
 * rejection (dt /= 2.5, stop after 50 consecutive rejections), accepted-step
 * dt = 1e-14 / torque / 2, stop when torque < 1e-3 or t > finaltime; writes
 * <sim_name>/dynamics.dat, Mdynamics.dat and panic.dat in the reference's
 * text formats.  M is overwritten with the final state.  The interactive
 * pause after step 1 is not reproduced.
 * Replaces: int time_marching(...)  ode_functions.hpp:17-24,
 *           ode_functions.cpp:158-365. */
int time_marching(unsigned char *material, float *M, float finaltime, float timestep,
                  int xdim, int ydim, int zdim, float dx, float dy, float dz, int P,
                  float mu_0, float Ms, float Aexch, float alfa, float gamma,
                  int demag, int exchange, int external, int use_fmm, int use_gpu,
                  const char *sim_name, int verbosity);

/* Direct-sum potential of device arrays that allocates nothing (no
 * counterpart in the reference: for callers that cannot allocate device
 * memory -- code under stream/graph capture, tight serving loops).  Same
 * kernels and chunking as calc_potential_exact, so the same results bit for
 * bit; charge_dev / potential_dev are device arrays of xdim*ydim*zdim
 * floats; enqueued on `stream` (a hipStream_t, NULL = the default stream),
 * returns once enqueued.  The chunk partial sums use a fixed 16 MiB device
 * scratch per GPU: grids needing more (above 2,097,152 cells) are rejected
 * with an error.  Calls that overlap on one GPU share that scratch and race.
 * Variant of calc_potential_exact (fmm_calc.cu:412-461). */
int mms_potential_exact_noalloc(const float *charge_dev, float *potential_dev, int xdim, int ydim, int zdim,
                                float dx, float dy, float dz, void *stream);

/* ------------------------------------------------------------------------ */
/* Part 2: resident API (no counterpart in the reference: it replaces the     */
/* reference's function-level statics with an explicit handle)              */
/* ------------------------------------------------------------------------ */

typedef struct mms_ctx mms_ctx;

enum mms_solver {
    MMS_SOLVER_DIRECT = 0, /* O(N^2) direct sum                                */
    MMS_SOLVER_FMM = 1,    /* P2M, M2M, M2L, L2L, L2P + near-field P2P          */
    MMS_SOLVER_DIRECT_REFGPU = 2 /* O(N^2) direct sum in the float arithmetic
                              and summation order of the reference's CUDA
                              kernel (potential_calc.cu:18-116): 1024
                              contiguous-chunk partial sums per target and
                              the halving tree, bit for bit             */
};
enum mms_m2l_precision {
    MMS_M2L_FP32 = 0,   /* FP32 operands, FP32 accumulation                   */
    MMS_M2L_BF16 = 1,   /* BF16 operands, FP32 accumulation (MFMA)            */
    MMS_M2L_BF16X3 = 2, /* hi/lo BF16 split of both operands, 3 MFMA products */
    MMS_M2L_FP16 = 3,   /* FP16 operands, FP32 accumulation (MFMA): operator
                           rows/columns and each workgroup's source window are
                           scaled by powers of two into FP16 range          */
    MMS_M2L_BF16X3_FP16NEAR = 4, /* BF16X3 M2L with the near field (leaf
                           width 8) in scaled FP16, one MFMA product     */
    MMS_M2L_FP16X2 = 5, /* hi/lo FP16 split of the scaled operator against
                           the scaled FP16 source window, 2 MFMA products;
                           near field (leaf width 8) in scaled FP16       */
    MMS_M2L_FP16X2_BF16X3NEAR = 6, /* FP16X2 M2L with the near field (leaf
                           width 8) in the BF16 hi/lo split of operator and
                           charges, 3 MFMA products                      */
    MMS_M2L_FP16X2_FP16X2NEAR = 7 /* FP16X2 M2L with the near field (leaf
                           width 8) as FP16 hi/lo operator planes against
                           the scaled FP16 charges, 2 MFMA products      */
};
enum mms_external {
    MMS_EXT_NONE = 0,
    MMS_EXT_RAMP = 1,   /* the reference ramp, ode_functions.cpp:61 */
    MMS_EXT_CONST = 2   /* constant hext[3] in A/m                  */
};

typedef struct {
    int xdim, ydim, zdim;       /* grid; cell index z*ydim*xdim + y*xdim + x   */
    float dx, dy, dz;           /* cell size (m)                               */
    int solver;                 /* enum mms_solver                             */
    int P;                      /* FMM expansion order                         */
    int m2l_precision;          /* enum mms_m2l_precision                      */
    int leaf_width;             /* FMM leaf box width in cells (2, 4 or 8; 0 = 4) */
    float mu_0, Ms, Aexch, alfa, gamma;
    int demag, exchange;        /* 0/1 switches                                */
    int external;               /* enum mms_external                           */
    float hext[3];              /* A/m, for MMS_EXT_CONST                      */
    int m2l_cols;               /* M2L targets per wavefront in 16-column MFMA tiles (4 or 8; 0 = 4) */
    int direct_split;           /* direct sum: source-chunk multiplier (1, 2 or 4; 0 = 1) */
} mms_params;

/* Create a context owning device state for one film.  material: one byte per
 * cell (non-zero = magnetic), may be NULL (all magnetic). */
int mms_create(mms_ctx **ctx, const mms_params *p, const unsigned char *material);
int mms_destroy(mms_ctx *ctx);

/* State transfer (host AoS x,y,z float, A/m).  set resets t to 0. */
int mms_set_m(mms_ctx *ctx, const float *M);
int mms_get_m(mms_ctx *ctx, float *M);
/* Set the simulation time used by the ramp field. */
int mms_set_time(mms_ctx *ctx, double t);

/* Advance nsteps fixed-dt RK4 steps on the device (4 field evaluations per
 * step; the stage-1 field of each step is the post-step field of the previous
 * one, as in the reference time loop).  Enqueued on the context's stream;
 * returns once the work is enqueued.  use_graph != 0 replays one captured
 * step as a hipGraph. */
int mms_steps(mms_ctx *ctx, int nsteps, float dt, int use_graph);
int mms_sync(mms_ctx *ctx);

/* Field at the current state (host AoS output, A/m); also energy (eV) and
 * max torque |M x H|/Ms^2 as in the reference time loop. */
int mms_hfield(mms_ctx *ctx, float *H);
int mms_energy_torque(mms_ctx *ctx, double *E, double *torque);

/* Device-pointer potential solve: rho_dev, phi_dev are device float arrays
 * of xdim*ydim*zdim cells (charge density in A/m^2, potential as
 * calc_potential_exact).  Enqueued on the context's stream. */
int mms_potential_dev(mms_ctx *ctx, const float *rho_dev, float *phi_dev);

/* Host-array potential through the context's solver (synchronous). */
int mms_potential(mms_ctx *ctx, const float *charge, float *potential);

/* FP64 direct-sum potential at selected cells (device FP64 arithmetic), used
 * to measure solver error on large grids: out[i] = phi(probe[i]). */
int mms_potential_probes_f64(mms_ctx *ctx, const float *charge, const int64_t *probes,
                             int nprobe, double *out);

/* Autotuner: time the solver's launch shapes on the context's current
 * charge (call after mms_set_m) -- FMM leaf width in {4, 8} x M2L column tiles
 * in {4, 8}; direct sum source-chunk multiplier in {1, 2, 4} -- `reps` solves
 * each, keep the fastest.  Accuracy settings (P, M2L precision) are not
 * changed.  report (may be NULL) receives one line per candidate. */
int mms_autotune(mms_ctx *ctx, int reps, char *report, int report_len);
/* Current launch shape: leaf width, M2L column tiles, direct split. */
int mms_get_shape(mms_ctx *ctx, int *leaf_width, int *m2l_cols, int *direct_split);

/* Timing: mean device time (ms) per launch of each kernel class over the last
 * mms_profile_end() window; names are returned through mms_profile_name(). */
int mms_profile_begin(mms_ctx *ctx);
int mms_profile_end(mms_ctx *ctx);
int mms_profile_count(mms_ctx *ctx);
const char *mms_profile_name(mms_ctx *ctx, int i);
double mms_profile_ms(mms_ctx *ctx, int i);
long mms_profile_launches(mms_ctx *ctx, int i);

/* ------------------------------------------------------------------------ */
/* Part 2b: learned far-field surrogate (no counterpart in the reference)    */
/* ------------------------------------------------------------------------ */
/* A small convolutional neural operator maps each layer's charge density to
 * its potential: 3x3 conv lift (3 input planes: the layer's own charge, the
 * summed charges of the layers at distance 1, the summed charges of the
 * layers at distance >= 2 -> 64 channels), three 3x3 convs 64 -> 64 (the
 * second with a spectral-mixing branch over 32 x 32 real Fourier modes), 1x1
 * projection; BF16 weights and activations, FP32 accumulation on the matrix
 * cores.  It may stand in for the solver at RK4 stages 2-3 while the solver
 * runs at stages 1 and 4.                                                    */

/* Number of FP32 weights mms_surrogate_init takes, in the order: lift
 * [64][3][3][3] + bias[64]; conv1, conv2 [64][64][3][3] + bias[64] each;
 * mode mixing [32 ky][32 kx][64 co][64 ci]; conv3 [64][64][3][3] + bias[64];
 * projection [64] + bias. */
long mms_surrogate_weight_count(void);
/* Load the network (weights NULL: a seeded untrained initialisation).  The
 * network sees charge / charge_scale and outputs potential / potential_scale
 * (<= 0: the film's natural units Ms/dx and Ms dx). */
int mms_surrogate_init(mms_ctx *ctx, const float *weights, float charge_scale, float potential_scale);
/* One inference: host arrays (synchronous) or device arrays (enqueued). */
int mms_surrogate_potential(mms_ctx *ctx, const float *charge, float *potential);
int mms_surrogate_potential_dev(mms_ctx *ctx, const float *rho_dev, float *phi_dev);
/* Use in mms_steps: mode 0 off; 1 the surrogate at RK4 stages 2-3 of every
 * step (the stage-1 judge residual is still computed and kept on the device;
 * step graphs allowed); 2 judged: at stage 1 the surrogate and the solver
 * both run, and stages 2-3 take the surrogate only when the relative L2
 * residual ||phi_sur - phi|| / ||phi|| <= tol, else the solver (decided on
 * the host each step, no step graph). */
int mms_surrogate_mode(mms_ctx *ctx, int mode, float tol);
/* Last stage-1 residual (-1 if none), steps judged and accepted (mode 2). */
int mms_surrogate_stats(mms_ctx *ctx, double *last_residual, long *judged, long *accepted);

/* Training (FP32 parameters, BF16 matrix-core forward and backward passes,
 * Adam with betas 0.9 / 0.999 and eps 1e-8).  The loss of a batch of films is
 * the mean over its films of ||phi_sur - phi||^2 / ||phi||^2. */
typedef struct mms_sur_trainer mms_sur_trainer;
/* Film state `index` of a seeded family (SplitMix64 from seed and index): a
 * disc, rectangle or ring mask (the same in every layer) holding a vortex
 * (chirality, polarity +-1, core radius 2 cells) or uniform in-plane domains
 * of 8, 16 or 32 cells, with 2% noise; |M| = Ms on the mask, 0 elsewhere.  M
 * host AoS [n][3]; material (may be NULL) one byte per cell; kind (may be
 * NULL) = 2 * shape (0 disc, 1 rectangle, 2 ring) + state (0 vortex, 1
 * domains).  Host only. */
int mms_surrogate_state(int xdim, int ydim, int zdim, float Ms, unsigned long long seed, long index, float *M,
                        unsigned char *material, int *kind);
/* A trainer for films of xdim x ydim x zdim cubic cells of side dx, batches
 * of `batch` films, charge and potential in the natural units Ms / dx and
 * Ms dx; weights NULL: the seeded initialisation of mms_surrogate_init. */
int mms_surrogate_trainer_create(mms_sur_trainer **t, int xdim, int ydim, int zdim, float dx, float Ms, int batch,
                                 const float *weights);
int mms_surrogate_trainer_destroy(mms_sur_trainer *t);
/* Append npairs training pairs: the states mms_surrogate_state(seed, i) for
 * i = pairs already held, ..., their charge (the divergence of
 * Hfield/rk4_step) and potential (FMM of order P; P <= 0: the direct sum),
 * computed on the device. */
int mms_surrogate_trainer_generate(mms_sur_trainer *t, int npairs, unsigned long long seed, int P);
/* Append npairs pairs from host arrays [npairs][n]. */
int mms_surrogate_trainer_add(mms_sur_trainer *t, int npairs, const float *charge, const float *potential);
long mms_surrogate_trainer_pairs(mms_sur_trainer *t);
int mms_surrogate_trainer_get_pair(mms_sur_trainer *t, long i, float *charge, float *potential);
/* Gradient of the loss of the batch of pairs idx[0..batch) (API weight order,
 * mms_surrogate_weight_count floats; may be NULL) and the loss, no update. */
int mms_surrogate_trainer_gradient(mms_sur_trainer *t, const long *idx, float *grad, double *loss);
/* One Adam step at rate lr on the batch idx; loss before the step. */
int mms_surrogate_trainer_step(mms_sur_trainer *t, const long *idx, float lr, double *loss);
/* Relative L2 residual ||phi_sur - phi|| / ||phi|| of each film of the batch idx. */
int mms_surrogate_trainer_eval(mms_sur_trainer *t, const long *idx, double *residual);
/* The network's potential of the batch idx, host [batch][n]. */
int mms_surrogate_trainer_predict(mms_sur_trainer *t, const long *idx, float *potential);
/* Current weights in API order (for mms_surrogate_init). */
int mms_surrogate_trainer_weights(mms_sur_trainer *t, float *weights);

/* ------------------------------------------------------------------------ */
/* Part 3: many independent films on one GPU (no counterpart in the          */
/* reference, which runs one film per process)                              */
/* ------------------------------------------------------------------------ */

typedef struct mms_batch mms_batch;
typedef struct {
    int xdim, ydim;   /* film extent (zdim, cell size and material constants are common) */
    float hext[3];    /* constant external field of this film, A/m                        */
} mms_film_spec;

/* Create a batch.  `common` gives zdim, dx/dy/dz, P, m2l_precision,
 * m2l_cols, leaf_width, mu_0, Ms, Aexch, alfa, gamma, demag, exchange;
 * solver = MMS_SOLVER_FMM lets films of >= 4096 cells use the FMM (padded to
 * power-of-two size classes), smaller films use the direct sum; material is
 * the films' flags packed in film order (NULL = all magnetic).  Films do not
 * interact. */
int mms_batch_create(mms_batch **b, const mms_params *common, int nfilm, const mms_film_spec *films,
                     const unsigned char *material);
int mms_batch_destroy(mms_batch *b);
/* Packed host AoS M of every film in film order (film f's cells are
 * contiguous, index z*ydim*xdim + y*xdim + x within the film). */
int mms_batch_set_m(mms_batch *b, const float *M);
int mms_batch_get_m(mms_batch *b, float *M);
/* nsteps fixed-dt RK4 steps of every film; one launch per kernel class per
 * stage over all films; use_graph != 0 replays a captured step. */
int mms_batch_steps(mms_batch *b, int nsteps, float dt, int use_graph);
int mms_batch_sync(mms_batch *b);
/* Demag potential of every film from packed host charge densities (A/m^2,
 * film order, as mms_batch_set_m packs M) with the batch's own solvers (the
 * direct sum or the film's FMM size class); synchronous. */
int mms_batch_potential(mms_batch *b, const float *charge, float *potential);
int mms_batch_info(mms_batch *b, int *n_direct, int *n_fmm, int *n_classes, long *cells, int *launches_per_step);
/* Autotuner: per FMM size class, leaf width x M2L column tiles on the current state. */
int mms_batch_autotune(mms_batch *b, int reps, char *report, int report_len);
int mms_batch_profile_begin(mms_batch *b);
/* mean ms and launch count per kernel class (arrays of mms_profile_count entries) */
int mms_batch_profile_end(mms_batch *b, double *ms_per_class, long *launches_per_class);

/* Message for the last failing call on this host thread ("" if none). */
const char *mms_last_error(void);
/* Library build identifier (target arch, version). */
const char *mms_version(void);

#ifdef __cplusplus
}
#endif

#endif /* MMS_GPU_H */
