// api.cpp -- C ABI of libmmsgpu.so: contexts, the reference-mirroring entry
// points (calc_potential_exact, fmm_calc, Hfield, rk4_step, time_marching) and
// the resident fixed-step driver.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <sys/time.h>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "mms_internal.hpp"

namespace mms {

static thread_local std::string g_err;
void set_error(const std::string &msg) { g_err = msg; }

const char *kclass_name(int k) {
    static const char *names[K_COUNT] = {"div",      "stage",  "direct", "direct_reduce", "upward", "m2m",
                                         "m2l",      "downward", "p2p",  "l2p",           "hfield", "energy",
                                         "probes",   "misc",     "surrogate"};
    return (k >= 0 && k < K_COUNT) ? names[k] : "?";
}

int Profiler::begin(int k, hipEvent_t *a, hipEvent_t *b) {
    (void)k;
    for (hipEvent_t *e : {a, b}) {
        if (!pool.empty()) {
            *e = pool.back();
            pool.pop_back();
        } else if (hipEventCreate(e) != hipSuccess) {
            *a = nullptr;
            return 1;
        }
    }
    if (hipEventRecord(*a, stream) != hipSuccess) { *a = nullptr; return 1; }
    return 0;
}
int Profiler::end_record(int k, hipEvent_t a, hipEvent_t b) {
    if (hipEventRecord(b, stream) != hipSuccess) return 1;
    pending.push_back({k, a, b});
    if (pending.size() > 4096) return flush();
    return 0;
}
int Profiler::flush() {
    for (const Pending &p : pending) {
        float t = 0.0f;
        if (hipEventSynchronize(p.b) == hipSuccess && hipEventElapsedTime(&t, p.a, p.b) == hipSuccess) {
            ms[p.k] += t;
            count[p.k]++;
        }
        pool.push_back(p.a);
        pool.push_back(p.b);
    }
    pending.clear();
    return 0;
}
void Profiler::reset() {
    for (int k = 0; k < K_COUNT; k++) { ms[k] = 0; count[k] = 0; }
}
Profiler::~Profiler() {
    for (const Pending &p : pending) { pool.push_back(p.a); pool.push_back(p.b); }
    for (hipEvent_t e : pool) (void)hipEventDestroy(e);
}

}  // namespace mms

using namespace mms;

// ============================================================== context ==
static int ctx_alloc(mms_ctx *c, const unsigned char *material) {
    const long n = c->g.n();
    for (int i = 0; i < 3; i++) MMS_HIP(hipMalloc(&c->m32[i], sizeof(float) * 3 * n));
    MMS_HIP(hipMalloc(&c->mn64, sizeof(double) * 3 * n));
    MMS_HIP(hipMalloc(&c->mnew64, sizeof(double) * 3 * n));
    MMS_HIP(hipMalloc(&c->acc32, sizeof(float) * 3 * n));
    MMS_HIP(hipMalloc(&c->rho, sizeof(float) * n));
    MMS_HIP(hipMalloc(&c->phi, sizeof(float) * n));
    MMS_HIP(hipMalloc(&c->matl, n));
    MMS_HIP(hipMalloc(&c->t_dev, sizeof(double)));
    MMS_HIP(hipMalloc(&c->h_aos, sizeof(float) * 3 * n));
    MMS_HIP(hipMalloc(&c->aux_aos, sizeof(float) * 9 * n));
    c->red_blocks = 1024;
    MMS_HIP(hipMalloc(&c->red, sizeof(double) * 5 * c->red_blocks));
    MMS_HIP(hipMemsetAsync(c->m32[0], 0, sizeof(float) * 3 * n, c->stream));
    MMS_HIP(hipMemsetAsync(c->m32[1], 0, sizeof(float) * 3 * n, c->stream));
    MMS_HIP(hipMemsetAsync(c->m32[2], 0, sizeof(float) * 3 * n, c->stream));
    MMS_HIP(hipMemsetAsync(c->mn64, 0, sizeof(double) * 3 * n, c->stream));
    MMS_HIP(hipMemsetAsync(c->acc32, 0, sizeof(float) * 3 * n, c->stream));
    MMS_HIP(hipMemsetAsync(c->t_dev, 0, sizeof(double), c->stream));
    if (material) {
        MMS_HIP(hipMemcpyAsync(c->matl, material, n, hipMemcpyHostToDevice, c->stream));
    } else {
        MMS_HIP(hipMemsetAsync(c->matl, 1, n, c->stream));
    }
    MMS_HIP(hipStreamSynchronize(c->stream));
    if (material) {
        c->host_matl.assign(material, material + n);
    } else {
        c->host_matl.assign(n, 1);
    }
    c->all_magnetic = true;
    for (long i = 0; i < n; i++)
        if (!c->host_matl[i]) { c->all_magnetic = false; break; }
    return 0;
}

static void ctx_free(mms_ctx *c) {
    if (c->graph) (void)hipGraphExecDestroy(c->graph);
    c->graph = nullptr;
    void *bufs[] = {c->m32[0], c->m32[1], c->m32[2], c->mn64, c->mnew64, c->acc32, c->rho, c->phi,
                    c->matl, c->t_dev, c->h_aos, c->aux_aos, c->red};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    c->direct.release();
    c->fmm.release();
    c->sur.release();
    if (c->phi_s) (void)hipFree(c->phi_s);
    c->phi_s = nullptr;
    if (c->stream) (void)hipStreamDestroy(c->stream);
    c->stream = nullptr;
}

int mms_ctx::potential(const float *rho_d, float *phi_d) {
    if (p.solver == MMS_SOLVER_FMM) return fmm_potential(g, fmm, rho_d, phi_d, stream, &prof);
    if (p.solver == MMS_SOLVER_DIRECT_REFGPU) return direct_refgpu_potential(g, rho_d, phi_d, stream, &prof);
    return direct_potential(g, direct, rho_d, phi_d, stream, &prof);
}

// one fixed-dt RK4 step on the device: four (div, potential, stage) passes;
// the stage-0 field is evaluated at M_n.  State returns to m32[cur], which
// keeps the FP32 copy of M_n through stages 0-2 (the other two slots carry
// the stage values).
int mms_ctx::one_step(float dt) {
    const int a = cur, b = (cur + 1) % 3, d = (cur + 2) % 3;
    const float *in[4] = {m32[a], m32[b], m32[d], m32[b]};
    float *out[4] = {m32[b], m32[d], m32[b], m32[a]};
    bool accept = (sur_mode == 1);
    for (int s = 0; s < 4; s++) {
        if (p.demag) {
            if (launch_div(g, in[s], rho, stream, &prof)) return 1;
            // the learned surrogate stands in at RK4 stages 2-3; the solver
            // runs at stages 1 and 4 and, at stage 1, judges the surrogate
            const bool use_sur = sur_mode != 0 && (s == 1 || s == 2) && accept;
            if (use_sur) {
                if (surrogate_potential(sur, rho, phi, sur_rscale, sur_oscale, stream, &prof)) return 1;
            } else if (potential(rho, phi)) {
                return 1;
            }
            if (sur_mode != 0 && s == 0) {
                if (surrogate_potential(sur, rho, phi_s, sur_rscale, sur_oscale, stream, &prof)) return 1;
                double r = 0.0;
                if (surrogate_residual(sur, phi_s, phi, g.n(), sur_mode == 2 ? &r : nullptr, stream)) return 1;
                if (sur_mode == 2) {
                    sur_last = r;
                    sur_judged++;
                    accept = r <= sur_tol;
                    if (accept) sur_accepted++;
                }
            }
        }
        // an all-magnetic film skips the per-cell material byte
        if (launch_stage(g, mat, s, (double)dt, phi, in[s], out[s], mn64, m32[a], acc32, all_magnetic ? nullptr : matl,
                         t_dev, stream, &prof))
            return 1;
    }
    if (p.external == MMS_EXT_RAMP && launch_add_time(t_dev, (double)dt, stream)) return 1;
    return 0;
}

static int validate_params(const mms_params *p) {
    MMS_CHECK(p != nullptr, "null parameters");
    MMS_CHECK(p->xdim > 0 && p->ydim > 0 && p->zdim > 0, "grid dimensions must be positive");
    MMS_CHECK((long)p->xdim * p->ydim * p->zdim < (1L << 31), "grid too large (>= 2^31 cells)");
    MMS_CHECK(p->dx > 0 && p->dy > 0 && p->dz > 0, "cell sizes must be positive");
    MMS_CHECK(p->solver == MMS_SOLVER_DIRECT || p->solver == MMS_SOLVER_FMM || p->solver == MMS_SOLVER_DIRECT_REFGPU,
              "unknown solver");
    MMS_CHECK(p->m2l_precision >= MMS_M2L_FP32 && p->m2l_precision <= MMS_M2L_FP16X2_FP16X2NEAR, "unknown M2L precision");
    return 0;
}

extern "C" int mms_create(mms_ctx **out, const mms_params *p, const unsigned char *material) {
    if (!out) { set_error("null context pointer"); return 1; }
    *out = nullptr;
    if (validate_params(p)) return 1;
    std::unique_ptr<mms_ctx> c(new mms_ctx());
    c->p = *p;
    if (c->p.leaf_width == 0) c->p.leaf_width = 4;
    c->g.xd = p->xdim; c->g.yd = p->ydim; c->g.zd = p->zdim;
    c->g.dx = p->dx; c->g.dy = p->dy; c->g.dz = p->dz;
    c->g.ay = p->dy / p->dx; c->g.az = p->dz / p->dx;
    c->mat.mu_0 = p->mu_0; c->mat.Ms = p->Ms; c->mat.Aexch = p->Aexch; c->mat.alfa = p->alfa;
    c->mat.gamma = p->gamma; c->mat.demag = p->demag; c->mat.exchange = p->exchange;
    c->mat.external = p->external;
    for (int i = 0; i < 3; i++) c->mat.hext[i] = p->hext[i];
    MMS_HIP(hipGetDevice(&c->device));
    MMS_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    c->prof.stream = c->stream;
    if (ctx_alloc(c.get(), material)) { ctx_free(c.get()); return 1; }
    int st = 0;
    if (c->p.m2l_cols == 0) c->p.m2l_cols = 4;
    if (c->p.direct_split == 0) c->p.direct_split = 1;
    if (c->p.solver == MMS_SOLVER_FMM)
        st = c->fmm.init(c->g, c->p.P, c->p.leaf_width, c->p.m2l_precision, c->p.m2l_cols);
    else if (c->p.solver == MMS_SOLVER_DIRECT)
        st = c->direct.init(c->g, c->p.direct_split);
    if (st) { ctx_free(c.get()); return 1; }
    *out = c.release();
    return 0;
}

extern "C" int mms_destroy(mms_ctx *c) {
    if (!c) return 0;
    ctx_free(c);
    delete c;
    return 0;
}

static int upload_m(mms_ctx *c, const float *M) {
    const long n = c->g.n();
    float *staging = c->aux_aos;  // device AoS staging
    MMS_HIP(hipMemcpyAsync(staging, M, sizeof(float) * 3 * n, hipMemcpyHostToDevice, c->stream));
    if (launch_soa_from_aos(staging, c->m32[c->cur], c->mn64, n, c->stream)) return 1;
    return 0;
}

extern "C" int mms_set_m(mms_ctx *c, const float *M) {
    MMS_CHECK(c && M, "null argument");
    if (upload_m(c, M)) return 1;
    c->t_host = 0.0;
    MMS_HIP(hipMemsetAsync(c->t_dev, 0, sizeof(double), c->stream));
    MMS_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int mms_get_m(mms_ctx *c, float *M) {
    MMS_CHECK(c && M, "null argument");
    const long n = c->g.n();
    if (launch_aos_from_soa(c->m32[c->cur], c->aux_aos, n, c->stream)) return 1;
    MMS_HIP(hipMemcpyAsync(M, c->aux_aos, sizeof(float) * 3 * n, hipMemcpyDeviceToHost, c->stream));
    MMS_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int mms_set_time(mms_ctx *c, double t) {
    MMS_CHECK(c, "null context");
    c->t_host = t;
    MMS_HIP(hipMemcpyAsync(c->t_dev, &c->t_host, sizeof(double), hipMemcpyHostToDevice, c->stream));
    MMS_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int mms_steps(mms_ctx *c, int nsteps, float dt, int use_graph) {
    MMS_CHECK(c, "null context");
    MMS_CHECK(nsteps >= 0, "negative step count");
    // the judged surrogate decides on the host after stage 1 of every step:
    // no step graph
    if (!use_graph || c->prof.on || c->sur_mode == 2) {
        for (int i = 0; i < nsteps; i++)
            if (c->one_step(dt)) return 1;
        c->t_host += (double)nsteps * dt;
        return 0;
    }
    if (!c->graph || c->graph_dt != dt) {
        if (c->graph) { (void)hipGraphExecDestroy(c->graph); c->graph = nullptr; }
        hipGraph_t gr;
        MMS_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        int st = c->one_step(dt);
        hipError_t e = hipStreamEndCapture(c->stream, &gr);
        if (st) return 1;
        MMS_HIP(e);
        hipError_t e2 = hipGraphInstantiate(&c->graph, gr, nullptr, nullptr, 0);
        (void)hipGraphDestroy(gr);
        MMS_HIP(e2);
        c->graph_dt = dt;
    }
    for (int i = 0; i < nsteps; i++) MMS_HIP(hipGraphLaunch(c->graph, c->stream));
    c->t_host += (double)nsteps * dt;
    return 0;
}

// ------------------------------------------------- learned surrogate --
extern "C" long mms_surrogate_weight_count(void) { return surrogate_weight_count(); }

extern "C" int mms_surrogate_init(mms_ctx *c, const float *weights, float charge_scale, float potential_scale) {
    MMS_CHECK(c, "null context");
    if (c->graph) { (void)hipGraphExecDestroy(c->graph); c->graph = nullptr; }
    // natural units of the film: charge ~ Ms / dx, potential ~ Ms dx
    c->sur_rscale = charge_scale > 0 ? 1.0f / charge_scale : c->g.dx / c->p.Ms;
    c->sur_oscale = potential_scale > 0 ? potential_scale : c->p.Ms * c->g.dx;
    if (c->sur.init(c->g, c->g.zd, weights, false, c->stream)) return 1;
    if (!c->phi_s) MMS_HIP(hipMalloc(&c->phi_s, sizeof(float) * c->g.n()));
    return 0;
}

extern "C" int mms_surrogate_potential_dev(mms_ctx *c, const float *rho_dev, float *phi_dev) {
    MMS_CHECK(c && rho_dev && phi_dev, "null argument");
    return surrogate_potential(c->sur, rho_dev, phi_dev, c->sur_rscale, c->sur_oscale, c->stream, &c->prof);
}

extern "C" int mms_surrogate_potential(mms_ctx *c, const float *charge, float *potential) {
    MMS_CHECK(c && charge && potential, "null argument");
    const long n = c->g.n();
    MMS_HIP(hipMemcpyAsync(c->rho, charge, sizeof(float) * n, hipMemcpyHostToDevice, c->stream));
    if (mms_surrogate_potential_dev(c, c->rho, c->phi)) return 1;
    MMS_HIP(hipMemcpyAsync(potential, c->phi, sizeof(float) * n, hipMemcpyDeviceToHost, c->stream));
    MMS_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int mms_surrogate_mode(mms_ctx *c, int mode, float tol) {
    MMS_CHECK(c, "null context");
    MMS_CHECK(mode >= 0 && mode <= 2, "surrogate mode must be 0, 1 or 2");
    MMS_CHECK(mode == 0 || c->sur.d_h0, "surrogate not initialised (mms_surrogate_init)");
    if (c->graph) { (void)hipGraphExecDestroy(c->graph); c->graph = nullptr; }
    c->sur_mode = mode;
    c->sur_tol = tol;
    c->sur_judged = c->sur_accepted = 0;
    c->sur_last = -1.0;
    return 0;
}

extern "C" int mms_surrogate_stats(mms_ctx *c, double *last_residual, long *judged, long *accepted) {
    MMS_CHECK(c, "null context");
    if (c->sur_mode == 1 && c->sur.d_part) {  // the last step's stage-1 residual, left on the device
        std::vector<double> h(2048);
        MMS_HIP(hipMemcpyAsync(h.data(), c->sur.d_part, sizeof(double) * 2048, hipMemcpyDeviceToHost, c->stream));
        MMS_HIP(hipStreamSynchronize(c->stream));
        double d2 = 0, b2 = 0;
        for (int i = 0; i < 1024; i++) { d2 += h[2 * i]; b2 += h[2 * i + 1]; }
        c->sur_last = b2 > 0 ? sqrt(d2 / b2) : 0.0;
    }
    if (last_residual) *last_residual = c->sur_last;
    if (judged) *judged = c->sur_judged;
    if (accepted) *accepted = c->sur_accepted;
    return 0;
}

extern "C" int mms_sync(mms_ctx *c) {
    MMS_CHECK(c, "null context");
    MMS_HIP(hipStreamSynchronize(c->stream));
    if (c->prof.on) c->prof.flush();
    return 0;
}

// field at the current state into c->h_aos (and rho/phi)
static int field_now(mms_ctx *c, const float *m) {
    if (c->p.demag) {
        if (launch_div(c->g, m, c->rho, c->stream, &c->prof)) return 1;
        if (c->potential(c->rho, c->phi)) return 1;
    }
    return launch_hfield(c->g, c->mat, c->phi, m, c->t_dev, c->h_aos, c->stream, &c->prof);
}

extern "C" int mms_hfield(mms_ctx *c, float *H) {
    MMS_CHECK(c && H, "null argument");
    if (field_now(c, c->m32[c->cur])) return 1;
    MMS_HIP(hipMemcpyAsync(H, c->h_aos, sizeof(float) * 3 * c->g.n(), hipMemcpyDeviceToHost, c->stream));
    MMS_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

// energy (eV) and torque from the AoS field h at state m (reference units)
static int energy_torque_of(mms_ctx *c, const float *m, const float *h, double *E, double *T) {
    const long n = c->g.n();
    if (launch_energy_h(m, h, c->matl, n, c->red, c->red_blocks, c->stream, &c->prof)) return 1;
    std::vector<double> part(2 * c->red_blocks);
    MMS_HIP(hipMemcpyAsync(part.data(), c->red, sizeof(double) * 2 * c->red_blocks, hipMemcpyDeviceToHost,
                           c->stream));
    MMS_HIP(hipStreamSynchronize(c->stream));
    double e = 0, t = 0;
    for (int i = 0; i < c->red_blocks; i++) {
        e += part[2 * i];
        t = t > part[2 * i + 1] ? t : part[2 * i + 1];
    }
    const double dV = (double)(c->g.dx * c->g.dy * c->g.dz);
    *E = e * (-0.5 * (double)c->p.mu_0 * dV / 1.6e-19);
    *T = t / c->p.Ms / c->p.Ms;
    return 0;
}

extern "C" int mms_energy_torque(mms_ctx *c, double *E, double *T) {
    MMS_CHECK(c && E && T, "null argument");
    if (field_now(c, c->m32[c->cur])) return 1;
    return energy_torque_of(c, c->m32[c->cur], c->h_aos, E, T);
}

extern "C" int mms_potential_dev(mms_ctx *c, const float *rho_dev, float *phi_dev) {
    MMS_CHECK(c && rho_dev && phi_dev, "null argument");
    return c->potential(rho_dev, phi_dev);
}

extern "C" int mms_potential(mms_ctx *c, const float *charge, float *potential) {
    MMS_CHECK(c && charge && potential, "null argument");
    const long n = c->g.n();
    MMS_HIP(hipMemcpyAsync(c->rho, charge, sizeof(float) * n, hipMemcpyHostToDevice, c->stream));
    if (c->potential(c->rho, c->phi)) return 1;
    MMS_HIP(hipMemcpyAsync(potential, c->phi, sizeof(float) * n, hipMemcpyDeviceToHost, c->stream));
    MMS_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" int mms_potential_probes_f64(mms_ctx *c, const float *charge, const int64_t *probes, int np,
                                        double *out) {
    MMS_CHECK(c && charge && probes && out && np >= 0, "bad argument");
    const long n = c->g.n();
    for (int i = 0; i < np; i++) MMS_CHECK(probes[i] >= 0 && probes[i] < n, "probe index out of range");
    int64_t *dp = nullptr;
    double *dout = nullptr;
    MMS_HIP(hipMemcpyAsync(c->rho, charge, sizeof(float) * n, hipMemcpyHostToDevice, c->stream));
    MMS_HIP(hipMalloc(&dp, sizeof(int64_t) * (np > 0 ? np : 1)));
    MMS_HIP(hipMalloc(&dout, sizeof(double) * (np > 0 ? np : 1)));
    int st = 0;
    if (hipMemcpyAsync(dp, probes, sizeof(int64_t) * np, hipMemcpyHostToDevice, c->stream) != hipSuccess) st = 1;
    if (!st) st = direct_probes_f64(c->g, c->rho, dp, np, dout, c->stream, &c->prof);
    if (!st && hipMemcpyAsync(out, dout, sizeof(double) * np, hipMemcpyDeviceToHost, c->stream) != hipSuccess) st = 1;
    if (!st && hipStreamSynchronize(c->stream) != hipSuccess) st = 1;
    (void)hipFree(dp);
    (void)hipFree(dout);
    if (st) set_error("probe evaluation failed");
    return st;
}

// ------------------------------------------------------------ autotuner --
static int time_solve(mms_ctx *c, int reps, double *ms) {
    hipEvent_t a, b;
    MMS_HIP(hipEventCreate(&a));
    MMS_HIP(hipEventCreate(&b));
    int st = c->potential(c->rho, c->phi);  // warm-up
    if (!st && hipEventRecord(a, c->stream) != hipSuccess) st = 1;
    for (int i = 0; i < reps && !st; i++) st = c->potential(c->rho, c->phi);
    if (!st && hipEventRecord(b, c->stream) != hipSuccess) st = 1;
    float t = 0.0f;
    if (!st && (hipEventSynchronize(b) != hipSuccess || hipEventElapsedTime(&t, a, b) != hipSuccess)) st = 1;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    *ms = t / (reps > 0 ? reps : 1);
    return st;
}

extern "C" int mms_autotune(mms_ctx *c, int reps, char *report, int report_len) {
    MMS_CHECK(c, "null context");
    if (reps < 1) reps = 3;
    if (c->graph) { (void)hipGraphExecDestroy(c->graph); c->graph = nullptr; }
    std::string rep;
    // charge of the current state
    if (launch_div(c->g, c->m32[c->cur], c->rho, c->stream, nullptr)) return 1;
    char line[160];
    if (c->p.solver == MMS_SOLVER_FMM) {
        int best_w = c->p.leaf_width, best_c = c->p.m2l_cols;
        double best = 1e30;
        const int side = c->g.xd > c->g.yd ? c->g.xd : c->g.yd;
        int S = 1;
        while (S < side) S <<= 1;
        for (int w : {4, 8}) {
            if (S / w < 4) continue;
            for (int cols : {4, 8}) {
                if (c->fmm.init(c->g, c->p.P, w, c->p.m2l_precision, cols)) return 1;
                double ms = 0;
                if (time_solve(c, reps, &ms)) return 1;
                snprintf(line, sizeof line, "fmm leaf=%d m2l_cols=%d: %.4f ms/solve\n", w, cols, ms);
                rep += line;
                if (ms < best) { best = ms; best_w = w; best_c = cols; }
            }
        }
        c->p.leaf_width = best_w;
        c->p.m2l_cols = best_c;
        if (c->fmm.init(c->g, c->p.P, best_w, c->p.m2l_precision, best_c)) return 1;
        snprintf(line, sizeof line, "chosen: leaf=%d m2l_cols=%d\n", best_w, best_c);
        rep += line;
    } else if (c->p.solver == MMS_SOLVER_DIRECT_REFGPU) {
        rep += "reference-order direct sum: fixed launch shape (the reference's 1024-way split)\n";
    } else {
        int best_s = c->p.direct_split;
        double best = 1e30;
        for (int sp : {1, 2, 4}) {
            if (c->direct.init(c->g, sp)) return 1;
            double ms = 0;
            if (time_solve(c, reps, &ms)) return 1;
            snprintf(line, sizeof line, "direct split=%d: %.4f ms/solve\n", sp, ms);
            rep += line;
            if (ms < best) { best = ms; best_s = sp; }
        }
        c->p.direct_split = best_s;
        if (c->direct.init(c->g, best_s)) return 1;
        snprintf(line, sizeof line, "chosen: direct split=%d\n", best_s);
        rep += line;
    }
    if (report && report_len > 0) {
        strncpy(report, rep.c_str(), report_len - 1);
        report[report_len - 1] = 0;
    }
    return 0;
}

extern "C" int mms_get_shape(mms_ctx *c, int *leaf_width, int *m2l_cols, int *direct_split) {
    MMS_CHECK(c, "null context");
    if (leaf_width) *leaf_width = c->p.solver == MMS_SOLVER_FMM ? c->fmm.w : 0;
    if (m2l_cols) *m2l_cols = c->p.solver == MMS_SOLVER_FMM ? c->fmm.m2l_cols : 0;
    if (direct_split) *direct_split = c->p.solver == MMS_SOLVER_DIRECT ? c->direct.split : 0;
    return 0;
}

extern "C" int mms_profile_begin(mms_ctx *c) {
    MMS_CHECK(c, "null context");
    c->prof.flush();
    c->prof.reset();
    c->prof.on = true;
    return 0;
}
extern "C" int mms_profile_end(mms_ctx *c) {
    MMS_CHECK(c, "null context");
    MMS_HIP(hipStreamSynchronize(c->stream));
    c->prof.flush();
    c->prof.on = false;
    return 0;
}
extern "C" int mms_profile_count(mms_ctx *c) { (void)c; return K_COUNT; }
extern "C" const char *mms_profile_name(mms_ctx *c, int i) { (void)c; return kclass_name(i); }
extern "C" double mms_profile_ms(mms_ctx *c, int i) {
    if (!c || i < 0 || i >= K_COUNT || c->prof.count[i] == 0) return 0.0;
    return c->prof.ms[i] / (double)c->prof.count[i];
}
extern "C" long mms_profile_launches(mms_ctx *c, int i) {
    if (!c || i < 0 || i >= K_COUNT) return 0;
    return c->prof.count[i];
}

extern "C" const char *mms_last_error(void) { return g_err.c_str(); }
extern "C" const char *mms_version(void) { return "mmsgpu 0.1 gfx950"; }

// ============================================== reference entry points ==
// Contexts for the Part-1 calls are cached per (shape, spacing, solver, P,
// material constants); at most a few are kept, least recently used evicted.
namespace {
struct CacheEntry {
    mms_params p;
    std::vector<unsigned char> material;
    mms_ctx *ctx;
    unsigned long used;
};
std::mutex g_cache_mu;
std::vector<CacheEntry> g_cache;
unsigned long g_clock = 0;

bool same_params(const mms_params &a, const mms_params &b) { return memcmp(&a, &b, sizeof(mms_params)) == 0; }

mms_ctx *cached_ctx(const mms_params &p, const unsigned char *material) {
    const long n = (long)p.xdim * p.ydim * p.zdim;
    for (CacheEntry &e : g_cache) {
        if (!same_params(e.p, p)) continue;
        if (material) {
            if (e.material.size() != (size_t)n || memcmp(e.material.data(), material, n) != 0) continue;
        } else if (!e.material.empty()) {
            continue;
        }
        e.used = ++g_clock;
        return e.ctx;
    }
    if (g_cache.size() >= 4) {
        size_t victim = 0;
        for (size_t i = 1; i < g_cache.size(); i++)
            if (g_cache[i].used < g_cache[victim].used) victim = i;
        mms_destroy(g_cache[victim].ctx);
        g_cache.erase(g_cache.begin() + victim);
    }
    mms_ctx *c = nullptr;
    if (mms_create(&c, &p, material)) return nullptr;
    CacheEntry e{p, {}, c, ++g_clock};
    if (material) e.material.assign(material, material + n);
    g_cache.push_back(e);
    return c;
}

// MMS_GPU_REFERENCE_ORDER=1: every direct-sum solve of the reference entry
// points (calc_potential_exact, Hfield, rk4_step, time_marching without the
// FMM) runs in the reference CUDA kernel's summation order, as the
// reference's use_gpu = 1 path does, instead of the tuned kernel
int direct_solver() {
    const char *e = std::getenv("MMS_GPU_REFERENCE_ORDER");
    return (e && e[0] == '1') ? MMS_SOLVER_DIRECT_REFGPU : MMS_SOLVER_DIRECT;
}

mms_params base_params(int xd, int yd, int zd, float dx, float dy, float dz) {
    mms_params p;
    memset(&p, 0, sizeof(p));
    p.xdim = xd; p.ydim = yd; p.zdim = zd;
    p.dx = dx; p.dy = dy; p.dz = dz;
    p.solver = direct_solver();
    p.P = 4;
    p.leaf_width = 4;
    p.demag = 1;
    return p;
}

// choose the FMM leaf width for a grid (the smallest padded square needs at
// least 4 leaf boxes per side); returns 0 when the grid is too small for the
// FMM and the direct sum should be used instead.
int fmm_leaf_for(int xd, int yd) {
    const int side = xd > yd ? xd : yd;
    int s = 1;
    while (s < side) s <<= 1;
    if (s >= 16) return 4;
    if (s >= 8) return 2;
    return 0;
}

void set_solver(mms_params &p, int use_fmm, int P) {
    if (use_fmm) {
        const int w = fmm_leaf_for(p.xdim, p.ydim);
        if (w > 0) {
            p.solver = MMS_SOLVER_FMM;
            p.P = P;
            p.leaf_width = w;
            return;
        }
    }
    p.solver = direct_solver();
    p.P = 0;
}
}  // namespace

extern "C" int calc_potential_exact(const float *charge, int xdim, int ydim, int zdim, float dx, float dy,
                                    float dz, float *potential, int use_gpu, int verbosity) {
    (void)use_gpu;
    std::lock_guard<std::mutex> lk(g_cache_mu);
    timeval t1, t2;
    gettimeofday(&t1, nullptr);
    mms_params p = base_params(xdim, ydim, zdim, dx, dy, dz);
    mms_ctx *c = cached_ctx(p, nullptr);
    if (!c || mms_potential(c, charge, potential)) return EXIT_FAILURE;
    gettimeofday(&t2, nullptr);
    if (verbosity >= 10)
        printf("Exact: took %f seconds\n", (t2.tv_sec + t2.tv_usec / 1e6) - (t1.tv_sec + t1.tv_usec / 1e6));
    return EXIT_SUCCESS;
}

extern "C" int calc_potential_exact_gpu(const float *charge, int xdim, int ydim, int zdim, float dx, float dy,
                                        float dz, float *potential) {
    std::lock_guard<std::mutex> lk(g_cache_mu);
    mms_params p = base_params(xdim, ydim, zdim, dx, dy, dz);
    p.solver = MMS_SOLVER_DIRECT_REFGPU;
    mms_ctx *c = cached_ctx(p, nullptr);
    if (!c || mms_potential(c, charge, potential)) return EXIT_FAILURE;
    return EXIT_SUCCESS;
}

extern "C" int fmm_calc(const float *charge, float *potential, int xdim, int ydim, int zdim, float dx, float dy,
                        float dz, int P, int use_gpu, int verbosity) {
    (void)use_gpu;
    if (P < 1 || P > 9) { set_error("fmm_calc: P must be in 1..9"); return EXIT_FAILURE; }
    std::lock_guard<std::mutex> lk(g_cache_mu);
    timeval t1, t2;
    gettimeofday(&t1, nullptr);
    mms_params p = base_params(xdim, ydim, zdim, dx, dy, dz);
    set_solver(p, 1, P);
    mms_ctx *c = cached_ctx(p, nullptr);
    if (!c || mms_potential(c, charge, potential)) return EXIT_FAILURE;
    gettimeofday(&t2, nullptr);
    if (verbosity >= 10)
        printf("FMM: took %f seconds\n", (t2.tv_sec + t2.tv_usec / 1e6) - (t1.tv_sec + t1.tv_usec / 1e6));
    return EXIT_SUCCESS;
}

static int mms_set_time_nolock(mms_ctx *c, double t) { return mms_set_time(c, t); }

static mms_params field_params(int xdim, int ydim, int zdim, float dx, float dy, float dz, int P, float mu_0,
                               float Ms, float Aexch, int demag, int exchange, int external, int use_fmm) {
    mms_params p = base_params(xdim, ydim, zdim, dx, dy, dz);
    set_solver(p, use_fmm, P);
    p.mu_0 = mu_0; p.Ms = Ms; p.Aexch = Aexch;
    p.demag = demag; p.exchange = exchange;
    p.external = external ? MMS_EXT_RAMP : MMS_EXT_NONE;
    return p;
}

extern "C" int Hfield(const float *M, float *H, float *charge, float *potential, float t, int xdim, int ydim,
                      int zdim, float dx, float dy, float dz, int P, float mu_0, float Ms, float Aexch, int demag,
                      int exchange, int external, int use_fmm, int use_gpu, int verbosity) {
    (void)use_gpu; (void)verbosity;
    if (use_fmm && (P < 1 || P > 9)) { set_error("Hfield: P must be in 1..9"); return EXIT_FAILURE; }
    std::lock_guard<std::mutex> lk(g_cache_mu);
    mms_params p = field_params(xdim, ydim, zdim, dx, dy, dz, P, mu_0, Ms, Aexch, demag, exchange, external, use_fmm);
    mms_ctx *c = cached_ctx(p, nullptr);
    if (!c) return EXIT_FAILURE;
    const long n = c->g.n();
    if (upload_m(c, M) || mms_set_time_nolock(c, (double)t)) return EXIT_FAILURE;
    if (field_now(c, c->m32[c->cur])) return EXIT_FAILURE;
    if (hipMemcpyAsync(H, c->h_aos, sizeof(float) * 3 * n, hipMemcpyDeviceToHost, c->stream) != hipSuccess)
        return EXIT_FAILURE;
    if (demag) {
        if (charge && hipMemcpyAsync(charge, c->rho, sizeof(float) * n, hipMemcpyDeviceToHost, c->stream) != hipSuccess)
            return EXIT_FAILURE;
        if (potential &&
            hipMemcpyAsync(potential, c->phi, sizeof(float) * n, hipMemcpyDeviceToHost, c->stream) != hipSuccess)
            return EXIT_FAILURE;
    }
    if (hipStreamSynchronize(c->stream) != hipSuccess) { set_error("Hfield: stream failure"); return EXIT_FAILURE; }
    return EXIT_SUCCESS;
}

extern "C" int rk4_step(float t, float dt, float *t2, const unsigned char *material, const float *M, float *M2,
                        float *H, float *slope, float *this_slope, float *charge, float *potential, int xdim,
                        int ydim, int zdim, float dx, float dy, float dz, int P, float mu_0, float Ms, float Aexch,
                        float alfa, float gamma, int demag, int exchange, int external, int use_fmm, int normalize,
                        int use_gpu, int verbosity) {
    (void)use_gpu; (void)verbosity;
    if (use_fmm && (P < 1 || P > 9)) { set_error("rk4_step: P must be in 1..9"); return EXIT_FAILURE; }
    if (!material || !M || !M2 || !H) { set_error("rk4_step: null material, M, M2 or H"); return EXIT_FAILURE; }
    std::lock_guard<std::mutex> lk(g_cache_mu);
    mms_params p = field_params(xdim, ydim, zdim, dx, dy, dz, P, mu_0, Ms, Aexch, demag, exchange, external, use_fmm);
    p.alfa = alfa; p.gamma = gamma;
    mms_ctx *c = cached_ctx(p, material);
    if (!c) return EXIT_FAILURE;
    const long n = c->g.n();
    hipStream_t s = c->stream;
    // device staging: aux_aos = [H in | k out | slope out]; the caller's M2
    // (FP32 SoA) goes to the unused FP64 proposal buffer: non-material cells
    // keep it, and this_slope's, as the reference leaves them untouched there
    float *h_in = c->aux_aos, *k_out = c->aux_aos + 3 * n, *s_out = c->aux_aos + 6 * n;
    float *m2_keep = reinterpret_cast<float *>(c->mnew64);
    if (mms_set_time_nolock(c, (double)t)) return EXIT_FAILURE;
    if (hipMemcpyAsync(c->h_aos, M, sizeof(float) * 3 * n, hipMemcpyHostToDevice, s) != hipSuccess ||
        launch_soa_from_aos(c->h_aos, c->m32[0], c->mn64, n, s) ||
        hipMemcpyAsync(c->h_aos, M2, sizeof(float) * 3 * n, hipMemcpyHostToDevice, s) != hipSuccess ||
        launch_soa_from_aos(c->h_aos, m2_keep, nullptr, n, s) ||
        hipMemcpyAsync(h_in, H, sizeof(float) * 3 * n, hipMemcpyHostToDevice, s) != hipSuccess ||
        (this_slope &&
         hipMemcpyAsync(k_out, this_slope, sizeof(float) * 3 * n, hipMemcpyHostToDevice, s) != hipSuccess)) {
        set_error("rk4_step: upload failed");
        return EXIT_FAILURE;
    }
    StageExtra ex{nullptr, nullptr, nullptr, nullptr, normalize, nullptr, c->m32[0], m2_keep};
    const float *in[4] = {c->m32[0], c->m32[1], c->m32[2], c->m32[1]};
    float *outp[4] = {c->m32[1], c->m32[2], c->m32[1], c->m32[2]};
    for (int st = 0; st < 4; st++) {
        ex.h_given = (st == 0) ? h_in : nullptr;
        ex.h_out = (st == 3) ? c->h_aos : nullptr;
        ex.k_out = (st == 3) ? k_out : nullptr;
        ex.slope_out = (st == 3) ? s_out : nullptr;
        if (st > 0 && demag) {
            if (launch_div(c->g, in[st], c->rho, s, &c->prof) || c->potential(c->rho, c->phi)) return EXIT_FAILURE;
        }
        if (launch_stage_ex(c->g, c->mat, st, (double)dt, c->phi, in[st], outp[st], c->mn64, c->acc32, c->matl,
                            c->t_dev, ex, s, &c->prof))
            return EXIT_FAILURE;
    }
    // outputs
    float *m_aos = c->aux_aos;  // h_in no longer needed
    if (launch_aos_from_soa(c->m32[2], m_aos, n, s)) return EXIT_FAILURE;
    bool ok = hipMemcpyAsync(M2, m_aos, sizeof(float) * 3 * n, hipMemcpyDeviceToHost, s) == hipSuccess &&
              hipMemcpyAsync(H, c->h_aos, sizeof(float) * 3 * n, hipMemcpyDeviceToHost, s) == hipSuccess;
    if (ok && slope) ok = hipMemcpyAsync(slope, s_out, sizeof(float) * 3 * n, hipMemcpyDeviceToHost, s) == hipSuccess;
    if (ok && this_slope)
        ok = hipMemcpyAsync(this_slope, k_out, sizeof(float) * 3 * n, hipMemcpyDeviceToHost, s) == hipSuccess;
    if (ok && demag && charge)
        ok = hipMemcpyAsync(charge, c->rho, sizeof(float) * n, hipMemcpyDeviceToHost, s) == hipSuccess;
    if (ok && demag && potential)
        ok = hipMemcpyAsync(potential, c->phi, sizeof(float) * n, hipMemcpyDeviceToHost, s) == hipSuccess;
    if (ok) ok = hipStreamSynchronize(s) == hipSuccess;
    if (!ok) { set_error("rk4_step: download failed"); return EXIT_FAILURE; }
    if (t2) *t2 = t + dt;
    return EXIT_SUCCESS;
}

// ---------------------------------------------------------- time loop --
static int mavg_of(mms_ctx *c, const float *m, double out[5]) {
    if (launch_mavg(m, c->matl, c->g.n(), c->red, c->red_blocks, c->stream)) return 1;
    std::vector<double> part(5 * c->red_blocks);
    MMS_HIP(hipMemcpyAsync(part.data(), c->red, sizeof(double) * 5 * c->red_blocks, hipMemcpyDeviceToHost, c->stream));
    MMS_HIP(hipStreamSynchronize(c->stream));
    for (int q = 0; q < 5; q++) out[q] = 0;
    for (int i = 0; i < c->red_blocks; i++)
        for (int q = 0; q < 5; q++) out[q] += part[5 * i + q];
    return 0;
}

static int write_vector3d(FILE *fh, const float *M, int zd, int yd, int xd) {
    for (int z = 0; z < zd; z++)
        for (int x = 0; x < xd; x++)
            for (int y = 0; y < yd; y++) {
                const float *v = M + 3 * ((long)z * yd * xd + (long)y * xd + x);
                fprintf(fh, "%g %g %g \n", v[0], v[1], v[2]);
            }
    return 0;
}

extern "C" int time_marching(unsigned char *material, float *M, float finaltime, float timestep, int xdim, int ydim,
                             int zdim, float dx, float dy, float dz, int P, float mu_0, float Ms, float Aexch,
                             float alfa, float gamma, int demag, int exchange, int external, int use_fmm,
                             int use_gpu, const char *sim_name, int verbosity) {
    (void)use_gpu;
    if (use_fmm && (P < 1 || P > 9)) { set_error("time_marching: P must be in 1..9"); return EXIT_FAILURE; }
    mms_params p = field_params(xdim, ydim, zdim, dx, dy, dz, P, mu_0, Ms, Aexch, demag, exchange, external, use_fmm);
    p.alfa = alfa; p.gamma = gamma;
    mms_ctx *c = nullptr;
    if (mms_create(&c, &p, material)) return EXIT_FAILURE;
    std::unique_ptr<mms_ctx, int (*)(mms_ctx *)> guard(c, mms_destroy);
    const long n = c->g.n();
    hipStream_t s = c->stream;
    char fname[1100];
    const char *dir = sim_name ? sim_name : ".";
    mkdir(dir, 0755);
    snprintf(fname, sizeof fname, "%s/%s", dir, "dynamics.dat");
    FILE *fh = fopen(fname, "w");
    snprintf(fname, sizeof fname, "%s/%s", dir, "Mdynamics.dat");
    FILE *fhM = fopen(fname, "w");
    snprintf(fname, sizeof fname, "%s/%s", dir, "panic.dat");
    FILE *fhp = fopen(fname, "w");
    if (!fh || !fhM || !fhp) {
        if (fh) fclose(fh);
        if (fhM) fclose(fhM);
        if (fhp) fclose(fhp);
        set_error(std::string("time_marching: cannot open output files in ") + dir);
        return EXIT_FAILURE;
    }
    int status = 0;
    if (upload_m(c, M)) status = 1;
    // buffers: m32[cur] = M_n (f32), mn64 = M_n (f64); h_n = the field H of
    // the reference loop (at M_n, evaluated at the time of the step that
    // produced M_n), h_2 = the post-step field H2 at the proposal.  Both are
    // AoS; aux_aos[0, 3n) stays free as the download staging of mms_get_m.
    float *h_n = c->aux_aos + 3 * n, *h_2 = c->h_aos;
    double dt = timestep, t = 0.0;
    int tindex = 0, consecutive_error_count = 0;
    bool accepted = false;
    std::vector<float> hostM(3 * n);
    auto field_into = [&](const float *m, float *h_out) -> int {
        if (demag) {
            if (launch_div(c->g, m, c->rho, s, &c->prof) || c->potential(c->rho, c->phi)) return 1;
        }
        return launch_hfield(c->g, c->mat, c->phi, m, c->t_dev, h_out, s, &c->prof);
    };
    float E = 0;
    double Ed = 0, torque0 = 0;
    if (!status && mms_set_time_nolock(c, 0.0)) status = 1;
    if (!status && field_into(c->m32[c->cur], h_n)) status = 1;
    // the reference's proposal M2 starts at zero and is never written on
    // non-material cells (ode_functions.cpp:165,102), so every stage and
    // accepted state holds 0 there; the initial field above used M as given
    if (!status && !c->all_magnetic && launch_zero_nonmagnetic(c->m32[c->cur], c->matl, n, s)) status = 1;
    while (!status && t <= finaltime) {
        if (mms_set_time_nolock(c, t)) { status = 1; break; }
        if (energy_torque_of(c, c->m32[c->cur], h_n, &Ed, &torque0)) { status = 1; break; }
        E = (float)Ed;
        // RK4 step: stages read/write m32 slots other than cur; stage 3
        // proposes mnew64; k1 takes the field H as given (H3 = H,
        // ode_functions.cpp:246-247), not recomputed at the current time
        const int a = c->cur, b = (c->cur + 1) % 3, d = (c->cur + 2) % 3;
        const float *in[4] = {c->m32[a], c->m32[b], c->m32[d], c->m32[b]};
        float *outp[4] = {c->m32[b], c->m32[d], c->m32[b], c->m32[d]};
        StageExtra ex{nullptr, nullptr, nullptr, nullptr, 1, c->mnew64, c->m32[a], nullptr};
        for (int st = 0; st < 4 && !status; st++) {
            ex.h_given = (st == 0) ? h_n : nullptr;
            if (st > 0 && demag) {
                if (launch_div(c->g, in[st], c->rho, s, &c->prof) || c->potential(c->rho, c->phi)) { status = 1; break; }
            }
            if (launch_stage_ex(c->g, c->mat, st, dt, c->phi, in[st], outp[st], c->mn64, c->acc32, c->matl, c->t_dev,
                                ex, s, &c->prof))
                status = 1;
        }
        if (status) break;
        // post-step field H2 at the proposal (m32[d]), at the unadvanced t
        // as in the reference (ode_functions.cpp:267)
        double E2d = 0, torqued = 0;
        if (field_into(c->m32[d], h_2) || energy_torque_of(c, c->m32[d], h_2, &E2d, &torqued)) { status = 1; break; }
        const float E2 = (float)E2d, torque = (float)torqued;
        if (E2 - E > .1) {
            // rejected: M and H stay as they were
            dt = (float)(dt / 2.5);
            printf("PANIC!! energy is increasing, so halving the step.\n");
            fprintf(fhp, "PANIC!! energy is increasing, so halving the step. %d, %g, %g, %g, %g \n", tindex, t, dt,
                    E2, torque);
            consecutive_error_count++;
            if (consecutive_error_count > 50) {
                fprintf(stdout, "PANIC!! Stuck in energy violation errors. Terminating the simulation.\n");
                fprintf(fhp, "PANIC!! Stuck in energy violation errors. Terminating the simulation.\n");
                break;
            }
        } else {
            consecutive_error_count = 0;
            accepted = true;
            const double t2 = (double)((float)t + (float)dt);
            dt = (float)(1e-14 / torque / 2);
            t = t2;
            tindex++;
            c->cur = d;
            std::swap(c->mn64, c->mnew64);
            std::swap(h_n, h_2);
            // log line
            double avg[5];
            if (mavg_of(c, c->m32[c->cur], avg)) { status = 1; break; }
            if (mms_get_m(c, hostM.data())) { status = 1; break; }
            write_vector3d(fhM, hostM.data(), zdim, ydim, xdim);
            const double cnt = avg[4] > 0 ? avg[4] : 1;
            fprintf(fh, "%d, %g, %g, %g, %g, %g, %g, %g, %g \n", tindex, t, dt, E2, avg[0] / cnt / Ms,
                    avg[1] / cnt / Ms, avg[2] / cnt / Ms, avg[3] / cnt / Ms, torque);
            if (verbosity >= 1)
                fprintf(stdout, "%d, %g, %g, %g, %g, %g, %g, %g, %g \n", tindex, t, dt, E2, avg[0] / cnt / Ms,
                        avg[1] / cnt / Ms, avg[2] / cnt / Ms, avg[3] / cnt / Ms, torque);
            if (torque < 1e-3) {
                printf("Torque is too low. Breaking simulation...\n");
                break;
            }
        }
        fflush(NULL);
    }
    // M is overwritten with the last accepted state (unchanged if none was)
    if (!status && accepted && mms_get_m(c, M)) status = 1;
    fclose(fh);
    fclose(fhM);
    fclose(fhp);
    return status ? EXIT_FAILURE : EXIT_SUCCESS;
}
