// batch.cpp -- many independent films advanced together on one GPU.
//
// All films are packed into one set of device arrays (film f occupies cells
// [off_f, off_f + n_f) of every SoA plane; planes are ntot cells apart).  Per
// RK4 stage the batch issues one launch per kernel class over all films:
//   div (tile list) -> direct sum (films below the direct threshold, one block
//   per (film, 256 targets)) -> FMM (one plan per padded power-of-two size
//   class; every FMM kernel takes the film index in blockIdx.z) -> fused
//   field/LLG stage (tile list, per-film external field).
// A whole RK4 step is captured once as a hipGraph and replayed.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cstdlib>
#include <stdio.h>
#include <string.h>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "mms_internal.hpp"

using namespace mms;

struct mms_batch {
    mms_params p;
    Material mat;
    Grid gcommon;  // spacing, zd, ay/az (xd/yd unused)
    int nfilm = 0;
    long ntot = 0;
    std::vector<BatchFilm> films;
    std::vector<int> solver;  // per film: 0 direct, 1 fmm
    hipStream_t stream = nullptr;
    // the direct sum of the small films and each FMM size class are
    // independent: branches 1.. run on their own streams (default priority),
    // forked after the charge kernel and joined before the stage kernel
    std::vector<hipStream_t> s_branch;
    hipEvent_t ev_fork = nullptr;
    std::vector<hipEvent_t> ev_join;
    float *m32[3] = {nullptr, nullptr, nullptr};  // state slot cur + two stage slots
    double *mn64 = nullptr;
    float *acc32 = nullptr;  // FP32 slope accumulator
    float *rho = nullptr, *phi = nullptr, *aos = nullptr;
    uint8_t *matl = nullptr;
    BatchFilm *d_films = nullptr;
    int4 *d_tiles = nullptr;
    int ntiles = 0;
    int2 *d_direct = nullptr;
    int ndirect_items = 0, ndirect_films = 0;
    float *d_dtab = nullptr;  // 1/r tables of the direct films (at their cell offsets), or null
    size_t dtab_lds = 0;      // LDS of the table kernel (largest direct film)
    std::vector<std::unique_ptr<FmmPlan>> fmm;  // one per size class
    std::vector<Grid> fmm_grid;
    std::vector<int> fmm_side;
    int cur = 0;
    hipGraphExec_t graph = nullptr;
    float graph_dt = 0.0f;
    int launches_per_step = 0;
    Profiler prof;
    int one_step(float dt);
};

// films with fewer cells than this use the direct sum, larger ones the FMM
static const long kDirectMaxCells = 4096;

int mms_batch::one_step(float dt) {
    // m32[a] keeps the FP32 copy of M_n through stages 0-2; stage 3 writes M_{n+1} there
    const int a = cur, b = (cur + 1) % 3, d = (cur + 2) % 3;
    const float *in[4] = {m32[a], m32[b], m32[d], m32[b]};
    float *out[4] = {m32[b], m32[d], m32[b], m32[a]};
    const double scale = (double)(gcommon.dx * gcommon.dy * gcommon.dz) / (double)gcommon.dx;
    int launches = 0;
    for (int s = 0; s < 4; s++) {
        if (p.demag) {
            if (launch_div_batch(in[s], rho, d_tiles, ntiles, d_films, ntot, gcommon.dx, gcommon.dy, gcommon.dz,
                                 gcommon.zd, stream, &prof))
                return 1;
            launches++;
            // branches: the FMM classes, then the direct sum; one kernel at a
            // time inside a profiling window
            const bool conc = !s_branch.empty() && !prof.on;
            if (conc) {
                MMS_HIP(hipEventRecord(ev_fork, stream));
                for (hipStream_t sb : s_branch) MMS_HIP(hipStreamWaitEvent(sb, ev_fork, 0));
            }
            int br = 0;
            auto branch = [&]() { hipStream_t sb = (conc && br > 0) ? s_branch[br - 1] : stream; br++; return sb; };
            for (size_t k = 0; k < fmm.size(); k++) {
                if (fmm_potential(fmm_grid[k], *fmm[k], rho, phi, branch(), &prof)) return 1;
                launches += 2 * fmm[k]->Lleaf - 1;
            }
            if (ndirect_items) {
                if (direct_batch_potential(rho, phi, d_direct, ndirect_items, d_films, gcommon.ay, gcommon.az, scale,
                                           d_dtab, dtab_lds, branch(), &prof))
                    return 1;
                launches++;
            }
            if (conc) {
                for (size_t i = 0; i < s_branch.size(); i++) {
                    MMS_HIP(hipEventRecord(ev_join[i], s_branch[i]));
                    MMS_HIP(hipStreamWaitEvent(stream, ev_join[i], 0));
                }
            }
        }
        if (launch_stage_batch(gcommon, mat, s, (double)dt, phi, in[s], out[s], mn64, m32[a], acc32, matl, d_tiles, ntiles,
                               d_films, ntot, stream, &prof))
            return 1;
        launches++;
    }
    launches_per_step = launches;
    return 0;
}

static void batch_free(mms_batch *b) {
    if (b->graph) (void)hipGraphExecDestroy(b->graph);
    void *bufs[] = {b->m32[0], b->m32[1], b->m32[2], b->mn64, b->acc32, b->rho, b->phi, b->aos, b->matl,
                    b->d_films, b->d_tiles, b->d_direct, b->d_dtab};
    for (void *x : bufs)
        if (x) (void)hipFree(x);
    for (auto &f : b->fmm) f->release();
    for (hipStream_t sb : b->s_branch) (void)hipStreamDestroy(sb);
    for (hipEvent_t e : b->ev_join) (void)hipEventDestroy(e);
    if (b->ev_fork) (void)hipEventDestroy(b->ev_fork);
    if (b->stream) (void)hipStreamDestroy(b->stream);
}

static int batch_build(mms_batch *b, const mms_film_spec *specs, const unsigned char *material) {
    const mms_params &p = b->p;
    b->films.resize(b->nfilm);
    b->solver.resize(b->nfilm);
    long off = 0;
    std::vector<int4> tiles;
    std::vector<int2> direct_items;
    std::map<int, std::vector<int>> classes;  // padded side -> films
    for (int f = 0; f < b->nfilm; f++) {
        const mms_film_spec &sp = specs[f];
        MMS_CHECK(sp.xdim > 0 && sp.ydim > 0, "film dimensions must be positive");
        BatchFilm &bf = b->films[f];
        bf.xd = sp.xdim; bf.yd = sp.ydim; bf.zd = p.zdim; bf.pad = 0; bf.off = off;
        bf.hx = sp.hext[0]; bf.hy = sp.hext[1]; bf.hz = sp.hext[2]; bf.pad2 = 0;
        const long n = (long)sp.xdim * sp.ydim * p.zdim;
        off += n;
        for (long c = 0; c < (long)sp.xdim * sp.ydim; c += STENCIL_TILE_COLUMNS) tiles.push_back({f, (int)c, 0, 0});
        int side = 16;
        while (side < sp.xdim || side < sp.ydim) side <<= 1;
        if (p.solver == MMS_SOLVER_DIRECT || n < kDirectMaxCells) {
            b->solver[f] = 0;
            for (int t = 0; t < (int)((n + 255) / 256); t++) direct_items.push_back({f, t});
            b->ndirect_films++;
        } else {
            b->solver[f] = 1;
            classes[side].push_back(f);
        }
    }
    b->ntot = off;
    MMS_CHECK(b->ntot < (1L << 31), "batch too large");
    const long N = b->ntot;
    MMS_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    b->prof.stream = b->stream;
    for (int i = 0; i < 3; i++) MMS_HIP(hipMalloc(&b->m32[i], sizeof(float) * 3 * N));
    MMS_HIP(hipMalloc(&b->mn64, sizeof(double) * 3 * N));
    MMS_HIP(hipMalloc(&b->acc32, sizeof(float) * 3 * N));
    MMS_HIP(hipMalloc(&b->rho, sizeof(float) * N));
    MMS_HIP(hipMalloc(&b->phi, sizeof(float) * N));
    MMS_HIP(hipMalloc(&b->aos, sizeof(float) * 3 * N));
    MMS_HIP(hipMalloc(&b->matl, N));
    MMS_HIP(hipMemset(b->m32[0], 0, sizeof(float) * 3 * N));
    MMS_HIP(hipMemset(b->m32[1], 0, sizeof(float) * 3 * N));
    MMS_HIP(hipMemset(b->m32[2], 0, sizeof(float) * 3 * N));
    MMS_HIP(hipMemset(b->mn64, 0, sizeof(double) * 3 * N));
    MMS_HIP(hipMemset(b->acc32, 0, sizeof(float) * 3 * N));
    MMS_HIP(hipMemset(b->phi, 0, sizeof(float) * N));
    if (material) {
        MMS_HIP(hipMemcpy(b->matl, material, N, hipMemcpyHostToDevice));
    } else {
        MMS_HIP(hipMemset(b->matl, 1, N));
    }
    MMS_HIP(hipMalloc(&b->d_films, sizeof(BatchFilm) * b->nfilm));
    MMS_HIP(hipMemcpy(b->d_films, b->films.data(), sizeof(BatchFilm) * b->nfilm, hipMemcpyHostToDevice));
    {
        // XCD-aware order: workgroup i runs on XCD i mod 8, so list slot i
        // holds logical tile (i mod 8) * per + i / 8 and each XCD walks a
        // contiguous run of tiles (whole films, column runs in order); padding
        // slots carry film -1
        const long nt = (long)tiles.size(), per = (nt + 7) / 8;
        std::vector<int4> order(8 * per);
        for (long i = 0; i < 8 * per; i++) {
            const long t = (i & 7) * per + (i >> 3);
            order[i] = (t < nt) ? tiles[t] : int4{-1, 0, 0, 0};
        }
        tiles.swap(order);
    }
    b->ntiles = (int)tiles.size();
    MMS_HIP(hipMalloc(&b->d_tiles, sizeof(int4) * tiles.size()));
    MMS_HIP(hipMemcpy(b->d_tiles, tiles.data(), sizeof(int4) * tiles.size(), hipMemcpyHostToDevice));
    // the largest films' items first: their workgroups are the longest, and
    // the smaller ones fill in behind them
    std::stable_sort(direct_items.begin(), direct_items.end(), [&](const int2 &u, const int2 &v) {
        const long nu = (long)b->films[u.x].xd * b->films[u.x].yd, nv = (long)b->films[v.x].xd * b->films[v.x].yd;
        return nu > nv;
    });
    b->ndirect_items = (int)direct_items.size();
    if (b->ndirect_items) {
        MMS_HIP(hipMalloc(&b->d_direct, sizeof(int2) * direct_items.size()));
        MMS_HIP(hipMemcpy(b->d_direct, direct_items.data(), sizeof(int2) * direct_items.size(), hipMemcpyHostToDevice));
        // 1/r tables G[|dz|][|dy|][|dx|] of the direct films (FP64, rounded
        // once) when every one fits the table kernel's LDS
        size_t lds = 0;
        for (int f = 0; f < b->nfilm; f++)
            if (b->solver[f] == 0) lds = std::max(lds, direct_table_lds(b->films[f].xd, b->films[f].yd, p.zdim));
        if (lds <= kDirectTableLds) {
            std::vector<float> tab((size_t)N, 0.0f);
            const double ay = b->gcommon.ay, az = b->gcommon.az;
            for (int f = 0; f < b->nfilm; f++) {
                if (b->solver[f] != 0) continue;
                const BatchFilm &bf = b->films[f];
                float *t = tab.data() + bf.off;
                for (int dz = 0; dz < bf.zd; dz++)
                    for (int dy = 0; dy < bf.yd; dy++)
                        for (int dx = 0; dx < bf.xd; dx++) {
                            const double r2 = (double)dx * dx + (dy * ay) * (dy * ay) + (dz * az) * (dz * az);
                            t[((long)dz * bf.yd + dy) * bf.xd + dx] = (dx | dy | dz) ? (float)(1.0 / std::sqrt(r2)) : 0.0f;
                        }
            }
            MMS_HIP(hipMalloc(&b->d_dtab, sizeof(float) * N));
            MMS_HIP(hipMemcpy(b->d_dtab, tab.data(), sizeof(float) * N, hipMemcpyHostToDevice));
            b->dtab_lds = lds;
        }
    }
    for (auto &kv : classes) {
        Grid g = b->gcommon;
        g.xd = g.yd = kv.first;
        std::vector<FilmDesc> fl;
        for (int f : kv.second) fl.push_back({b->films[f].xd, b->films[f].yd, b->films[f].off});
        std::unique_ptr<FmmPlan> plan(new FmmPlan());
        const int w = (p.leaf_width == 8 && kv.first >= 32) ? 8 : 4;
        if (plan->init_batch(g, p.P, w, p.m2l_precision, p.m2l_cols ? p.m2l_cols : 4, fl)) return 1;
        b->fmm.push_back(std::move(plan));
        b->fmm_grid.push_back(g);
        b->fmm_side.push_back(kv.first);
    }
    // one stream per solver branch beyond the first (MMS_GPU_SERIAL=1: all on one)
    const int nbranch = (int)b->fmm.size() + (b->ndirect_items ? 1 : 0);
    const char *serial = std::getenv("MMS_GPU_SERIAL");
    if (nbranch > 1 && !(serial && serial[0] == '1')) {
        MMS_HIP(hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming));
        for (int i = 1; i < nbranch; i++) {
            hipStream_t sb;
            hipEvent_t e;
            MMS_HIP(hipStreamCreateWithFlags(&sb, hipStreamNonBlocking));
            b->s_branch.push_back(sb);
            MMS_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            b->ev_join.push_back(e);
        }
    }
    return 0;
}

extern "C" int mms_batch_create(mms_batch **out, const mms_params *common, int nfilm, const mms_film_spec *films,
                                const unsigned char *material) {
    MMS_CHECK(out && common && films && nfilm > 0, "bad argument");
    *out = nullptr;
    MMS_CHECK(common->zdim > 0 && common->dx > 0 && common->dy > 0 && common->dz > 0, "bad common parameters");
    std::unique_ptr<mms_batch> b(new mms_batch());
    b->p = *common;
    if (b->p.P <= 0) b->p.P = 4;
    b->nfilm = nfilm;
    b->gcommon.xd = b->gcommon.yd = 0;
    b->gcommon.zd = common->zdim;
    b->gcommon.dx = common->dx; b->gcommon.dy = common->dy; b->gcommon.dz = common->dz;
    b->gcommon.ay = common->dy / common->dx; b->gcommon.az = common->dz / common->dx;
    b->mat.mu_0 = common->mu_0; b->mat.Ms = common->Ms; b->mat.Aexch = common->Aexch; b->mat.alfa = common->alfa;
    b->mat.gamma = common->gamma; b->mat.demag = common->demag; b->mat.exchange = common->exchange;
    b->mat.external = MMS_EXT_CONST;
    b->mat.hext[0] = b->mat.hext[1] = b->mat.hext[2] = 0.0f;
    if (batch_build(b.get(), films, material)) { batch_free(b.get()); return 1; }
    *out = b.release();
    return 0;
}

extern "C" int mms_batch_destroy(mms_batch *b) {
    if (!b) return 0;
    batch_free(b);
    delete b;
    return 0;
}

extern "C" int mms_batch_set_m(mms_batch *b, const float *M) {
    MMS_CHECK(b && M, "null argument");
    MMS_HIP(hipMemcpyAsync(b->aos, M, sizeof(float) * 3 * b->ntot, hipMemcpyHostToDevice, b->stream));
    if (launch_soa_from_aos(b->aos, b->m32[b->cur], b->mn64, b->ntot, b->stream)) return 1;
    MMS_HIP(hipStreamSynchronize(b->stream));
    return 0;
}

extern "C" int mms_batch_get_m(mms_batch *b, float *M) {
    MMS_CHECK(b && M, "null argument");
    if (launch_aos_from_soa(b->m32[b->cur], b->aos, b->ntot, b->stream)) return 1;
    MMS_HIP(hipMemcpyAsync(M, b->aos, sizeof(float) * 3 * b->ntot, hipMemcpyDeviceToHost, b->stream));
    MMS_HIP(hipStreamSynchronize(b->stream));
    return 0;
}

extern "C" int mms_batch_steps(mms_batch *b, int nsteps, float dt, int use_graph) {
    MMS_CHECK(b && nsteps >= 0, "bad argument");
    if (!use_graph || b->prof.on) {
        for (int i = 0; i < nsteps; i++)
            if (b->one_step(dt)) return 1;
        return 0;
    }
    if (!b->graph || b->graph_dt != dt) {
        if (b->graph) { (void)hipGraphExecDestroy(b->graph); b->graph = nullptr; }
        hipGraph_t gr;
        MMS_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        int st = b->one_step(dt);
        hipError_t e = hipStreamEndCapture(b->stream, &gr);
        if (st) return 1;
        MMS_HIP(e);
        hipError_t e2 = hipGraphInstantiate(&b->graph, gr, nullptr, nullptr, 0);
        (void)hipGraphDestroy(gr);
        MMS_HIP(e2);
        b->graph_dt = dt;
    }
    for (int i = 0; i < nsteps; i++) MMS_HIP(hipGraphLaunch(b->graph, b->stream));
    return 0;
}

extern "C" int mms_batch_potential(mms_batch *b, const float *charge, float *potential) {
    MMS_CHECK(b && charge && potential, "null argument");
    MMS_HIP(hipMemcpyAsync(b->rho, charge, sizeof(float) * b->ntot, hipMemcpyHostToDevice, b->stream));
    if (b->ndirect_items) {
        const double scale = (double)(b->gcommon.dx * b->gcommon.dy * b->gcommon.dz) / (double)b->gcommon.dx;
        if (direct_batch_potential(b->rho, b->phi, b->d_direct, b->ndirect_items, b->d_films, b->gcommon.ay,
                                   b->gcommon.az, scale, b->d_dtab, b->dtab_lds, b->stream, &b->prof))
            return 1;
    }
    for (size_t k = 0; k < b->fmm.size(); k++)
        if (fmm_potential(b->fmm_grid[k], *b->fmm[k], b->rho, b->phi, b->stream, &b->prof)) return 1;
    MMS_HIP(hipMemcpyAsync(potential, b->phi, sizeof(float) * b->ntot, hipMemcpyDeviceToHost, b->stream));
    MMS_HIP(hipStreamSynchronize(b->stream));
    return 0;
}

extern "C" int mms_batch_sync(mms_batch *b) {
    MMS_CHECK(b, "null batch");
    MMS_HIP(hipStreamSynchronize(b->stream));
    if (b->prof.on) b->prof.flush();
    return 0;
}

extern "C" int mms_batch_info(mms_batch *b, int *n_direct, int *n_fmm, int *n_classes, long *cells,
                              int *launches_per_step) {
    MMS_CHECK(b, "null batch");
    if (n_direct) *n_direct = b->ndirect_films;
    if (n_fmm) *n_fmm = b->nfilm - b->ndirect_films;
    if (n_classes) *n_classes = (int)b->fmm.size();
    if (cells) *cells = b->ntot;
    if (launches_per_step) *launches_per_step = b->launches_per_step;
    return 0;
}

// autotuner for the batch: per FMM size class, leaf width x M2L column tiles
extern "C" int mms_batch_autotune(mms_batch *b, int reps, char *report, int report_len) {
    MMS_CHECK(b, "null batch");
    if (reps < 1) reps = 3;
    if (b->graph) { (void)hipGraphExecDestroy(b->graph); b->graph = nullptr; }
    if (launch_div_batch(b->m32[b->cur], b->rho, b->d_tiles, b->ntiles, b->d_films, b->ntot, b->gcommon.dx,
                         b->gcommon.dy, b->gcommon.dz, b->gcommon.zd, b->stream, nullptr))
        return 1;
    std::string rep;
    char line[200];
    for (size_t k = 0; k < b->fmm.size(); k++) {
        std::vector<FilmDesc> fl;
        for (int f = 0; f < b->nfilm; f++)
            if (b->solver[f] == 1) {
                int side = 16;
                while (side < b->films[f].xd || side < b->films[f].yd) side <<= 1;
                if (side == b->fmm_side[k]) fl.push_back({b->films[f].xd, b->films[f].yd, b->films[f].off});
            }
        int best_w = 4, best_c = 8;
        double best = 1e30;
        for (int w : {4, 8}) {
            if (b->fmm_side[k] / w < 4) continue;
            for (int cols : {4, 8}) {
                if (b->fmm[k]->init_batch(b->fmm_grid[k], b->p.P, w, b->p.m2l_precision, cols, fl)) return 1;
                hipEvent_t e0, e1;
                MMS_HIP(hipEventCreate(&e0));
                MMS_HIP(hipEventCreate(&e1));
                int st = fmm_potential(b->fmm_grid[k], *b->fmm[k], b->rho, b->phi, b->stream, nullptr);
                if (!st && hipEventRecord(e0, b->stream) != hipSuccess) st = 1;
                for (int r = 0; r < reps && !st; r++)
                    st = fmm_potential(b->fmm_grid[k], *b->fmm[k], b->rho, b->phi, b->stream, nullptr);
                if (!st && hipEventRecord(e1, b->stream) != hipSuccess) st = 1;
                float ms = 0;
                if (!st && (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
                    st = 1;
                (void)hipEventDestroy(e0);
                (void)hipEventDestroy(e1);
                if (st) return 1;
                ms /= reps;
                snprintf(line, sizeof line, "class %d (%zu films) leaf=%d m2l_cols=%d: %.4f ms/solve\n",
                         b->fmm_side[k], fl.size(), w, cols, ms);
                rep += line;
                if (ms < best) { best = ms; best_w = w; best_c = cols; }
            }
        }
        if (b->fmm[k]->init_batch(b->fmm_grid[k], b->p.P, best_w, b->p.m2l_precision, best_c, fl)) return 1;
        snprintf(line, sizeof line, "class %d chosen: leaf=%d m2l_cols=%d\n", b->fmm_side[k], best_w, best_c);
        rep += line;
    }
    if (report && report_len > 0) {
        strncpy(report, rep.c_str(), report_len - 1);
        report[report_len - 1] = 0;
    }
    return 0;
}

extern "C" int mms_batch_profile_begin(mms_batch *b) {
    MMS_CHECK(b, "null batch");
    b->prof.flush();
    b->prof.reset();
    b->prof.on = true;
    return 0;
}

extern "C" int mms_batch_profile_end(mms_batch *b, double *ms_per_class, long *launches_per_class) {
    MMS_CHECK(b, "null batch");
    MMS_HIP(hipStreamSynchronize(b->stream));
    b->prof.flush();
    b->prof.on = false;
    for (int k = 0; k < K_COUNT; k++) {
        if (ms_per_class) ms_per_class[k] = b->prof.count[k] ? b->prof.ms[k] / b->prof.count[k] : 0.0;
        if (launches_per_class) launches_per_class[k] = b->prof.count[k];
    }
    return 0;
}
