// surrogate_train.cpp -- training of the learned far-field surrogate: the
// seeded film states of the training set, its (charge, potential) pairs
// from the library's own solvers, and the optimiser loop over the training
// kernels of surrogate.hip.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "mms_internal.hpp"

using namespace mms;

namespace {

// SplitMix64 -> uniform in [0, 1)
struct Draw {
    uint64_t s;
    double u() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return (double)(z >> 11) * (1.0 / 9007199254740992.0);
    }
};

}  // namespace

// Film state `index` of the seeded family (see mms_surrogate_state in the
// public header): mask disc / rectangle / ring, state vortex / domains.
extern "C" int mms_surrogate_state(int xd, int yd, int zd, float Ms, unsigned long long seed, long index, float *M,
                                   unsigned char *material, int *kind) {
    MMS_CHECK(xd > 0 && yd > 0 && zd > 0 && M && index >= 0, "bad argument");
    Draw r{(uint64_t)seed ^ (0xD1B54A32D192ED03ull * (uint64_t)(index + 1))};
    const int shape = std::min(2, (int)(3 * r.u()));
    const int state = std::min(1, (int)(2 * r.u()));
    const double L = std::min(xd, yd);
    double cx, cy, rad = 0, rin = 0, x0 = 0, y0 = 0, w = 0, h = 0;
    if (shape == 1) {
        w = (0.3 + 0.65 * r.u()) * xd;
        h = (0.3 + 0.65 * r.u()) * yd;
        x0 = r.u() * (xd - w);
        y0 = r.u() * (yd - h);
        cx = x0 + 0.5 * w;
        cy = y0 + 0.5 * h;
    } else {
        rad = (0.25 + 0.23 * r.u()) * L;
        cx = rad + r.u() * (xd - 2 * rad);
        cy = rad + r.u() * (yd - 2 * rad);
        if (shape == 2) rin = (0.3 + 0.4 * r.u()) * rad;
    }
    const int chir = r.u() < 0.5 ? -1 : 1, pol = r.u() < 0.5 ? -1 : 1;
    const int D = 8 << std::min(2, (int)(3 * r.u()));
    const int ndx = (xd + D - 1) / D, ndy = (yd + D - 1) / D;
    std::vector<double> theta((size_t)ndx * ndy);
    for (auto &t : theta) t = 2.0 * M_PI * r.u();
    const long plane = (long)xd * yd;
    std::vector<float> m2((size_t)plane * 3, 0.0f);
    std::vector<unsigned char> mk((size_t)plane, 0);
    for (int y = 0; y < yd; y++)
        for (int x = 0; x < xd; x++) {
            const double px = x + 0.5 - cx, py = y + 0.5 - cy, d2 = px * px + py * py;
            bool in;
            if (shape == 1) in = x + 0.5 >= x0 && x + 0.5 < x0 + w && y + 0.5 >= y0 && y + 0.5 < y0 + h;
            else in = d2 <= rad * rad && d2 > rin * rin;
            double m[3];
            if (state == 0) {  // vortex around the shape's centre, core radius 2 cells
                const double mz = pol * std::exp(-d2 / 4.0), s = std::sqrt(std::max(0.0, 1.0 - mz * mz));
                const double ph = std::atan2(py, px);
                m[0] = -chir * s * std::sin(ph);
                m[1] = chir * s * std::cos(ph);
                m[2] = mz;
            } else {           // uniform in-plane angle per D x D domain
                const double t = theta[(size_t)(y / D) * ndx + x / D];
                m[0] = std::cos(t);
                m[1] = std::sin(t);
                m[2] = 0.0;
            }
            double nn = 0;
            for (int k = 0; k < 3; k++) {  // 2% noise, renormalised
                m[k] += 0.02 * (r.u() - 0.5);
                nn += m[k] * m[k];
            }
            nn = 1.0 / std::sqrt(nn);
            const long c = (long)y * xd + x;
            mk[c] = in ? 1 : 0;
            for (int k = 0; k < 3; k++) m2[c * 3 + k] = in ? (float)(Ms * m[k] * nn) : 0.0f;
        }
    for (int z = 0; z < zd; z++) {
        std::memcpy(M + (size_t)z * plane * 3, m2.data(), sizeof(float) * plane * 3);
        if (material) std::memcpy(material + (size_t)z * plane, mk.data(), plane);
    }
    if (kind) *kind = shape * 2 + state;
    return 0;
}

struct mms_sur_trainer {
    int xd = 0, yd = 0, zd = 0, batch = 0;
    float dx = 0, Ms = 0, rscale = 1, oscale = 1;
    long npf = 0;
    hipStream_t stream = nullptr;
    SurrogatePlan sp;
    float *d_m = nullptr, *d_v = nullptr, *d_grad = nullptr;
    long step = 0;
    // training set [cap][npf] and 1 / sum phi^2 per pair
    long npairs = 0, cap = 0;
    float *d_rho = nullptr, *d_phi = nullptr;
    std::vector<float> inv_norm;
    // batch
    float *d_brho = nullptr, *d_bphi = nullptr, *d_bphs = nullptr, *d_binv = nullptr;
    std::vector<float> h_binv;
    mms_ctx *solver = nullptr;
    int solver_P = -1;
    ~mms_sur_trainer() {
        if (solver) mms_destroy(solver);
        sp.release();
        void *bufs[] = {d_m, d_v, d_grad, d_rho, d_phi, d_brho, d_bphi, d_bphs, d_binv};
        for (void *b : bufs)
            if (b) (void)hipFree(b);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

int reserve(mms_sur_trainer *t, long want) {
    if (want <= t->cap) return 0;
    const long cap = std::max(want, 2 * t->cap);
    float *r = nullptr, *p = nullptr;
    MMS_HIP(hipMalloc(&r, sizeof(float) * cap * t->npf));
    MMS_HIP(hipMalloc(&p, sizeof(float) * cap * t->npf));
    if (t->npairs) {
        MMS_HIP(hipMemcpyAsync(r, t->d_rho, sizeof(float) * t->npairs * t->npf, hipMemcpyDeviceToDevice, t->stream));
        MMS_HIP(hipMemcpyAsync(p, t->d_phi, sizeof(float) * t->npairs * t->npf, hipMemcpyDeviceToDevice, t->stream));
        MMS_HIP(hipStreamSynchronize(t->stream));
    }
    if (t->d_rho) (void)hipFree(t->d_rho);
    if (t->d_phi) (void)hipFree(t->d_phi);
    t->d_rho = r;
    t->d_phi = p;
    t->cap = cap;
    return 0;
}

int append_norm(mms_sur_trainer *t, long i) {
    double n2 = 0;
    if (surrogate_sqnorm(t->sp, t->d_phi + i * t->npf, t->npf, &n2, t->stream)) return 1;
    t->inv_norm.push_back(n2 > 0 ? (float)(1.0 / n2) : 0.0f);
    return 0;
}

// gather the batch idx[0..batch) and run the forward pass
int forward_batch(mms_sur_trainer *t, const long *idx) {
    MMS_CHECK(idx, "null batch index");
    const long n = t->npf;
    for (int b = 0; b < t->batch; b++) {
        MMS_CHECK(idx[b] >= 0 && idx[b] < t->npairs, "batch index outside the training set");
        MMS_HIP(hipMemcpyAsync(t->d_brho + b * n, t->d_rho + idx[b] * n, sizeof(float) * n, hipMemcpyDeviceToDevice,
                               t->stream));
        MMS_HIP(hipMemcpyAsync(t->d_bphi + b * n, t->d_phi + idx[b] * n, sizeof(float) * n, hipMemcpyDeviceToDevice,
                               t->stream));
        t->h_binv[b] = t->inv_norm[idx[b]];
    }
    MMS_HIP(hipMemcpyAsync(t->d_binv, t->h_binv.data(), sizeof(float) * t->batch, hipMemcpyHostToDevice, t->stream));
    return surrogate_potential(t->sp, t->d_brho, t->d_bphs, t->rscale, t->oscale, t->stream, nullptr);
}

int gradient(mms_sur_trainer *t, const long *idx, double *loss) {
    if (forward_batch(t, idx)) return 1;
    if (surrogate_backward(t->sp, t->d_brho, t->d_bphs, t->d_bphi, t->d_binv, 2.0f / t->batch, t->rscale, t->oscale,
                           t->d_grad, t->stream))
        return 1;
    float lsum = 0;
    MMS_HIP(hipMemcpyAsync(&lsum, t->sp.d_hsum + SurrogatePlan::kHeadPart - 1, sizeof(float), hipMemcpyDeviceToHost,
                           t->stream));
    MMS_HIP(hipStreamSynchronize(t->stream));
    if (loss) *loss = (double)lsum / t->batch;
    return 0;
}

}  // namespace

extern "C" int mms_surrogate_trainer_create(mms_sur_trainer **out, int xdim, int ydim, int zdim, float dx, float Ms,
                                            int batch, const float *weights) {
    MMS_CHECK(out, "null trainer pointer");
    *out = nullptr;
    MMS_CHECK(xdim >= 1 && ydim >= 1 && zdim >= 1 && batch >= 1 && dx > 0 && Ms > 0, "bad trainer shape");
    std::unique_ptr<mms_sur_trainer> t(new mms_sur_trainer());
    t->xd = xdim; t->yd = ydim; t->zd = zdim; t->batch = batch; t->dx = dx; t->Ms = Ms;
    t->npf = (long)xdim * ydim * zdim;
    // the film's natural units, as mms_surrogate_init with scales <= 0
    t->rscale = dx / Ms;
    t->oscale = Ms * dx;
    MMS_HIP(hipStreamCreateWithFlags(&t->stream, hipStreamNonBlocking));
    Grid g{};
    g.xd = xdim; g.yd = ydim; g.zd = zdim * batch;
    g.dx = g.dy = g.dz = dx;
    g.ay = g.az = 1.0f;
    if (t->sp.init(g, zdim, weights, true, t->stream)) return 1;
    const long np = surrogate_param_count(), nb = t->npf * batch;
    MMS_HIP(hipMalloc(&t->d_m, sizeof(float) * np));
    MMS_HIP(hipMalloc(&t->d_v, sizeof(float) * np));
    MMS_HIP(hipMalloc(&t->d_grad, sizeof(float) * np));
    MMS_HIP(hipMemsetAsync(t->d_m, 0, sizeof(float) * np, t->stream));
    MMS_HIP(hipMemsetAsync(t->d_v, 0, sizeof(float) * np, t->stream));
    MMS_HIP(hipMemsetAsync(t->d_grad, 0, sizeof(float) * np, t->stream));
    MMS_HIP(hipMalloc(&t->d_brho, sizeof(float) * nb));
    MMS_HIP(hipMalloc(&t->d_bphi, sizeof(float) * nb));
    MMS_HIP(hipMalloc(&t->d_bphs, sizeof(float) * nb));
    MMS_HIP(hipMalloc(&t->d_binv, sizeof(float) * batch));
    t->h_binv.assign(batch, 0.0f);
    MMS_HIP(hipStreamSynchronize(t->stream));
    *out = t.release();
    return 0;
}

extern "C" int mms_surrogate_trainer_destroy(mms_sur_trainer *t) {
    delete t;
    return 0;
}

extern "C" int mms_surrogate_trainer_generate(mms_sur_trainer *t, int npairs, unsigned long long seed, int P) {
    MMS_CHECK(t && npairs >= 0, "bad argument");
    if (!t->solver || t->solver_P != P) {
        if (t->solver) mms_destroy(t->solver);
        t->solver = nullptr;
        mms_params p{};
        p.xdim = t->xd; p.ydim = t->yd; p.zdim = t->zd;
        p.dx = p.dy = p.dz = t->dx;
        p.solver = P > 0 ? MMS_SOLVER_FMM : MMS_SOLVER_DIRECT;
        p.P = P > 0 ? P : 4;
        p.m2l_precision = MMS_M2L_FP32;
        p.leaf_width = 4;
        p.mu_0 = 4e-7f * (float)M_PI;
        p.Ms = t->Ms;
        p.Aexch = 1.3e-11f; p.alfa = 0.5f; p.gamma = 2.211e5f;
        p.demag = 1;
        if (mms_create(&t->solver, &p, nullptr)) return 1;
        t->solver_P = P;
    }
    mms_ctx *c = t->solver;
    if (reserve(t, t->npairs + npairs)) return 1;
    std::vector<float> M((size_t)t->npf * 3);
    for (int k = 0; k < npairs; k++) {
        const long i = t->npairs;
        if (mms_surrogate_state(t->xd, t->yd, t->zd, t->Ms, seed, i, M.data(), nullptr, nullptr)) return 1;
        if (mms_set_m(c, M.data())) return 1;
        if (launch_div(c->g, c->m32[c->cur], c->rho, c->stream, nullptr) || c->potential(c->rho, c->phi)) return 1;
        MMS_HIP(hipMemcpyAsync(t->d_rho + i * t->npf, c->rho, sizeof(float) * t->npf, hipMemcpyDeviceToDevice,
                               c->stream));
        MMS_HIP(hipMemcpyAsync(t->d_phi + i * t->npf, c->phi, sizeof(float) * t->npf, hipMemcpyDeviceToDevice,
                               c->stream));
        MMS_HIP(hipStreamSynchronize(c->stream));
        t->npairs++;
        if (append_norm(t, i)) return 1;
    }
    return 0;
}

extern "C" int mms_surrogate_trainer_add(mms_sur_trainer *t, int npairs, const float *charge, const float *potential) {
    MMS_CHECK(t && npairs >= 0 && charge && potential, "bad argument");
    if (reserve(t, t->npairs + npairs)) return 1;
    const long n = t->npf;
    MMS_HIP(hipMemcpyAsync(t->d_rho + t->npairs * n, charge, sizeof(float) * npairs * n, hipMemcpyHostToDevice,
                           t->stream));
    MMS_HIP(hipMemcpyAsync(t->d_phi + t->npairs * n, potential, sizeof(float) * npairs * n, hipMemcpyHostToDevice,
                           t->stream));
    MMS_HIP(hipStreamSynchronize(t->stream));
    for (int k = 0; k < npairs; k++) {
        t->npairs++;
        if (append_norm(t, t->npairs - 1)) return 1;
    }
    return 0;
}

extern "C" long mms_surrogate_trainer_pairs(mms_sur_trainer *t) { return t ? t->npairs : -1; }

extern "C" int mms_surrogate_trainer_get_pair(mms_sur_trainer *t, long i, float *charge, float *potential) {
    MMS_CHECK(t && i >= 0 && i < t->npairs, "pair index outside the training set");
    if (charge)
        MMS_HIP(hipMemcpyAsync(charge, t->d_rho + i * t->npf, sizeof(float) * t->npf, hipMemcpyDeviceToHost, t->stream));
    if (potential)
        MMS_HIP(hipMemcpyAsync(potential, t->d_phi + i * t->npf, sizeof(float) * t->npf, hipMemcpyDeviceToHost,
                               t->stream));
    MMS_HIP(hipStreamSynchronize(t->stream));
    return 0;
}

extern "C" int mms_surrogate_trainer_gradient(mms_sur_trainer *t, const long *idx, float *grad, double *loss) {
    MMS_CHECK(t, "null trainer");
    if (gradient(t, idx, loss)) return 1;
    if (grad) {
        std::vector<float> g((size_t)surrogate_param_count());
        MMS_HIP(hipMemcpy(g.data(), t->d_grad, sizeof(float) * g.size(), hipMemcpyDeviceToHost));
        surrogate_param_to_api(g.data(), grad);
    }
    return 0;
}

extern "C" int mms_surrogate_trainer_step(mms_sur_trainer *t, const long *idx, float lr, double *loss) {
    MMS_CHECK(t && lr >= 0, "bad argument");
    if (gradient(t, idx, loss)) return 1;
    if (surrogate_adam(t->sp, t->d_m, t->d_v, t->d_grad, ++t->step, lr, t->stream)) return 1;
    MMS_HIP(hipStreamSynchronize(t->stream));
    return 0;
}

extern "C" int mms_surrogate_trainer_eval(mms_sur_trainer *t, const long *idx, double *residual) {
    MMS_CHECK(t && residual, "bad argument");
    if (forward_batch(t, idx)) return 1;
    if (surrogate_losses(t->sp, t->d_bphs, t->d_bphi, t->d_binv, 0.0f, t->oscale, false, t->stream)) return 1;
    const int nblk = SurrogatePlan::kHeadBlocks, hp = SurrogatePlan::kHeadPart;
    std::vector<float> h((size_t)t->batch * nblk * hp);
    MMS_HIP(hipMemcpyAsync(h.data(), t->sp.d_hpart, sizeof(float) * h.size(), hipMemcpyDeviceToHost, t->stream));
    MMS_HIP(hipStreamSynchronize(t->stream));
    for (int b = 0; b < t->batch; b++) {
        double s = 0;
        for (int k = 0; k < nblk; k++) s += h[((size_t)b * nblk + k) * hp + hp - 1];
        residual[b] = std::sqrt(s);
    }
    return 0;
}

extern "C" int mms_surrogate_trainer_weights(mms_sur_trainer *t, float *weights) {
    MMS_CHECK(t && weights, "bad argument");
    std::vector<float> p((size_t)surrogate_param_count());
    MMS_HIP(hipStreamSynchronize(t->stream));
    MMS_HIP(hipMemcpy(p.data(), t->sp.d_param, sizeof(float) * p.size(), hipMemcpyDeviceToHost));
    surrogate_param_to_api(p.data(), weights);
    return 0;
}

extern "C" int mms_surrogate_trainer_predict(mms_sur_trainer *t, const long *idx, float *potential) {
    MMS_CHECK(t && potential, "bad argument");
    if (forward_batch(t, idx)) return 1;
    MMS_HIP(hipMemcpyAsync(potential, t->d_bphs, sizeof(float) * t->npf * t->batch, hipMemcpyDeviceToHost, t->stream));
    MMS_HIP(hipStreamSynchronize(t->stream));
    return 0;
}
