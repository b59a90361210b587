// fmm_operators.cpp -- FP64 construction of the regular-grid FMM operators.
//
// Solid harmonics (complex, all |m| <= n), with P_n^m the associated Legendre
// function including the Condon-Shortley phase:
//   R_n^m(r) = r^n P_n^m(cos t) e^{i m p} / (n+m)!       (regular)
//   I_n^m(r) = (n-m)! P_n^m(cos t) e^{i m p} / r^{n+1}   (irregular)
//   R_n^{-m} = (-1)^m conj(R_n^m),  I_n^{-m} = (-1)^m conj(I_n^m)
// They satisfy
//   1/|r - s| = sum_{n,m} conj(R_n^m(s)) I_n^m(r)                 (|s| < |r|)
//   R_n^m(a + b) = sum_{j,k} R_j^k(a) R_{n-j}^{m-k}(b)
//   I_l^m(D + t) = sum_{j,k} (-1)^j conj(R_j^k(t)) I_{l+j}^{m+k}(D)  (|t| < |D|)
// from which, for a box centre c and sources q_i at c + s_i,
//   P2M  M_n^m = sum_i q_i conj(R_n^m(s_i))
//   M2M  M'_n^m = sum_{j,k} conj(R_j^k(d)) M_{n-j}^{m-k}          d = c_child - c_parent
//   M2L  L_j^k = (-1)^{j+k} sum_{l,m} M_l^m I_{l+j}^{m-k}(D)       D = c_local - c_mult
//   L2L  L'_a^b = sum_{j,k} L_j^k R_{j-a}^{k-b}(e)                 e = c_child - c_parent
//   L2P  phi(t) = sum_{j,k} L_j^k R_j^k(t)
// The reference tree-code uses the same addition theorem in its unnormalised
// form (fmm_calc.cu:142-147 P2M, :194-207 M2P with numerics.cu:107-114); the
// M2M/M2L/L2L translations have no counterpart there.
//
// Real storage: sources and targets of one layer lie in the plane of their box
// centre, so only (n+m even) coefficients are nonzero (P2M) or needed (L2P),
// and M_n^{-m} = (-1)^m conj(M_n^m).  A box keeps Re/Im of the (n, m >= 0,
// n+m even) coefficients, Im dropped for m = 0: K = (P+1)(P+2)/2 reals.
// Every operator is real-linear in that basis and is built by applying the
// complex formula to the K unit vectors.  Coefficients are scaled per level
// by the box width w_L (M~_n = M_n / w_L^n, L~_j = L_j w_L^j) so every
// operator is O(1) at every level; distances are in units of dx.
#include <complex>
#include <cmath>
#include <vector>
#include "mms_internal.hpp"

namespace mms {
namespace {

typedef std::complex<double> cd;

struct Coef { int n, m; };

std::vector<Coef> coef_list(int P) {
    std::vector<Coef> v;
    for (int n = 0; n <= P; n++)
        for (int m = 0; m <= n; m++)
            if ((n + m) % 2 == 0) v.push_back({n, m});
    return v;
}

double fact(int k) {
    double f = 1.0;
    for (int i = 2; i <= k; i++) f *= i;
    return f;
}

// P[n][m] (m >= 0) with Condon-Shortley phase, three-term recurrences
void legendre(int nmax, double x, std::vector<double> &P) {
    const int W = nmax + 1;
    P.assign((size_t)W * W, 0.0);
    const double s = std::sqrt(std::max(0.0, 1.0 - x * x));
    P[0] = 1.0;
    for (int m = 1; m <= nmax; m++) P[m * W + m] = -(2 * m - 1) * s * P[(m - 1) * W + (m - 1)];
    for (int m = 0; m < nmax; m++) P[(m + 1) * W + m] = (2 * m + 1) * x * P[m * W + m];
    for (int m = 0; m <= nmax; m++)
        for (int n = m + 2; n <= nmax; n++)
            P[n * W + m] = ((2 * n - 1) * x * P[(n - 1) * W + m] - (n + m - 1) * P[(n - 2) * W + m]) / (n - m);
}

// Full table H[n][m + nmax] for |m| <= n <= nmax; kind 'R' or 'I'
struct Harm {
    int nmax;
    std::vector<cd> v;
    cd &at(int n, int m) { return v[(size_t)n * (2 * nmax + 1) + (m + nmax)]; }
    cd get(int n, int m) const {
        if (n < 0 || n > nmax || m < -n || m > n) return cd(0, 0);
        return v[(size_t)n * (2 * nmax + 1) + (m + nmax)];
    }
};

Harm harmonics(int nmax, const double r[3], char kind) {
    Harm h;
    h.nmax = nmax;
    h.v.assign((size_t)(nmax + 1) * (2 * nmax + 1), cd(0, 0));
    const double rr = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    const double ct = rr > 0 ? r[2] / rr : 1.0;
    const double ph = std::atan2(r[1], r[0]);
    std::vector<double> P;
    legendre(nmax, ct, P);
    const int W = nmax + 1;
    for (int n = 0; n <= nmax; n++)
        for (int m = 0; m <= n; m++) {
            cd e = std::polar(1.0, m * ph);
            cd val;
            if (kind == 'R') val = std::pow(rr, n) * P[n * W + m] * e / fact(n + m);
            else val = fact(n - m) * P[n * W + m] * e / std::pow(rr, n + 1);
            h.at(n, m) = val;
            h.at(n, -m) = ((m & 1) ? -1.0 : 1.0) * std::conj(val);
        }
    return h;
}

// complex coefficient table (|m| <= n <= P) from a real K-vector
Harm from_real(const std::vector<Coef> &lst, int P, const double *v) {
    Harm h;
    h.nmax = P;
    h.v.assign((size_t)(P + 1) * (2 * P + 1), cd(0, 0));
    int i = 0;
    for (const Coef &c : lst) {
        double re = v[i++], im = 0.0;
        if (c.m > 0) im = v[i++];
        h.at(c.n, c.m) = cd(re, im);
        h.at(c.n, -c.m) = ((c.m & 1) ? -1.0 : 1.0) * cd(re, -im);
    }
    return h;
}

void to_real(const std::vector<Coef> &lst, const Harm &h, double *out) {
    int i = 0;
    for (const Coef &c : lst) {
        cd z = h.get(c.n, c.m);
        out[i++] = z.real();
        if (c.m > 0) out[i++] = z.imag();
    }
}

const int CHILD_X[4] = {0, 1, 0, 1};  // child order: Morton, x in bit 0
const int CHILD_Y[4] = {0, 0, 1, 1};

}  // namespace

int fmm_ncoef(int P) { return (P + 1) * (P + 2) / 2; }

int fmm_build_ops(int P, int w, double ay, FmmHostOps &ops) {
    const std::vector<Coef> lst = coef_list(P);
    const int K = fmm_ncoef(P);
    ops.P = P;
    ops.K = K;
    ops.w = w;
    // P2M / L2P at the leaf, scaled by the leaf width w
    ops.p2m.assign((size_t)K * w * w, 0.0);
    ops.l2p.assign((size_t)w * w * K, 0.0);
    std::vector<double> col(K);
    for (int iy = 0; iy < w; iy++)
        for (int ix = 0; ix < w; ix++) {
            const int c = iy * w + ix;
            const double s[3] = {(ix - (w - 1) / 2.0) / w, (iy - (w - 1) / 2.0) * ay / w, 0.0};
            Harm R = harmonics(P, s, 'R');
            int k = 0;
            for (const Coef &cf : lst) {
                cd z = std::conj(R.get(cf.n, cf.m));
                ops.p2m[(size_t)k * w * w + c] = z.real();
                cd rr = R.get(cf.n, cf.m);
                if (cf.m == 0) {
                    ops.l2p[(size_t)c * K + k] = rr.real();
                    k++;
                } else {
                    ops.l2p[(size_t)c * K + k] = 2.0 * rr.real();
                    ops.p2m[(size_t)(k + 1) * w * w + c] = z.imag();
                    ops.l2p[(size_t)c * K + k + 1] = -2.0 * rr.imag();
                    k += 2;
                }
            }
        }
    // M2M and L2L for the 4 child positions (offsets in parent-width units,
    // child width = parent width / 2)
    ops.m2m.assign((size_t)4 * K * K, 0.0);
    ops.l2l.assign((size_t)4 * K * K, 0.0);
    std::vector<double> e(K), out(K);
    for (int ci = 0; ci < 4; ci++) {
        const double d[3] = {(CHILD_X[ci] - 0.5) * 0.5, (CHILD_Y[ci] - 0.5) * 0.5 * ay, 0.0};
        Harm Rd = harmonics(2 * P, d, 'R');
        for (int b = 0; b < K; b++) {
            std::fill(e.begin(), e.end(), 0.0);
            e[b] = 1.0;
            Harm Mc = from_real(lst, P, e.data());
            Harm Mp;
            Mp.nmax = P;
            Mp.v.assign((size_t)(P + 1) * (2 * P + 1), cd(0, 0));
            for (int n = 0; n <= P; n++)
                for (int m = -n; m <= n; m++) {
                    cd s(0, 0);
                    for (int j = 0; j <= n; j++)
                        for (int k = -j; k <= j; k++)
                            s += std::conj(Rd.get(j, k)) * Mc.get(n - j, m - k) * std::pow(0.5, n - j);
                    Mp.at(n, m) = s;
                }
            to_real(lst, Mp, out.data());
            for (int a = 0; a < K; a++) ops.m2m[((size_t)ci * K + a) * K + b] = out[a];
            Harm Lp = from_real(lst, P, e.data());
            Harm Lc;
            Lc.nmax = P;
            Lc.v.assign((size_t)(P + 1) * (2 * P + 1), cd(0, 0));
            for (int a = 0; a <= P; a++)
                for (int bb = -a; bb <= a; bb++) {
                    cd s(0, 0);
                    for (int j = a; j <= P; j++)
                        for (int k = -j; k <= j; k++) s += Lp.get(j, k) * Rd.get(j - a, k - bb);
                    Lc.at(a, bb) = s * std::pow(0.5, a);
                }
            to_real(lst, Lc, out.data());
            for (int a = 0; a < K; a++) ops.l2l[((size_t)ci * K + a) * K + b] = out[a];
        }
    }
    return 0;
}

// M2L operators of one level: [2zd-1][49][K][K], offset slot (oy+3)*7 + (ox+3)
// for the source box at (bx+ox, by+oy) relative to the target box; slots with
// max(|ox|,|oy|) < 2 stay zero.  dz index = (z_target - z_source) + zd - 1.
std::vector<double> FmmHostOps::m2l_level(int level, int Lmax, int zd, double ay, double az) const {
    const std::vector<Coef> lst = coef_list(P);
    const double wl = std::ldexp(1.0, Lmax - level);
    const size_t KK = (size_t)K * K;
    std::vector<double> T((size_t)(2 * zd - 1) * 49 * KK, 0.0);
    std::vector<double> e(K), out(K);
    for (int dzi = 0; dzi < 2 * zd - 1; dzi++) {
        const int dz = dzi - (zd - 1);
        for (int oy = -3; oy <= 3; oy++)
            for (int ox = -3; ox <= 3; ox++) {
                if (std::max(std::abs(ox), std::abs(oy)) < 2) continue;
                const double D[3] = {-(double)ox, -(double)oy * ay, dz * az / wl};
                Harm I = harmonics(2 * P, D, 'I');
                double *blk = &T[((size_t)dzi * 49 + (size_t)(oy + 3) * 7 + (ox + 3)) * KK];
                for (int b = 0; b < K; b++) {
                    std::fill(e.begin(), e.end(), 0.0);
                    e[b] = 1.0;
                    Harm M = from_real(lst, P, e.data());
                    Harm L;
                    L.nmax = P;
                    L.v.assign((size_t)(P + 1) * (2 * P + 1), cd(0, 0));
                    for (const Coef &jk : lst) {
                        cd s(0, 0);
                        for (int l = 0; l <= P; l++)
                            for (int m = -l; m <= l; m++) {
                                cd mv = M.get(l, m);
                                if (mv == cd(0, 0)) continue;
                                s += mv * I.get(l + jk.n, m - jk.m);
                            }
                        L.at(jk.n, jk.m) = (((jk.n + jk.m) & 1) ? -1.0 : 1.0) * s / wl;
                    }
                    to_real(lst, L, out.data());
                    for (int a = 0; a < K; a++) blk[(size_t)a * K + b] = out[a];
                }
            }
    }
    return T;
}

}  // namespace mms
