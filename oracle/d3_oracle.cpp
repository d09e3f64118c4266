// D3 dispersion ORACLE — test infrastructure only.
//
// A plain FP64 C++/OpenMP restatement of the reference's algorithm (LipidL/LASP-D3), used by
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the correctness judge and
// the CPU baseline.  Nothing in the product library links, loads or calls this file.
//
// It enumerates neighbours by brute force over periodic images of wrapped coordinates (the
// reference's ALL_ITERATE scheme, src/d3_kernel.cu:633-667, with the image range of
// src/d3_buffer.cu:63-84) and evaluates every formula literally, including the 25-term C6
// interpolation with its max-exponent shift (src/d3_kernel.cu:709-800).  Where the reference
// has a bug, the oracle follows the physics (decisions listed per function below).
//
// Parity pins (tests/test_oracle.py): closed forms (dimer energy/force/stress, isolated-pair CN,
// single-reference C6, BJ r->0 limit, equilateral ATM angle factor, and mixed-element clusters
// whose atoms have different CNs: per-pair 25-term C6 sums, dE/dCN, a heteronuclear ATM
// triangle's energy and dE/dCN), central finite differences
// of forces and of stress under homogeneous strain, translation/permutation/supercell
// invariance, sum(F) = 0, and the reference's seven property tests (test/d3_test.cpp:73-341).
//
// Units: inputs in Angstrom, cutoffs in bohr; outputs E in eV, F in eV/Angstrom, stress in
// eV/Angstrom^3, per-atom energy in eV, CN dimensionless, dE/dCN in Hartree.
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

constexpr double BOHR = 0.52917726;              // src/d3.cu:50
constexpr double HARTREE_EV = 27.211396641308;   // src/d3.cu:204
constexpr double K1 = 16.0, K3 = 4.0;            // src/d3_kernel.cuh:17-19

// Smooth CN cutoff (the reference's optional SMOOTH_CUTOFF_CN, src/d3_kernel.cu:550-561 and
// 1525-1533): each CN term times s(r) = tanh^3(rc_CN - r), r and rc_CN in bohr.
int g_smooth_cn = 0;

struct Tab {
    const float *c6ab, *r0ab, *rcov, *r2r4;
    int nel, nref, zoff;
    double c6(int sa, int sb, int a, int b, int m) const {
        return c6ab[((((size_t)sa * nel + sb) * nref + a) * nref + b) * 3 + m];
    }
};

struct Sys {
    int64_t n;
    std::vector<double> x;  // wrapped positions, bohr, n*3
    std::vector<int> s;     // table slot per atom
    double H[3][3];         // rows = lattice vectors (bohr)
    double V;               // signed det(H), src/d3_kernel.cu:90-96
    int rep[3];             // images -rep..rep per axis
    double rc2, rccn, rc3;
    int damping;
    double s6, s8, p1, p2;  // s6, s8 (damping-specific), and (sr6, sr8) or (a1, a2)
};

void inv3(const double m[3][3], double o[3][3]) {
    double det = m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) - m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
                 m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
    double id = 1.0 / det;
    o[0][0] = (m[1][1] * m[2][2] - m[1][2] * m[2][1]) * id;
    o[0][1] = (m[0][2] * m[2][1] - m[0][1] * m[2][2]) * id;
    o[0][2] = (m[0][1] * m[1][2] - m[0][2] * m[1][1]) * id;
    o[1][0] = (m[1][2] * m[2][0] - m[1][0] * m[2][2]) * id;
    o[1][1] = (m[0][0] * m[2][2] - m[0][2] * m[2][0]) * id;
    o[1][2] = (m[0][2] * m[1][0] - m[0][0] * m[1][2]) * id;
    o[2][0] = (m[1][0] * m[2][1] - m[1][1] * m[2][0]) * id;
    o[2][1] = (m[0][1] * m[2][0] - m[0][0] * m[2][1]) * id;
    o[2][2] = (m[0][0] * m[1][1] - m[0][1] * m[1][0]) * id;
}

// Enumerates every image (j, T) of every atom with 0 < |r_i - r_j - T|^2 <= rcut^2 (or the
// strict / floored variants selected by the caller).  Wrapped coordinates plus the image
// range rep_k = floor(rc * |b_k|) + 1 cover every pair within rc (src/d3_buffer.cu:63-84 gives
// the same range; the reference does not wrap in ALL_ITERATE mode, quirk Q5, which only
// relabels images and leaves the set of pair vectors unchanged).
template <class F>
void for_images(const Sys &S, int64_t i, double rcut, F &&f) {
    const double rc2 = rcut * rcut;
    const double *xi = &S.x[i * 3];
    for (int t0 = -S.rep[0]; t0 <= S.rep[0]; ++t0)
        for (int t1 = -S.rep[1]; t1 <= S.rep[1]; ++t1)
            for (int t2 = -S.rep[2]; t2 <= S.rep[2]; ++t2) {
                double T[3];
                for (int c = 0; c < 3; ++c) T[c] = t0 * S.H[0][c] + t1 * S.H[1][c] + t2 * S.H[2][c];
                for (int64_t j = 0; j < S.n; ++j) {
                    double d[3] = {xi[0] - S.x[j * 3] - T[0], xi[1] - S.x[j * 3 + 1] - T[1], xi[2] - S.x[j * 3 + 2] - T[2]};
                    double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
                    if (r2 <= rc2) f(j, d, r2);
                }
            }
}

// C6 by Gaussian-weighted reference interpolation and both CN derivatives, literally as
// src/d3_kernel.cu:709-800 (calculate_c6ab_dual): validity iff both CN refs > -1, L_ab shifted
// by the maximum exponent, C6 = Z/W, NaN/Inf -> 0.
void c6_dual(const Tab &t, int sa, int sb, double cna, double cnb, double &c6, double &d1, double &d2) {
    double mx = -1e300;
    for (int a = 0; a < t.nref; ++a)
        for (int b = 0; b < t.nref; ++b) {
            double r1 = t.c6(sa, sb, a, b, 1), r2 = t.c6(sa, sb, a, b, 2);
            if (r1 > -1.0 && r2 > -1.0) {
                double e = -K3 * ((cna - r1) * (cna - r1) + (cnb - r2) * (cnb - r2));
                mx = std::max(mx, e);
            }
        }
    double Z = 0, W = 0, cz1 = 0, cz2 = 0, dw1 = 0, dw2 = 0;
    for (int a = 0; a < t.nref; ++a)
        for (int b = 0; b < t.nref; ++b) {
            double cref = t.c6(sa, sb, a, b, 0), r1 = t.c6(sa, sb, a, b, 1), r2 = t.c6(sa, sb, a, b, 2);
            if (r1 > -1.0 && r2 > -1.0) {
                double L = std::exp(-K3 * ((cna - r1) * (cna - r1) + (cnb - r2) * (cnb - r2)) - mx);
                double l1 = -2.0 * K3 * (cna - r1) * L, l2 = -2.0 * K3 * (cnb - r2) * L;
                Z += cref * L;
                W += L;
                cz1 += cref * l1;
                cz2 += cref * l2;
                dw1 += l1;
                dw2 += l2;
            }
        }
    d1 = (W * W > 0) ? (cz1 * W - Z * dw1) / (W * W) : 0.0;
    d2 = (W * W > 0) ? (cz2 * W - Z * dw2) / (W * W) : 0.0;
    c6 = (W > 0) ? Z / W : 0.0;
    if (!std::isfinite(d1)) d1 = 0;
    if (!std::isfinite(d2)) d2 = 0;
    if (!std::isfinite(c6)) c6 = 0;
}

// Damping factors f6, f8 and their r-derivatives, src/d3_kernel.cu:359-424.
void damping(int type, double r, double R0, double p1, double p2, double &f6, double &f8, double &df6, double &df8) {
    if (type == 0) {  // zero damping: f_n = 1/(1 + 6 (s_r,n R0 / r)^alpha_n), alpha = 14, 16
        double b6 = p1 * R0 / r, b8 = p2 * R0 / r;
        double b6_14 = std::pow(b6, 14), b8_16 = std::pow(b8, 16);
        f6 = 1.0 / (1.0 + 6.0 * b6_14);
        f8 = 1.0 / (1.0 + 6.0 * b8_16);
        df6 = 6.0 * 14.0 * f6 * f6 * b6_14 / r;
        df8 = 6.0 * 16.0 * f8 * f8 * b8_16 / r;
    } else {  // Becke-Johnson: f_n = r^n / (r^n + (a1 R0 + a2)^n)
        double a = p1 * R0 + p2;
        double a6 = std::pow(a, 6), a8 = std::pow(a, 8);
        double r6 = std::pow(r, 6), r8 = std::pow(r, 8);
        f6 = r6 / (r6 + a6);
        f8 = r8 / (r8 + a8);
        df6 = 6.0 * a6 * std::pow(r, 5) / ((r6 + a6) * (r6 + a6));
        df8 = 8.0 * a8 * std::pow(r, 7) / ((r8 + a8) * (r8 + a8));
    }
    if (!std::isfinite(f6)) f6 = 0;
    if (!std::isfinite(f8)) f8 = 0;
    if (!std::isfinite(df6)) df6 = 0;
    if (!std::isfinite(df8)) df8 = 0;
}

struct Out {
    std::vector<double> cn, g, e, f, sig;  // per atom (sig: 9 per atom, summed at the end)
    uint64_t n2b = 0, ncn = 0, ntri = 0;   // ordered-pair / ordered-center counts
};

// CN_i = sum over 0 < r^2 <= rc_CN^2 of 1/(1 + exp(-k1((Rcov_i+Rcov_j)/r - 1))),
// src/d3_kernel.cu:544-574 and 626.
void pass_cn(const Sys &S, const Tab &t, Out &o, const int64_t *centers, int64_t nc) {
    uint64_t cnt = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : cnt)
    for (int64_t q = 0; q < nc; ++q) {
        int64_t i = centers ? centers[q] : q;
        double ri = t.rcov[S.s[i]], acc = 0;
        for_images(S, i, S.rccn, [&](int64_t j, const double *, double r2) {
            if (r2 <= 0.0) return;
            double r = std::sqrt(r2);
            double e = std::exp(-K1 * ((ri + t.rcov[S.s[j]]) / r - 1.0));
            double sm = 1.0;
            if (g_smooth_cn) {
                const double th = std::tanh(S.rccn - r);
                sm = th * th * th;
            }
            acc += sm / (1.0 + e);
            ++cnt;
        });
        o.cn[i] = acc;
    }
    o.ncn += cnt;
}

// Two-body energy, dE/dCN, force and virial over 0 < r^2 < rc^2 (strict),
// src/d3_kernel.cu:902-986 and 997-1197.  Accumulates E+ (positive), F = grad E+.
void pass_2b(const Sys &S, const Tab &t, Out &o, const int64_t *centers, int64_t nc) {
    uint64_t cnt = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : cnt)
    for (int64_t q = 0; q < nc; ++q) {
        int64_t i = centers ? centers[q] : q;
        const int si = S.s[i];
        double E = 0, G = 0, F[3] = {0, 0, 0}, W[9] = {0};
        for_images(S, i, S.rc2, [&](int64_t j, const double *d, double r2) {
            if (!(r2 < S.rc2 * S.rc2 && r2 > 0.0)) return;
            const int sj = S.s[j];
            double c6, dc6i, dc6j;
            c6_dual(t, si, sj, o.cn[i], o.cn[j], c6, dc6i, dc6j);
            double q4 = (double)t.r2r4[si] * (double)t.r2r4[sj];
            double c8 = 3.0 * c6 * q4, dc8i = 3.0 * dc6i * q4;
            double R0 = (S.damping == 0) ? (double)t.r0ab[(size_t)si * t.nel + sj] : std::sqrt(3.0 * q4);
            double r = std::sqrt(r2);
            double f6, f8, df6, df8;
            damping(S.damping, r, R0, S.p1, S.p2, f6, f8, df6, df8);
            double r6 = r2 * r2 * r2, r8 = r6 * r2;
            E += 0.5 * (S.s6 * c6 * f6 / r6 + S.s8 * c8 * f8 / r8);
            G += S.s6 * f6 * dc6i / r6 + S.s8 * f8 * dc8i / r8;
            // phi = (1/r) d/dr [s6 C6 f6 r^-6 + s8 C8 f8 r^-8]
            double phi = S.s6 * c6 * (df6 / r6 - 6.0 * f6 / (r6 * r)) / r + S.s8 * c8 * (df8 / r8 - 8.0 * f8 / (r8 * r)) / r;
            for (int a = 0; a < 3; ++a) {
                F[a] += phi * d[a];
                for (int b = 0; b < 3; ++b) W[a * 3 + b] += -d[a] * phi * d[b] / (2.0 * S.V);
            }
            ++cnt;
        });
        o.e[i] += E;
        o.g[i] += G;
        for (int a = 0; a < 3; ++a) o.f[i * 3 + a] += F[a];
        for (int a = 0; a < 9; ++a) o.sig[i * 9 + a] += W[a];
    }
    o.n2b += cnt;
}

// Axilrod-Teller-Muto three-body term, src/d3_kernel.cu:1199-1509.
// Triangles: every edge with 1e-12 <= r^2 <= rc3^2 (src/d3_kernel.cu:1268,1324).
// The reference enumerates i >= j >= k with scale 1, 1/2, 1/6 for index-degenerate triplets;
// this oracle enumerates, for each centre i, every unordered pair {(j,T),(k,T')} of distinct
// neighbour images, which counts each periodic triangle exactly three times (once per vertex),
// so the centre takes e/3 of the energy (the same per-atom split as the reference), the full
// gradient of e with respect to its own vertex, and its own-vertex part of dE/dCN.  Forces are
// therefore physically consistent also for index-degenerate triplets (the reference omits the
// scale factor in its forces, quirk Q3).  Virial: 1/3 of the triangle's virial per centre.
// The image range includes rc3 (the reference sizes it from max(rc, rc_CN) only, quirk Q4).
void pass_atm(const Sys &S, const Tab &t, Out &o, const int64_t *centers, int64_t nc) {
    uint64_t cnt = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : cnt)
    for (int64_t q = 0; q < nc; ++q) {
        int64_t i = centers ? centers[q] : q;
        const int si = S.s[i];
        struct Nb {
            int64_t j;
            double d[3], r2, c6, dc6;
        };
        std::vector<Nb> nb;
        for_images(S, i, S.rc3, [&](int64_t j, const double *d, double r2) {
            if (r2 < 1e-12) return;
            Nb e;
            e.j = j;
            std::memcpy(e.d, d, sizeof(e.d));
            e.r2 = r2;
            double d2;
            c6_dual(t, si, S.s[j], o.cn[i], o.cn[j], e.c6, e.dc6, d2);
            nb.push_back(e);
        });
        const double rc3sq = S.rc3 * S.rc3;
        double E = 0, G = 0, F[3] = {0, 0, 0}, Wv[9] = {0};
        for (size_t p = 0; p < nb.size(); ++p)
            for (size_t qq = p + 1; qq < nb.size(); ++qq) {
                const Nb &J = nb[p], &K = nb[qq];
                // r_jk = r_j - r_k = (r_i - d_ij) - (r_i - d_ik) = d_ik - d_ij
                double djk[3] = {K.d[0] - J.d[0], K.d[1] - J.d[1], K.d[2] - J.d[2]};
                double c = djk[0] * djk[0] + djk[1] * djk[1] + djk[2] * djk[2];
                if (c > rc3sq || c < 1e-12) continue;
                double a = J.r2, b = K.r2;
                double c6ij = J.c6, c6ik = K.c6, c6jk, dd1, dd2;
                c6_dual(t, S.s[J.j], S.s[K.j], o.cn[J.j], o.cn[K.j], c6jk, dd1, dd2);
                double C9 = -std::sqrt(c6ij * c6ik * c6jk);
                // angle factor A = (1 + 3 cos_a cos_b cos_c) / (r_ab r_ac r_bc)^3
                double abc = a * b * c, R = std::sqrt(abc);
                double Q = (a + b - c) * (a - b + c) * (-a + b + c);
                double A = 1.0 / (abc * R) + 3.0 * Q / (8.0 * abc * abc * R);
                // dA/da and dA/db (a = r_ij^2, b = r_ik^2, c = r_jk^2)
                double dQa = (a - b + c) * (-a + b + c) + (a + b - c) * (-a + b + c) - (a + b - c) * (a - b + c);
                double dQb = (a - b + c) * (-a + b + c) - (a + b - c) * (-a + b + c) + (a + b - c) * (a - b + c);
                double dQc = -(a - b + c) * (-a + b + c) + (a + b - c) * (-a + b + c) + (a + b - c) * (a - b + c);
                double R5 = abc * abc * R;
                double dAa = -1.5 / (abc * R) / a + 3.0 / 8.0 * (dQa / R5 - 2.5 * Q / R5 / a);
                double dAb = -1.5 / (abc * R) / b + 3.0 / 8.0 * (dQb / R5 - 2.5 * Q / R5 / b);
                double dAc = -1.5 / (abc * R) / c + 3.0 / 8.0 * (dQc / R5 - 2.5 * Q / R5 / c);
                // damping f = 1/(1 + 6 x^16), x = (4/3) (R0ij R0ik R0jk)^(1/3) / (r_ij r_ik r_jk)^(1/3)
                double R0 = std::cbrt((double)t.r0ab[(size_t)si * t.nel + S.s[J.j]] *
                                      (double)t.r0ab[(size_t)si * t.nel + S.s[K.j]] *
                                      (double)t.r0ab[(size_t)S.s[J.j] * t.nel + S.s[K.j]]);
                double rbar = std::pow(abc, 1.0 / 6.0);
                double x = 4.0 / 3.0 * R0 / rbar;
                double x16 = std::pow(x, 16);
                double f = 1.0 / (1.0 + 6.0 * x16);
                double e = C9 * A * f;
                // t_xy = (1/r_xy) de/dr_xy = 2 de/d(r_xy^2);  df/da = 16 x^16 f^2 / a
                double tij = 2.0 * C9 * (dAa * f + A * 16.0 * x16 * f * f / a);
                double tik = 2.0 * C9 * (dAb * f + A * 16.0 * x16 * f * f / b);
                double tjk = 2.0 * C9 * (dAc * f + A * 16.0 * x16 * f * f / c);
                E += e / 3.0;
                // dC9/dCN_i (own vertex) = C9/2 (dC6ij/dCNi / C6ij + dC6ik/dCNi / C6ik)
                double dC9 = 0.0;
                if (c6ij != 0.0) dC9 += 0.5 * C9 * J.dc6 / c6ij;
                if (c6ik != 0.0) dC9 += 0.5 * C9 * K.dc6 / c6ik;
                G += dC9 * A * f;
                for (int u = 0; u < 3; ++u) {
                    F[u] += tij * J.d[u] + tik * K.d[u];
                    for (int v = 0; v < 3; ++v)
                        Wv[u * 3 + v] += -(tij * J.d[u] * J.d[v] + tik * K.d[u] * K.d[v] + tjk * djk[u] * djk[v]) / (3.0 * S.V);
                }
                ++cnt;
            }
        o.e[i] += E;
        o.g[i] += G;
        for (int u = 0; u < 3; ++u) o.f[i * 3 + u] += F[u];
        for (int u = 0; u < 9; ++u) o.sig[i * 9 + u] += Wv[u];
    }
    o.ntri += cnt;
}

// dE/dCN chain rule: F_i += sum_j (g_i + g_j) c'(r)/r d_ij, virial -(g_i+g_j) c'/r d d / (2V),
// over 0 < r^2 <= rc_CN^2, src/d3_kernel.cu:1511-1556 and 1571-1698.
void pass_chain(const Sys &S, const Tab &t, Out &o, const int64_t *centers, int64_t nc) {
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t q = 0; q < nc; ++q) {
        int64_t i = centers ? centers[q] : q;
        double ri = t.rcov[S.s[i]], F[3] = {0, 0, 0}, W[9] = {0};
        for_images(S, i, S.rccn, [&](int64_t j, const double *d, double r2) {
            if (r2 <= 0.0) return;
            double r = std::sqrt(r2), rr = ri + t.rcov[S.s[j]];
            double e = std::exp(-K1 * (rr / r - 1.0));
            double dcn = -K1 * e * rr / ((1.0 + e) * (1.0 + e) * r2 * r);  // c'(r)/r
            if (g_smooth_cn) {  // d(c s)/dr / r = (c'/r) s + c s'/r,  s' = -3 tanh^2 (1 - tanh^2)
                const double th = std::tanh(S.rccn - r);
                dcn = dcn * th * th * th - 3.0 * th * th * (1.0 - th * th) / ((1.0 + e) * r);
            }
            double s = (o.g[i] + o.g[j]) * dcn;
            if (!std::isfinite(s)) s = 0.0;
            for (int a = 0; a < 3; ++a) {
                F[a] += s * d[a];
                for (int b = 0; b < 3; ++b) W[a * 3 + b] += -d[a] * s * d[b] / (2.0 * S.V);
            }
        });
        for (int a = 0; a < 3; ++a) o.f[i * 3 + a] += F[a];
        for (int a = 0; a < 9; ++a) o.sig[i * 9 + a] += W[a];
    }
}

int setup(Sys &S, const double *pos_A, const int32_t *Z, int64_t n, const double *cell_A, const Tab &t, double rc2,
          double rccn, double rc3, int32_t damping, const double *params) {
    S.n = n;
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) S.H[a][b] = cell_A[a * 3 + b] / BOHR;
    S.V = S.H[0][0] * (S.H[1][1] * S.H[2][2] - S.H[1][2] * S.H[2][1]) -
          S.H[0][1] * (S.H[1][0] * S.H[2][2] - S.H[1][2] * S.H[2][0]) +
          S.H[0][2] * (S.H[1][0] * S.H[2][1] - S.H[1][1] * S.H[2][0]);
    if (!(std::fabs(S.V) > 0)) return 3;
    double Hi[3][3];
    inv3(S.H, Hi);
    double rmax = std::max(rc2, std::max(rccn, rc3));
    for (int k = 0; k < 3; ++k) {
        // |b_k| = norm of column k of H^-1 = 1 / perpendicular height k
        double nk = std::sqrt(Hi[0][k] * Hi[0][k] + Hi[1][k] * Hi[1][k] + Hi[2][k] * Hi[2][k]);
        S.rep[k] = (int)std::floor(rmax * nk) + 1;
    }
    S.x.resize(n * 3);
    S.s.resize(n);
    for (int64_t i = 0; i < n; ++i) {
        if (Z[i] < 1 || Z[i] - t.zoff >= t.nel || Z[i] - t.zoff < 0) return 2;
        S.s[i] = Z[i] - t.zoff;
        double r[3], fr[3];
        for (int c = 0; c < 3; ++c) r[c] = pos_A[i * 3 + c] / BOHR;
        for (int k = 0; k < 3; ++k) fr[k] = r[0] * Hi[0][k] + r[1] * Hi[1][k] + r[2] * Hi[2][k];
        for (int c = 0; c < 3; ++c) {
            double v = r[c];
            for (int k = 0; k < 3; ++k) v -= std::floor(fr[k]) * S.H[k][c];
            S.x[i * 3 + c] = v;
        }
    }
    S.rc2 = rc2;
    S.rccn = rccn;
    S.rc3 = rc3;
    S.damping = damping;
    // params = {s6, s8_zero, s8_bj, sr6, sr8, a1, a2}, src/d3_types.h:36-44
    S.s6 = params[0];
    if (damping == 0) {
        S.s8 = params[1];
        S.p1 = params[3];
        S.p2 = params[4];
    } else {
        S.s8 = params[2];
        S.p1 = params[5];
        S.p2 = params[6];
    }
    return 0;
}

}  // namespace

extern "C" {

// Full evaluation.  rc3 <= 0 disables ATM.  Returns 0 on success, 2 for an unsupported element,
// 3 for a singular cell.  counts = {2-body ordered pairs, CN ordered pairs, ATM centre-pairs};
// each unique pair is counted twice and each unique triangle three times.
int d3o_compute(const double *pos_A, const int32_t *Z, int64_t n, const double *cell_A, const float *c6ab,
                const float *r0ab, const float *rcov, const float *r2r4, int32_t nel, int32_t nref, int32_t zoff,
                double rc2, double rccn, double rc3, int32_t damping, const double *params, double *energy,
                double *forces, double *stress, double *cn, double *dedcn, double *atom_energy, uint64_t *counts) {
    Tab t{c6ab, r0ab, rcov, r2r4, nel, nref, zoff};
    Sys S;
    int rc = setup(S, pos_A, Z, n, cell_A, t, rc2, rccn, rc3, damping, params);
    if (rc) return rc;
    Out o;
    o.cn.assign(n, 0);
    o.g.assign(n, 0);
    o.e.assign(n, 0);
    o.f.assign(n * 3, 0);
    o.sig.assign(n * 9, 0);
    pass_cn(S, t, o, nullptr, n);
    if (rc3 > 0) pass_atm(S, t, o, nullptr, n);  // ATM adds to dE/dCN before the chain rule
    pass_2b(S, t, o, nullptr, n);
    pass_chain(S, t, o, nullptr, n);
    double esum = 0, sig[9] = {0};
    for (int64_t i = 0; i < n; ++i) {
        esum += o.e[i];
        for (int a = 0; a < 9; ++a) sig[a] += o.sig[i * 9 + a];
    }
    // E = -Hartree * sum E+, F *= Hartree/bohr, stress *= Hartree/bohr^3 (src/d3.cu:203-219)
    *energy = -HARTREE_EV * esum;
    for (int64_t i = 0; i < n * 3; ++i) forces[i] = o.f[i] * HARTREE_EV / BOHR;
    for (int a = 0; a < 9; ++a) stress[a] = sig[a] * HARTREE_EV / (BOHR * BOHR * BOHR);
    for (int64_t i = 0; i < n; ++i) {
        if (cn) cn[i] = o.cn[i];
        if (dedcn) dedcn[i] = o.g[i];
        if (atom_energy) atom_energy[i] = -HARTREE_EV * o.e[i];
    }
    if (counts) {
        counts[0] = o.n2b;
        counts[1] = o.ncn;
        counts[2] = o.ntri;
    }
    return 0;
}

// CPU baseline timing helper: runs the four passes for the given centre atoms only (CN of the
// others taken from cn_in, dE/dCN of the others from g_in) with nthreads OpenMP threads, and
// returns the wall time in seconds.  Work per centre is what a full evaluation does per atom.
double d3o_time_centers(const double *pos_A, const int32_t *Z, int64_t n, const double *cell_A, const float *c6ab,
                        const float *r0ab, const float *rcov, const float *r2r4, int32_t nel, int32_t nref,
                        int32_t zoff, double rc2, double rccn, double rc3, int32_t damping, const double *params,
                        const int64_t *centers, int64_t nc, const double *cn_in, const double *g_in, int32_t nthreads,
                        uint64_t *counts) {
    Tab t{c6ab, r0ab, rcov, r2r4, nel, nref, zoff};
    Sys S;
    if (setup(S, pos_A, Z, n, cell_A, t, rc2, rccn, rc3, damping, params)) return -1.0;
    Out o;
    o.cn.assign(cn_in, cn_in + n);
    o.g.assign(g_in, g_in + n);
    o.e.assign(n, 0);
    o.f.assign(n * 3, 0);
    o.sig.assign(n * 9, 0);
    int prev = omp_get_max_threads();
    if (nthreads > 0) omp_set_num_threads(nthreads);
    double t0 = omp_get_wtime();
    pass_cn(S, t, o, centers, nc);
    o.cn.assign(cn_in, cn_in + n);  // keep the given CN so the sample sees realistic values
    if (rc3 > 0) pass_atm(S, t, o, centers, nc);
    pass_2b(S, t, o, centers, nc);
    o.g.assign(g_in, g_in + n);
    pass_chain(S, t, o, centers, nc);
    double t1 = omp_get_wtime();
    omp_set_num_threads(prev);
    if (counts) {
        counts[0] = o.n2b;
        counts[1] = o.ncn;
        counts[2] = o.ntri;
    }
    return t1 - t0;
}

// Per-centre results for checks of large cells: for the given centre atoms, CN from the CN pass
// (out_cn), and with the CN of every atom taken from cn_in and its total dE/dCN from g_in, the
// two-body + three-body per-atom energy (eV, out_e), their dE/dCN (Hartree, out_g, before the
// chain rule), and the force on the centre (eV/Angstrom, out_f: two-body, three-body and chain
// rule).  Each is what a full evaluation gives for that atom when cn_in and g_in are its CN and
// dE/dCN; the cost is the centres' share of a full evaluation.
int d3o_centers(const double *pos_A, const int32_t *Z, int64_t n, const double *cell_A, const float *c6ab,
                const float *r0ab, const float *rcov, const float *r2r4, int32_t nel, int32_t nref, int32_t zoff,
                double rc2, double rccn, double rc3, int32_t damping, const double *params, const int64_t *centers,
                int64_t nc, const double *cn_in, const double *g_in, double *out_cn, double *out_e, double *out_g,
                double *out_f) {
    Tab t{c6ab, r0ab, rcov, r2r4, nel, nref, zoff};
    Sys S;
    int rc = setup(S, pos_A, Z, n, cell_A, t, rc2, rccn, rc3, damping, params);
    if (rc) return rc;
    for (int64_t q = 0; q < nc; ++q)
        if (centers[q] < 0 || centers[q] >= n) return 4;
    Out o;
    o.cn.assign(n, 0);
    o.g.assign(n, 0);
    o.e.assign(n, 0);
    o.f.assign(n * 3, 0);
    o.sig.assign(n * 9, 0);
    pass_cn(S, t, o, centers, nc);
    for (int64_t q = 0; q < nc; ++q) out_cn[q] = o.cn[centers[q]];
    o.cn.assign(cn_in, cn_in + n);
    if (rc3 > 0) pass_atm(S, t, o, centers, nc);
    pass_2b(S, t, o, centers, nc);
    for (int64_t q = 0; q < nc; ++q) out_g[q] = o.g[centers[q]];
    o.g.assign(g_in, g_in + n);
    pass_chain(S, t, o, centers, nc);
    for (int64_t q = 0; q < nc; ++q) {
        const int64_t i = centers[q];
        out_e[q] = -HARTREE_EV * o.e[i];
        for (int a = 0; a < 3; ++a) out_f[3 * q + a] = o.f[3 * i + a] * HARTREE_EV / BOHR;
    }
    return 0;
}

int d3o_max_threads(void) { return omp_get_max_threads(); }

// Smooth CN cutoff on (1) or off (0, the reference's default build) for the following calls.
void d3o_set_cn_smooth(int on) { g_smooth_cn = on ? 1 : 0; }

}  // extern "C"
