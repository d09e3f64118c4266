// Learned stand-in for the three-body term on the engine's cell lists: features, network
// (d3_mlp.hip) and the backward pass to forces, virial and dE/dCN.  No counterpart in the
// reference (its only three-body path is atm_kernel, src/d3_kernel.cu:1199-1509).
//
// Units inside the engine: positions in bohr, E+ in Hartree with E = -Hartree E+ (the reference's
// sign convention); the network works in eV, so its energies and gradients enter the per-atom
// arrays multiplied by -1/Hartree.
//
// Neighbour weight of the radial sums: u_ij h_ij with u_ij = sqrt(C6_ij) q_j, q_j = C6_jj^(1/4), the
// D3 C6 at the atoms' CNs in the factorised form C6_ij = y_i[B] . w_j (y_i[B] = C6ref_{A B}^T w_i,
// the centre's rows in LDS, as the two-body pass forms them; k_sur_self computes q and dq/dCN per
// atom), and the edge damping h_ij = (1 + 6 ((4/3) R0_ij / r_ij)^16)^(-1/2), which multiplies the
// cutoff function (fc -> fc h, and its derivative).
//
// k_sur_feat: one wave per centre i.  Neighbours within RCF (every image, from the stencil of the
// fourth cutoff) are compacted into a per-wave LDS queue of (r, u fc(r), class); the queue is then
// read three entries per step with lane l < 63 taking entry l / 21 and radial centre l % 21, so
// that every lane evaluates one Gaussian per step; the three partial sums of each (class, centre)
// are added through LDS at the end.  x_i[0] = CN_i.
//
// k_sur_back: one wave per centre i, lane per queued neighbour j: with G = dE/dx from the
// network, dE/dr_ij = sum_m (G_i[1 + 21 c_j + m] u_ij + G_j[1 + 21 c_i + m] u_ji) phi_m'(r_ij) (x_i and
// x_j both depend on r_ij), gradient of atom i = sum_j dE/dr_ij d_ij / r_ij (no scatter: each
// atom's wave gathers its own), virial = 1/2 sum_j dE/dr_ij d d^T / r (each pair is seen from both
// ends).  dE/dCN_i = G_i[0] + sum_j phi_m (G_i[c_j, m] du_ij/dCN_i + G_j[c_i, m] du_ji/dCN_i) goes into
// g, so the chain-rule pass adds the CN part of the forces and stress.
//
// k_sur_tan: the push-forward u_i = sum_j (dx_i/dr_j) . dir_j of a direction field dir (per A of
// position step) through the features, for force matching: with dir = F - F_ref and w = -2,
// sum_i d(G_i . u_i)/dparams is the gradient of sum |F - F_ref|^2.  k_sur_dcn first forms every
// atom's CN push-forward dCN_i = sum_j c'(r) (d_ij / r) . (dir_i - dir_j) over the CN pair set (as
// k_chain; u_i[0]); then one wave per centre: the radial sums as in k_sur_feat with
// u_ij phi_m'(r) (d_ij / r) . (dir_i - dir_j) + phi_m(r) (du_ij/dCN_i dCN_i + du_ij/dCN_j dCN_j).
#include <hip/hip_runtime.h>

#include "d3_device.hpp"
#include "d3_sur.hpp"

namespace d3 {
namespace {

using namespace sur;

constexpr int SQ = 256;        // per-wave neighbour queue (entries)
constexpr int YMAX = 64 * MAX_NREF;  // C6 rows of one centre (types x references)

// Edge damping h = (1 + 6 ((4/3) R0 / r)^16)^(-1/2) and dh/dr (0 where (4/3) R0 / r >= 100).
__device__ __forceinline__ float edge_h(float r, float R0, float &dh) {
    const float t = fminf((4.0f / 3.0f) * R0 / r, 100.0f);
    const float t2 = t * t, t4 = t2 * t2, t8 = t4 * t4, s = t8 * t8;
    const float h = t < 100.0f ? rsqrtf(1.0f + 6.0f * s) : 0.0f;
    dh = 48.0f * s * h * h * h / r;
    return h;
}

// The centre's C6 rows y[B nr + b] = sum_a C6ref[A][B][a][b] w_i[a] and, if dy, the same with
// dw_i (one wave, rows in LDS).
template <class T>
__device__ void centre_rows(const Params<T> &P, int i, int ti, int lane, float *y, float *dy) {
    const int nr = P.nref, ne = P.ne;
    for (int idx = lane; idx < ne * nr; idx += WAVE) {
        const int B = idx / nr, b = idx % nr;
        T sy = T(0), sdy = T(0);
        for (int a = 0; a < nr; ++a) {
            const T c = P.c6[((ti * ne + B) * nr + a) * nr + b];
            sy += c * P.w[a * P.n + i];
            sdy += c * P.dw[a * P.n + i];
        }
        y[idx] = (float)sy;
        if (dy) dy[idx] = (float)sdy;
    }
}

// C6_ij = y_i[tj] . w_j, and with dw_j (when wanted) dC6_ij/dCN_j.
template <class T>
__device__ __forceinline__ float pair_c6(const Params<T> &P, const float *y, int j, int tj, float *d_j = nullptr) {
    const int nr = P.nref;
    float c = 0.f, dc = 0.f;
    for (int b = 0; b < nr; ++b) {
        c += y[tj * nr + b] * (float)P.w[b * P.n + j];
        if (d_j) dc += y[tj * nr + b] * (float)P.dw[b * P.n + j];
    }
    if (d_j) *d_j = dc;
    return c;
}

// q_i = C6_ii^(1/4) and dq_i/dCN_i per atom.
template <class T>
__global__ void k_sur_self(Params<T> P) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.n) return;
    const int ti = P.type_s[i], nr = P.nref, ne = P.ne;
    T c = T(0), dc = T(0);
    for (int a = 0; a < nr; ++a)
        for (int b = 0; b < nr; ++b) {
            const T cr = P.c6[((ti * ne + ti) * nr + a) * nr + b];
            c += P.w[a * P.n + i] * cr * P.w[b * P.n + i];
            dc += T(2) * P.dw[a * P.n + i] * cr * P.w[b * P.n + i];
        }
    const float cf = (float)c;
    const float q = cf > 0.f ? sqrtf(sqrtf(cf)) : 0.f;
    P.sur_q[i] = q;
    P.sur_q[P.n + i] = cf > 0.f ? 0.25f * q / cf * (float)dc : 0.f;
}

__device__ __forceinline__ float cutf(float r, float &dfc) {
    const float a = (float)(3.14159265358979323846 / RCF) * r;
    dfc = -0.5f * (float)(3.14159265358979323846 / RCF) * __sinf(a);
    return 0.5f * (1.0f + __cosf(a));
}

// Scans the stencil of cut CUT_SUR for centre C and calls flush(qn) whenever the wave's queue
// holds more than SQ - 64 entries and once at the end; push writes one entry.
template <class T, class Push, class Flush>
__device__ void scan_neighbours(const Params<T> &P, const Centre<T> &C, int lane, int &qn, Push push, Flush flush) {
    const T rc2 = T(RCF * RCF);
    const int s0 = C.S->st_off[CUT_SUR], s1 = s0 + C.S->st_cnt[CUT_SUR];
    for (int sb = s0; sb < s1; sb += WAVE) {
        const auto Bn = C.bins(P, sb, s1, T(RCF));
        const int nsb = s1 - sb < WAVE ? s1 - sb : WAVE;
        for (int sk = 0; sk < nsb; ++sk) {
            int j0, j1;
            T ox, oy, oz;
            Bn.get(sk, j0, j1, ox, oy, oz);
            for (int jb = j0; jb < j1; jb += WAVE) {
                const int j = jb + lane;
                bool in = false;
                T dx = T(0), dy = T(0), dz = T(0), r2 = T(0);
                if (j < j1) {
                    dx = ox - P.lx[j];
                    dy = oy - P.ly[j];
                    dz = oz - P.lz[j];
                    r2 = dx * dx + dy * dy + dz * dz;
                    in = r2 <= rc2 && r2 > T(1e-12);
                }
                const uint64_t m = ballot64(in);
                if (m == 0) continue;
                if (in) push(qn + mbcnt64(m), j, (float)dx, (float)dy, (float)dz, (float)xsqrt(r2));
                qn += __popcll(m);
                if (qn > SQ - WAVE) {
                    wave_sync();
                    flush(qn);
                    wave_sync();
                    qn = 0;
                }
            }
        }
    }
    wave_sync();
    flush(qn);
    wave_sync();
    qn = 0;
}

template <class T>
__global__ __launch_bounds__(BLOCK) void k_sur_feat(Params<T> P) {
    __shared__ float Qr[WPB][SQ], Qf[WPB][SQ];
    __shared__ int Qc[WPB][SQ];
    __shared__ float Red[WPB][3][64];
    __shared__ float Y[WPB][YMAX];
    const int lane = lane_id(), wv = threadIdx.x / WAVE;
    const int i = blockIdx.x * WPB + wv;
    if (i >= P.n) return;  // whole wave; no block barrier below
    Centre<T> C;
    C.load(P, i);
    centre_rows(P, i, C.ti, lane, Y[wv], nullptr);
    wave_sync();
    const int m = lane % NRBF, e = lane / NRBF;  // lane 63: e = 3, idle in the sums
    const float mu = (float)MU0 + (float)DMU * m, is2 = (float)(0.5 / (SIGMA * SIGMA));
    float a0 = 0.f, a1 = 0.f, a2 = 0.f;
    auto push = [&](int pos, int j, float, float, float, float r) {
        float dfc;
        const int tj = P.type_s[j];
        const float u = sqrtf(fmaxf(pair_c6(P, Y[wv], j, tj), 0.f)) * P.sur_q[j];
        float dh;
        const float hh = edge_h(r, (float)P.r0[C.ti * P.ne + tj], dh);
        Qr[wv][pos] = r;
        Qf[wv][pos] = cutf(r, dfc) * hh * u;
        Qc[wv][pos] = P.type_cls[tj];
    };
    auto flush = [&](int qn) {
        for (int q = 0; q < qn; q += 3) {
            const int k = q + e;
            if (e < 3 && k < qn) {
                const float u = Qr[wv][k] - mu;
                const float phi = __expf(-u * u * is2) * Qf[wv][k];
                const int c = Qc[wv][k];
                a0 += c == 0 ? phi : 0.f;
                a1 += c == 1 ? phi : 0.f;
                a2 += c == 2 ? phi : 0.f;
            }
        }
    };
    int qn = 0;
    scan_neighbours(P, C, lane, qn, push, flush);
    Red[wv][0][lane] = a0;
    Red[wv][1][lane] = a1;
    Red[wv][2][lane] = a2;
    wave_sync();
    float *x = P.feat + (size_t)i * NFEAT;
    if (lane < 3 * NRBF) {
        const int c = lane / NRBF, mm = lane % NRBF;
        x[1 + lane] = Red[wv][c][mm] + Red[wv][c][NRBF + mm] + Red[wv][c][2 * NRBF + mm];
    } else {
        x[0] = (float)P.cn[i];
        P.sur_cls[i] = P.type_cls[C.ti];
    }
}

template <class T>
__global__ __launch_bounds__(BLOCK) void k_sur_back(Params<T> P) {
    __shared__ float Qx[WPB][SQ], Qy[WPB][SQ], Qz[WPB][SQ], Qr[WPB][SQ];
    __shared__ int Qj[WPB][SQ];
    __shared__ float Gi[WPB][NFEAT];
    __shared__ float Y[WPB][YMAX], DY[WPB][YMAX];
    const int lane = lane_id(), wv = threadIdx.x / WAVE;
    const int i = blockIdx.x * WPB + wv;
    if (i >= P.n) return;
    Centre<T> C;
    C.load(P, i);
    const int ci = P.sur_cls[i];
    Gi[wv][lane] = P.sur_G[(size_t)i * NFEAT + lane];
    centre_rows(P, i, C.ti, lane, Y[wv], DY[wv]);
    wave_sync();
    const float qi = P.sur_q[i], dqi = P.sur_q[P.n + i];
    const float is2 = (float)(0.5 / (SIGMA * SIGMA)), isig2 = (float)(1.0 / (SIGMA * SIGMA));
    Acc<float> g[3], w[6], gc;
    for (int k = 0; k < 3; ++k) g[k].init();
    for (int k = 0; k < 6; ++k) w[k].init();
    gc.init();
    auto push = [&](int pos, int j, float dx, float dy, float dz, float r) {
        Qx[wv][pos] = dx;
        Qy[wv][pos] = dy;
        Qz[wv][pos] = dz;
        Qr[wv][pos] = r;
        Qj[wv][pos] = j;
    };
    auto flush = [&](int qn) {
        for (int q = lane; q < qn; q += WAVE) {
            const int j = Qj[wv][q];
            const float r = Qr[wv][q];
            const int cj = P.sur_cls[j];
            const float *gj = P.sur_G + (size_t)j * NFEAT + 1 + NRBF * ci;
            const float *gi = &Gi[wv][1 + NRBF * cj];
            // u_ij = sqrt(C6_ij) q_j weights x_i, u_ji = sqrt(C6_ij) q_i weights x_j; their CN_i
            // derivatives
            const int tj = P.type_s[j];
            float c6 = 0.f, dc6 = 0.f;
            for (int b = 0; b < P.nref; ++b) {
                const float wb = (float)P.w[b * P.n + j];
                c6 += Y[wv][tj * P.nref + b] * wb;
                dc6 += DY[wv][tj * P.nref + b] * wb;
            }
            const float sc = sqrtf(fmaxf(c6, 0.f)), qj = P.sur_q[j];
            const float isc = sc > 0.f ? 0.5f / sc : 0.f;
            const float uij = sc * qj, uji = sc * qi;
            const float duij = qj * dc6 * isc, duji = qi * dc6 * isc + sc * dqi;
            float dfc0, dh;
            const float fc0 = cutf(r, dfc0), hh = edge_h(r, (float)P.r0[C.ti * P.ne + tj], dh);
            const float fc = fc0 * hh, dfc = dfc0 * hh + fc0 * dh;  // damped cutoff function
            float s = 0.f, A = 0.f, B = 0.f;
            for (int mm = 0; mm < NRBF; ++mm) {
                const float u = r - ((float)MU0 + (float)DMU * mm);
                const float ga = __expf(-u * u * is2);
                s += (gi[mm] * uij + gj[mm] * uji) * ga * (dfc - fc * u * isig2);
                A += gi[mm] * ga;
                B += gj[mm] * ga;
            }
            gc.add(fc * (A * duij + B * duji));
            const float dx = Qx[wv][q], dy = Qy[wv][q], dz = Qz[wv][q];
            const float sr = s / r;
            g[0].add(sr * dx);
            g[1].add(sr * dy);
            g[2].add(sr * dz);
            const float hs = 0.5f * sr;
            w[0].add(hs * dx * dx);
            w[1].add(hs * dy * dy);
            w[2].add(hs * dz * dz);
            w[3].add(hs * dx * dy);
            w[4].add(hs * dx * dz);
            w[5].add(hs * dy * dz);
        }
    };
    int qn = 0;
    scan_neighbours(P, C, lane, qn, push, flush);
    double gt[3], wt[6];
    for (int k = 0; k < 3; ++k) gt[k] = wave_total(g[k]);
    for (int k = 0; k < 6; ++k) wt[k] = wave_total(w[k]);
    const double gcn = wave_total(gc);
    if (lane == 0) {
        const double h = -1.0 / HARTREE_EV;  // eV -> E+ (Hartree, opposite sign)
        for (int k = 0; k < 3; ++k) P.fa[3 * i + k] = h * gt[k];
        P.ga[i] = h * ((double)Gi[wv][0] + gcn);
        const double ei = h * (double)P.sur_E[i];
        P.eas[i] = ei;
        P.eat[i] = ei;
        // vir holds -(1/V) dE+/d(strain) = (1/(V Hartree)) dE/d(strain), and wt = dE/d(strain)
        const double iv = C.S->inv_vol;
        for (int k = 0; k < 6; ++k) P.vir[6 * i + k] -= h * wt[k] * iv;
        P.npair[3 * i + 2] = 0.f;
    }
}

// CN push-forward of every atom along dir (as k_chain walks the CN pairs): sur_dcn[i] =
// sum_j c'(r)/r d_ij . (dir_i - dir_j), per bohr of distance times dir's units.
template <class T>
__global__ __launch_bounds__(BLOCK) void k_sur_dcn(Params<T> P, const float *__restrict__ dir) {
    const int lane = lane_id(), wv = threadIdx.x / WAVE;
    const int i = blockIdx.x * WPB + wv;
    if (i >= P.n) return;  // whole wave
    Centre<T> C;
    C.load(P, i);
    const float di0 = dir[3 * i], di1 = dir[3 * i + 1], di2 = dir[3 * i + 2];
    const T rci = P.rcov[C.ti];
    float cacc = 0.f;
    const int s0 = C.S->st_off[CUT_CN], s1 = s0 + C.S->st_cnt[CUT_CN];
    for (int sb = s0; sb < s1; sb += WAVE) {
        const auto Bn = C.bins(P, sb, s1, P.rc_cn);
        const int nsb = s1 - sb < WAVE ? s1 - sb : WAVE;
        for (int sk = 0; sk < nsb; ++sk) {
            int j0, j1;
            T ox, oy, oz;
            Bn.get(sk, j0, j1, ox, oy, oz);
            for (int j = j0 + lane; j < j1; j += WAVE) {
                const T dx = ox - P.lx[j], dy = oy - P.ly[j], dz = oz - P.lz[j];
                const T r2 = dx * dx + dy * dy + dz * dz;
                if (r2 <= P.rc_cn2 && r2 > T(0)) {
                    const T ir = xrsqrt(r2);
                    const T rr = rci + P.rcov[P.type_s[j]];
                    const T ex = xexp(-T(K1) * (rr * ir - T(1)));
                    const T ope = T(1) + ex;
                    T dcn = -T(K1) * ex * rr * ir * ir * ir / (ope * ope);  // c'(r) / r
                    if (P.cn_smooth) {
                        T c = T(1) / ope;
                        smooth_cn_term(P.rc_cn, r2 * ir, ir, c, dcn);
                    }
                    dcn = finite_or_zero(dcn);
                    cacc += (float)dcn * ((float)dx * (di0 - dir[3 * j]) + (float)dy * (di1 - dir[3 * j + 1]) +
                                          (float)dz * (di2 - dir[3 * j + 2]));
                }
            }
        }
    }
    cacc = wave_sumf(cacc);
    if (lane == 0) P.sur_dcn[i] = cacc;
}

template <class T>
__global__ __launch_bounds__(BLOCK) void k_sur_tan(Params<T> P, const float *__restrict__ dir, float *__restrict__ U) {
    __shared__ float Qr[WPB][SQ], Qp[WPB][SQ], Qt[WPB][SQ], Qfc[WPB][SQ], Qdf[WPB][SQ];
    __shared__ int Qc[WPB][SQ];
    __shared__ float Red[WPB][3][64];
    __shared__ float Y[WPB][YMAX], DY[WPB][YMAX];
    const int lane = lane_id(), wv = threadIdx.x / WAVE;
    const int i = blockIdx.x * WPB + wv;
    if (i >= P.n) return;  // whole wave; no block barrier below
    Centre<T> C;
    C.load(P, i);
    centre_rows(P, i, C.ti, lane, Y[wv], DY[wv]);
    wave_sync();
    const float di0 = dir[3 * i], di1 = dir[3 * i + 1], di2 = dir[3 * i + 2];
    const float dcni = P.sur_dcn[i];
    const int m = lane % NRBF, e = lane / NRBF;  // lane 63: e = 3, idle in the sums
    const float mu = (float)MU0 + (float)DMU * m, is2 = (float)(0.5 / (SIGMA * SIGMA)),
                isig2 = (float)(1.0 / (SIGMA * SIGMA));
    float a0 = 0.f, a1 = 0.f, a2 = 0.f;
    auto push = [&](int pos, int j, float dx, float dy, float dz, float r) {
        // u_ij and its change through CN_i and CN_j along dir
        const int tj = P.type_s[j];
        float c6 = 0.f, dci = 0.f, dcj = 0.f;
        for (int b = 0; b < P.nref; ++b) {
            const float wb = (float)P.w[b * P.n + j], dwb = (float)P.dw[b * P.n + j];
            c6 += Y[wv][tj * P.nref + b] * wb;
            dci += DY[wv][tj * P.nref + b] * wb;
            dcj += Y[wv][tj * P.nref + b] * dwb;
        }
        const float sc = sqrtf(fmaxf(c6, 0.f)), qj = P.sur_q[j], dqj = P.sur_q[P.n + j];
        const float isc = sc > 0.f ? 0.5f / sc : 0.f;
        float dfc0, dh;
        const float fc0 = cutf(r, dfc0), hh = edge_h(r, (float)P.r0[C.ti * P.ne + tj], dh);
        Qfc[wv][pos] = fc0 * hh;
        Qdf[wv][pos] = dfc0 * hh + fc0 * dh;
        Qr[wv][pos] = r;
        Qp[wv][pos] = sc * qj * (dx * (di0 - dir[3 * j]) + dy * (di1 - dir[3 * j + 1]) + dz * (di2 - dir[3 * j + 2])) / r;
        Qt[wv][pos] = qj * dci * isc * dcni + (qj * dcj * isc + sc * dqj) * P.sur_dcn[j];
        Qc[wv][pos] = P.type_cls[tj];
    };
    auto flush = [&](int qn) {
        for (int q = 0; q < qn; q += 3) {
            const int k = q + e;
            if (e < 3 && k < qn) {
                const float r = Qr[wv][k], u = r - mu;
                const float fc = Qfc[wv][k], dfc = Qdf[wv][k];
                const float v = __expf(-u * u * is2) * ((dfc - fc * u * isig2) * Qp[wv][k] + fc * Qt[wv][k]);
                const int c = Qc[wv][k];
                a0 += c == 0 ? v : 0.f;
                a1 += c == 1 ? v : 0.f;
                a2 += c == 2 ? v : 0.f;
            }
        }
    };
    int qn = 0;
    scan_neighbours(P, C, lane, qn, push, flush);
    Red[wv][0][lane] = a0;
    Red[wv][1][lane] = a1;
    Red[wv][2][lane] = a2;
    wave_sync();
    // per A of step: the features are functions of bohr distances
    const float ib = (float)(1.0 / sur::BOHR_A);
    float *uo = U + (size_t)i * NFEAT;
    if (lane < 3 * NRBF) {
        const int c = lane / NRBF, mm = lane % NRBF;
        uo[1 + lane] = ib * (Red[wv][c][mm] + Red[wv][c][NRBF + mm] + Red[wv][c][2 * NRBF + mm]);
    } else {
        uo[0] = ib * dcni;
    }
}

}  // namespace

template <class T>
void launch_sur_tangent(const Params<T> &P, const float *dir, float *U, hipStream_t s) {
    if (P.n == 0) return;
    hipLaunchKernelGGL(k_sur_self<T>, dim3((P.n + 255) / 256), dim3(256), 0, s, P);
    hipLaunchKernelGGL(k_sur_dcn<T>, dim3((P.n + WPB - 1) / WPB), dim3(BLOCK), 0, s, P, dir);
    hipLaunchKernelGGL(k_sur_tan<T>, dim3((P.n + WPB - 1) / WPB), dim3(BLOCK), 0, s, P, dir, U);
}

template <class T>
void launch_sur_features(const Params<T> &P, hipStream_t s) {
    if (P.n == 0) return;
    hipLaunchKernelGGL(k_sur_self<T>, dim3((P.n + 255) / 256), dim3(256), 0, s, P);
    hipLaunchKernelGGL(k_sur_feat<T>, dim3((P.n + WPB - 1) / WPB), dim3(BLOCK), 0, s, P);
}

template <class T>
void launch_surrogate(const Params<T> &P, hipStream_t s) {
    if (P.n == 0) return;
    launch_sur_features<T>(P, s);
    sur::launch_infer(P.sur_P, P.sur_PK, P.n, P.feat, P.sur_cls, P.sur_E, P.sur_G, s);
    hipLaunchKernelGGL(k_sur_back<T>, dim3((P.n + WPB - 1) / WPB), dim3(BLOCK), 0, s, P);
    launch_atm_merge<T>(P, s);
}

template void launch_sur_features<double>(const Params<double> &, hipStream_t);
template void launch_sur_features<float>(const Params<float> &, hipStream_t);
template void launch_sur_tangent<double>(const Params<double> &, const float *, float *, hipStream_t);
template void launch_sur_tangent<float>(const Params<float> &, const float *, float *, hipStream_t);
template void launch_surrogate<double>(const Params<double> &, hipStream_t);
template void launch_surrogate<float>(const Params<float> &, hipStream_t);

}  // namespace d3
