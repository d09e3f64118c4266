// Learned stand-in for the three-body term on the engine's cell lists: features, network
// (d3_mlp.hip) and the backward pass to forces, virial and dE/dCN.  No counterpart in the
// reference (its only three-body path is atm_kernel, src/d3_kernel.cu:1199-1509).
//
// Units inside the engine: positions in bohr, E+ in Hartree with E = -Hartree E+ (the reference's
// sign convention); the network works in eV, so its energies and gradients enter the per-atom
// arrays multiplied by -1/Hartree.
//
// k_sur_feat: one wave per centre i.  Neighbours within RCF (every image, from the stencil of the
// fourth cutoff) are compacted into a per-wave LDS queue of (r, fc(r), class); the queue is then
// read three entries per step with lane l < 63 taking entry l / 21 and radial centre l % 21, so
// that every lane evaluates one Gaussian per step; the three partial sums of each (class, centre)
// are added through LDS at the end.  x_i[0] = CN_i.
//
// k_sur_back: one wave per centre i, lane per queued neighbour j: with G = dE/dx from the
// network, dE/dr_ij = sum_m (G_i[1 + 21 c_j + m] + G_j[1 + 21 c_i + m]) phi_m'(r_ij) (x_i and x_j
// both depend on r_ij), gradient of atom i = sum_j dE/dr_ij d_ij / r_ij (no scatter: each atom's
// wave gathers its own), virial = 1/2 sum_j dE/dr_ij d d^T / r (each pair is seen from both ends).
// The CN part of the forces goes through the chain-rule pass: dE/dCN_i = G_i[0] is added to g.
//
// k_sur_tan: the push-forward u_i = sum_j (dx_i/dr_j) . dir_j of a direction field dir (per A of
// position step) through the features, for force matching: with dir = F - F_ref and w = -2,
// sum_i d(G_i . u_i)/dparams is the gradient of sum |F - F_ref|^2.  One wave per centre: the
// radial sums as in k_sur_feat with phi_m'(r) (d_ij / r) . (dir_i - dir_j) in place of phi_m(r),
// then u_i[0] = sum_j c'(r) (d_ij / r) . (dir_i - dir_j) over the CN pair set (as k_chain).
#include <hip/hip_runtime.h>

#include "d3_device.hpp"
#include "d3_sur.hpp"

namespace d3 {
namespace {

using namespace sur;

constexpr int SQ = 256;  // per-wave neighbour queue (entries)

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
        const auto Bn = C.bins(P, sb, s1);
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
    const int lane = lane_id(), wv = threadIdx.x / WAVE;
    const int i = blockIdx.x * WPB + wv;
    if (i >= P.n) return;  // whole wave; no block barrier below
    Centre<T> C;
    C.load(P, i);
    const int m = lane % NRBF, e = lane / NRBF;  // lane 63: e = 3, idle in the sums
    const float mu = (float)MU0 + (float)DMU * m, is2 = (float)(0.5 / (SIGMA * SIGMA));
    float a0 = 0.f, a1 = 0.f, a2 = 0.f;
    auto push = [&](int pos, int j, float, float, float, float r) {
        float dfc;
        Qr[wv][pos] = r;
        Qf[wv][pos] = cutf(r, dfc);
        Qc[wv][pos] = P.type_cls[P.type_s[j]];
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
    const int lane = lane_id(), wv = threadIdx.x / WAVE;
    const int i = blockIdx.x * WPB + wv;
    if (i >= P.n) return;
    Centre<T> C;
    C.load(P, i);
    const int ci = P.sur_cls[i];
    Gi[wv][lane] = P.sur_G[(size_t)i * NFEAT + lane];
    wave_sync();
    const float is2 = (float)(0.5 / (SIGMA * SIGMA)), isig2 = (float)(1.0 / (SIGMA * SIGMA));
    Acc<float> g[3], w[6];
    for (int k = 0; k < 3; ++k) g[k].init();
    for (int k = 0; k < 6; ++k) w[k].init();
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
            float dfc;
            const float fc = cutf(r, dfc);
            float s = 0.f;
            for (int mm = 0; mm < NRBF; ++mm) {
                const float u = r - ((float)MU0 + (float)DMU * mm);
                const float ga = __expf(-u * u * is2);
                s += (gi[mm] + gj[mm]) * ga * (dfc - fc * u * isig2);
            }
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
    if (lane == 0) {
        const double h = -1.0 / HARTREE_EV;  // eV -> E+ (Hartree, opposite sign)
        for (int k = 0; k < 3; ++k) P.fa[3 * i + k] = h * gt[k];
        P.ga[i] = h * (double)Gi[wv][0];
        const double ei = h * (double)P.sur_E[i];
        P.eas[i] = ei;
        P.eat[i] = ei;
        // vir holds -(1/V) dE+/d(strain) = (1/(V Hartree)) dE/d(strain), and wt = dE/d(strain)
        const double iv = C.S->inv_vol;
        for (int k = 0; k < 6; ++k) P.vir[6 * i + k] -= h * wt[k] * iv;
        P.npair[3 * i + 2] = 0.f;
    }
}

template <class T>
__global__ __launch_bounds__(BLOCK) void k_sur_tan(Params<T> P, const float *__restrict__ dir, float *__restrict__ U) {
    __shared__ float Qr[WPB][SQ], Qp[WPB][SQ];
    __shared__ int Qc[WPB][SQ];
    __shared__ float Red[WPB][3][64];
    const int lane = lane_id(), wv = threadIdx.x / WAVE;
    const int i = blockIdx.x * WPB + wv;
    if (i >= P.n) return;  // whole wave; no block barrier below
    Centre<T> C;
    C.load(P, i);
    const float di0 = dir[3 * i], di1 = dir[3 * i + 1], di2 = dir[3 * i + 2];
    const int m = lane % NRBF, e = lane / NRBF;  // lane 63: e = 3, idle in the sums
    const float mu = (float)MU0 + (float)DMU * m, is2 = (float)(0.5 / (SIGMA * SIGMA)),
                isig2 = (float)(1.0 / (SIGMA * SIGMA));
    float a0 = 0.f, a1 = 0.f, a2 = 0.f;
    auto push = [&](int pos, int j, float dx, float dy, float dz, float r) {
        Qr[wv][pos] = r;
        Qp[wv][pos] = (dx * (di0 - dir[3 * j]) + dy * (di1 - dir[3 * j + 1]) + dz * (di2 - dir[3 * j + 2])) / r;
        Qc[wv][pos] = P.type_cls[P.type_s[j]];
    };
    auto flush = [&](int qn) {
        for (int q = 0; q < qn; q += 3) {
            const int k = q + e;
            if (e < 3 && k < qn) {
                const float r = Qr[wv][k], u = r - mu;
                float dfc;
                const float fc = cutf(r, dfc);
                const float v = __expf(-u * u * is2) * (dfc - fc * u * isig2) * Qp[wv][k];
                const int c = Qc[wv][k];
                a0 += c == 0 ? v : 0.f;
                a1 += c == 1 ? v : 0.f;
                a2 += c == 2 ? v : 0.f;
            }
        }
    };
    int qn = 0;
    scan_neighbours(P, C, lane, qn, push, flush);
    // CN row
    const T rci = P.rcov[C.ti];
    float cacc = 0.f;
    const int s0 = C.S->st_off[CUT_CN], s1 = s0 + C.S->st_cnt[CUT_CN];
    for (int sb = s0; sb < s1; sb += WAVE) {
        const auto Bn = C.bins(P, sb, s1);
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
                    const T dcn = finite_or_zero(-T(K1) * ex * rr * ir * ir * ir / (ope * ope));  // c'(r) / r
                    cacc += (float)dcn * ((float)dx * (di0 - dir[3 * j]) + (float)dy * (di1 - dir[3 * j + 1]) +
                                          (float)dz * (di2 - dir[3 * j + 2]));
                }
            }
        }
    }
    cacc = wave_sumf(cacc);
    Red[wv][0][lane] = a0;
    Red[wv][1][lane] = a1;
    Red[wv][2][lane] = a2;
    wave_sync();
    // per A of step: the features are functions of bohr distances
    const float ib = (float)(1.0 / sur::BOHR_A);
    float *u = U + (size_t)i * NFEAT;
    if (lane < 3 * NRBF) {
        const int c = lane / NRBF, mm = lane % NRBF;
        u[1 + lane] = ib * (Red[wv][c][mm] + Red[wv][c][NRBF + mm] + Red[wv][c][2 * NRBF + mm]);
    } else {
        u[0] = ib * cacc;
    }
}

}  // namespace

template <class T>
void launch_sur_tangent(const Params<T> &P, const float *dir, float *U, hipStream_t s) {
    if (P.n == 0) return;
    hipLaunchKernelGGL(k_sur_tan<T>, dim3((P.n + WPB - 1) / WPB), dim3(BLOCK), 0, s, P, dir, U);
}

template <class T>
void launch_sur_features(const Params<T> &P, hipStream_t s) {
    if (P.n == 0) return;
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
