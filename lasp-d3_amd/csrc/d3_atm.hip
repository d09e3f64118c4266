// Axilrod-Teller-Muto three-body term, replacing the reference's atm_kernel
// (src/d3_kernel.cu:1199-1509).  Every periodic triangle is evaluated exactly once.
//
// Enumeration.  The anchor of a triangle is its vertex with the smallest atom index (sorted
// order); among vertices of the same atom, the one whose other copies lie at lattice offsets
// t >_lex 0.  A centre i keeps only its "forward" neighbours within rc3 (atom index > i, or the
// atom i itself at an image t >_lex 0), so each triangle appears once, as a pair {j, k} of
// forward neighbours of its anchor with r_jk <= rc3.  (The reference enumerates i >= j >= k over
// every image and rescales index-degenerate triplets by 1/2 or 1/6.)
//
// Work layout.  One wave per centre (persistent blocks of 64 lanes).  Phase 1 compacts the
// forward neighbours (ballot/mbcnt) into per-block scratch; phase 2 adds C6_ij and its two
// logarithmic CN derivatives; phase 3 visits the pairs in 64 x 64 tiles: lane l owns neighbour
// k = 64 kb + l in registers, the j tile sits in LDS together with its C6 rows
// y_j[B] = C6ref_{A_j B}^T w_j and dy_j[B] = C6ref^T dw_j, and at step q lane l takes
// j = (l + q) mod 64, so no two lanes touch the same j accumulator in the same step.  Tile pairs
// whose bounding boxes are farther apart than rc3 are skipped.  Phase 4 adds the per-neighbour
// sums to the per-atom arrays with FP64 atomics.
//
// Per triangle (a = r_ij^2, b = r_ik^2, c = r_jk^2, d_j = r_i - r_j, d_k = r_i - r_k, v = r_j - r_k):
//   e = C9 A f,  C9 = -sqrt(C6_ij C6_ik C6_jk)
//   A = (abc)^(-3/2) + (3/8) Q (abc)^(-5/2),  Q = (a+b-c)(a-b+c)(-a+b+c)   [= (1 + 3 cos cos cos)/(r r r)^3]
//   f = 1 / (1 + 6 x^16),  x^6 = y = (4/3)^6 R0_ij^2 R0_ik^2 R0_jk^2 / (abc),  x^16 = y^3 y^(-1/3)
//   t_xy = (1/r_xy) de/dr_xy = 2 de/d(r_xy^2)
//        = 2 C9 [(P1 f + A fx) / (r_xy^2) + (3/8) iR5 f dQ/d(r_xy^2)],
//     P1 = -(3/2)(abc)^(-3/2) - (15/16) Q (abc)^(-5/2),  fx = 16 x^16 f^2
//   grad_i e = t_ij d_j + t_ik d_k,  grad_j e = -t_ij d_j + t_jk v,  grad_k e = -(grad_i + grad_j)
//   de/dCN_x = (C9/2)(dlnC6_xy/dCN_x + dlnC6_xz/dCN_x) A f
//   virial += -(t_ij d_j d_j^T + t_ik d_k d_k^T + t_jk v v^T)
//   each vertex's atom takes e/3 of the energy (the reference's split, src/d3_kernel.cu:1394-1399)
#include <hip/hip_runtime.h>

#include "d3_device.hpp"

namespace d3 {
namespace {

constexpr int AW = 64;                          // one wave per block
constexpr double K6 = 4096.0 / 729.0;           // (4/3)^6
enum { F_DX, F_DY, F_DZ, F_R2, F_IR2, F_C6, F_DLI, F_DLJ, F_R0SQ };  // record / LDS fields

// y^(-1/3): FP32 through the hardware exp2/log2; FP64 from the FP32 value and two Newton steps
// z <- z (4 - y z^3) / 3 (quadratic convergence, no division).
__device__ __forceinline__ float inv_cbrt(float y) {
    return __builtin_amdgcn_exp2f(-(1.0f / 3.0f) * __builtin_amdgcn_logf(y));
}
__device__ __forceinline__ double inv_cbrt(double y) {
    double z = (double)inv_cbrt((float)y);
    z = z * (4.0 - y * z * z * z) * (1.0 / 3.0);
    z = z * (4.0 - y * z * z * z) * (1.0 / 3.0);
    return z;
}
__device__ __forceinline__ float xrcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double xrcp(double x) { return 1.0 / x; }
// FP32: the hardware square root (1 ulp) without the correctly-rounded fix-up sequence.
__device__ __forceinline__ float fast_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ double fast_sqrt(double x) { return sqrt(x); }

// A wave-uniform float, moved to a scalar register.
__device__ __forceinline__ float uni(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, v)));
}

__device__ __forceinline__ float wmin(float v) {
#pragma unroll
    for (int o = AW / 2; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, AW));
    return v;
}
__device__ __forceinline__ float wmax(float v) {
#pragma unroll
    for (int o = AW / 2; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, AW));
    return v;
}

template <class T, int NR>
__global__ __launch_bounds__(AW) void k_atm(Params<T> P) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int lane = threadIdx.x;
    constexpr int nref = NR;
    const int ne = P.ne, ynr = ne * nref;
    // LDS: J tile fields [9][64] T | type [64] int | y [ynr][64] | dy [ynr][64] | acc [5][64] T |
    //      centre rows Cy [ynr], Cdy [ynr] | R0^2 table [ne*ne]
    T *Jf = reinterpret_cast<T *>(smem);
    int *Jt = reinterpret_cast<int *>(Jf + 9 * AW);
    T *Jy = reinterpret_cast<T *>(Jt + AW);
    T *Jdy = Jy + (size_t)ynr * AW;
    T *Jacc = Jdy + (size_t)ynr * AW;
    T *Cy = Jacc + 5 * AW;
    T *Cdy = Cy + ynr;
    T *R0sq = Cdy + ynr;

    const int cap = P.atm_cap;
    int *Lj = P.atm_i32 + (size_t)blockIdx.x * cap * 2;
    int *Lt = Lj + cap;
    T *LT = P.atm_T + (size_t)blockIdx.x * cap * ATM_NF;
    // per-neighbour FP64 accumulators La[5][cap]: F_p (3), H_p, Gc_p (see phase 4)
    double *La = P.atm_acc + (size_t)blockIdx.x * atm_acc_stride(cap);
    const T rc32 = P.rc_32;

    for (int idx = lane; idx < ne * ne; idx += AW) R0sq[idx] = P.r0[idx] * P.r0[idx];

    for (int i = blockIdx.x; i < P.n; i += gridDim.x) {
        Centre<T> C;
        C.load(P, i);
        const int ti = C.ti;
        __syncthreads();  // previous centre done with LDS and scratch
        for (int idx = lane; idx < ynr; idx += AW) {
            const int B = idx / nref, bb = idx % nref;
            T sy = T(0), sdy = T(0);
            for (int a = 0; a < nref; ++a) {
                const T c = P.c6[((ti * ne + B) * nref + a) * nref + bb];
                sy += c * P.w[a * P.n + i];
                sdy += c * P.dw[a * P.n + i];
            }
            Cy[idx] = sy;
            Cdy[idx] = sdy;
        }
        __syncthreads();

        // ---- phase 1: forward neighbours within rc3 (1e-12 <= r^2 <= rc3^2, src/d3_kernel.cu:1268)
        int count = 0;
        const int s0 = C.S->st_off[CUT_ATM], s1 = s0 + C.S->st_cnt[CUT_ATM];
        for (int s = s0; s < s1; ++s) {
            int j0, j1;
            T ox, oy, oz;
            C.bin(P, s, j0, j1, ox, oy, oz);
            if (j1 <= i) continue;  // bins are sorted: every atom here has index < i
            const StEntry d = P.st_d[s];
            const bool dlex = d.dx > 0 || (d.dx == 0 && (d.dy > 0 || (d.dy == 0 && d.dz > 0)));
            for (int jb = j0; jb < j1; jb += AW) {
                const int j = jb + lane;
                bool in = false;
                T dx = T(0), dy = T(0), dz = T(0), r2 = T(0);
                if (j < j1 && (j > i || (j == i && dlex))) {
                    dx = ox - P.lx[j];
                    dy = oy - P.ly[j];
                    dz = oz - P.lz[j];
                    r2 = dx * dx + dy * dy + dz * dz;
                    in = r2 <= rc32 && r2 >= T(1e-12);
                }
                const uint64_t m = ballot64(in);
                if (m == 0) continue;
                if (in) {
                    const int pos = count + mbcnt64(m);
                    if (pos < cap) {
                        Lj[pos] = j;
                        LT[F_DX * cap + pos] = dx;
                        LT[F_DY * cap + pos] = dy;
                        LT[F_DZ * cap + pos] = dz;
                        LT[F_R2 * cap + pos] = r2;
                    }
                }
                count += __popcll(m);
            }
        }
        if (count > cap) {  // uniform: flag, skip the centre, the host retries with a larger capacity
            if (lane == 0) {
                atomicOr(P.status, 2u);
                P.eat[i] = 0.0;
            }
            continue;
        }
        const int nf = count;
        __syncthreads();  // phase-1 slots were written by arbitrary lanes

        // ---- phase 2: C6_ij, dlnC6_ij/dCN_i, dlnC6_ij/dCN_j, 1/r^2, R0_ij^2; zero accumulators.
        // Slot p is always touched by lane p % 64, here and below.
        for (int p = lane; p < nf; p += AW) {
            const int j = Lj[p], tj = P.type_s[j];
            T c6 = T(0), dci = T(0), dcj = T(0);
            for (int b = 0; b < nref; ++b) {
                const T wj = P.w[b * P.n + j], dwj = P.dw[b * P.n + j];
                c6 += Cy[tj * nref + b] * wj;
                dci += Cdy[tj * nref + b] * wj;
                dcj += Cy[tj * nref + b] * dwj;
            }
            const T ic6 = c6 > T(0) ? T(1) / c6 : T(0);
            Lt[p] = tj;
            LT[F_IR2 * cap + p] = T(1) / LT[F_R2 * cap + p];
            LT[F_C6 * cap + p] = c6;
            LT[F_DLI * cap + p] = dci * ic6;
            LT[F_DLJ * cap + p] = dcj * ic6;
            LT[F_R0SQ * cap + p] = R0sq[ti * ne + tj];
            for (int q = 0; q < 5; ++q) La[q * cap + p] = 0.0;
        }
        __syncthreads();

        // ---- phase 3: pairs {j, k}, j < k, of forward neighbours.  Per-pair terms are summed in T
        // over one 64 x 64 tile (at most 64 terms per sum); tile sums go to the FP64 per-neighbour
        // accumulators La, which lane l alone touches for its slots.
        float cnt = 0.0f;
        const int ntile = (nf + AW - 1) / AW;
        for (int kb = 0; kb < ntile; ++kb) {
            const int k = kb * AW + lane;
            const bool kval = k < nf;
            T kx = T(0), ky = T(0), kz = T(0), kb2 = T(1), kib = T(1), kc6 = T(0), kr0 = T(0);
            T kw[NR], kdw[NR];
            int tk = 0;
            for (int a = 0; a < NR; ++a) kw[a] = kdw[a] = T(0);
            if (kval) {
                const int jk = Lj[k];
                kx = LT[F_DX * cap + k];
                ky = LT[F_DY * cap + k];
                kz = LT[F_DZ * cap + k];
                kb2 = LT[F_R2 * cap + k];
                kib = LT[F_IR2 * cap + k];
                kc6 = LT[F_C6 * cap + k];
                kr0 = LT[F_R0SQ * cap + k];
                tk = Lt[k];
                for (int a = 0; a < nref; ++a) {
                    kw[a] = P.w[a * P.n + jk];
                    kdw[a] = P.dw[a * P.n + jk];
                }
            }
            // bounding box of the k wave (valid lanes), wave-uniform
            const float kminx = uni(wmin(kval ? (float)kx : 3e38f)), kmaxx = uni(wmax(kval ? (float)kx : -3e38f));
            const float kminy = uni(wmin(kval ? (float)ky : 3e38f)), kmaxy = uni(wmax(kval ? (float)ky : -3e38f));
            const float kminz = uni(wmin(kval ? (float)kz : 3e38f)), kmaxz = uni(wmax(kval ? (float)kz : -3e38f));

            for (int jt = 0; jt <= kb; ++jt) {
                const int pj = jt * AW + lane;
                const bool jval = pj < nf;
                // bounding box of the j tile from the records (global scratch, same-lane slots)
                T jxs = T(0), jys = T(0), jzs = T(0);
                if (jval) {
                    jxs = LT[F_DX * cap + pj];
                    jys = LT[F_DY * cap + pj];
                    jzs = LT[F_DZ * cap + pj];
                }
                const float jminx = uni(wmin(jval ? (float)jxs : 3e38f)), jmaxx = uni(wmax(jval ? (float)jxs : -3e38f));
                const float jminy = uni(wmin(jval ? (float)jys : 3e38f)), jmaxy = uni(wmax(jval ? (float)jys : -3e38f));
                const float jminz = uni(wmin(jval ? (float)jzs : 3e38f)), jmaxz = uni(wmax(jval ? (float)jzs : -3e38f));
                const float gx = fmaxf(0.f, fmaxf(jminx - kmaxx, kminx - jmaxx));
                const float gy = fmaxf(0.f, fmaxf(jminy - kmaxy, kminy - jmaxy));
                const float gz = fmaxf(0.f, fmaxf(jminz - kmaxz, kminz - jmaxz));
                // skip only when clearly apart (1e-3 relative margin covers float rounding)
                if (gx * gx + gy * gy + gz * gz > (float)rc32 * 1.001f) continue;

                // stage the j tile
                wave_sync();
                if (jval) {
                    Jf[F_DX * AW + lane] = jxs;
                    Jf[F_DY * AW + lane] = jys;
                    Jf[F_DZ * AW + lane] = jzs;
                    Jf[F_R2 * AW + lane] = LT[F_R2 * cap + pj];
                    Jf[F_IR2 * AW + lane] = LT[F_IR2 * cap + pj];
                    Jf[F_C6 * AW + lane] = LT[F_C6 * cap + pj];
                    Jf[F_R0SQ * AW + lane] = LT[F_R0SQ * cap + pj];
                    const int jj = Lj[pj], tj = Lt[pj];
                    Jt[lane] = tj;
                    T wj[NR], dwj[NR];
                    for (int a = 0; a < nref; ++a) {
                        wj[a] = P.w[a * P.n + jj];
                        dwj[a] = P.dw[a * P.n + jj];
                    }
                    for (int B = 0; B < ne; ++B)
                        for (int bb = 0; bb < nref; ++bb) {
                            T sy = T(0), sdy = T(0);
                            for (int a = 0; a < nref; ++a) {
                                const T c = P.c6[((tj * ne + B) * nref + a) * nref + bb];
                                sy += c * wj[a];
                                sdy += c * dwj[a];
                            }
                            Jy[(B * nref + bb) * AW + lane] = sy;
                            Jdy[(B * nref + bb) * AW + lane] = sdy;
                        }
                }
                for (int q = 0; q < 5; ++q) Jacc[q * AW + lane] = T(0);
                wave_sync();

                // tile-local sums of neighbour k (folded into FP64 per tile)
                T kF0 = T(0), kF1 = T(0), kF2 = T(0), kH = T(0), kG = T(0);
                const int jlim = (jt == kb) ? lane : AW;  // diagonal tile: j < k
#pragma unroll 1
                for (int q = 0; q < AW; ++q) {
                    // lane l updates the j accumulator that lane l+1 updated one step earlier
                    wave_sync();
                    const int jl = (lane + q) & (AW - 1);
                    const int jg = jt * AW + jl;
                    if (!(kval && jl < jlim && jg < nf)) continue;
                    const T jx = Jf[F_DX * AW + jl], jy = Jf[F_DY * AW + jl], jz = Jf[F_DZ * AW + jl];
                    const T vx = kx - jx, vy = ky - jy, vz = kz - jz;  // v = r_j - r_k
                    const T c = vx * vx + vy * vy + vz * vz;
                    if (c > rc32 || c < T(1e-12)) continue;
                    const T a = Jf[F_R2 * AW + jl], ia = Jf[F_IR2 * AW + jl];
                    const T b = kb2, ib = kib;
                    const int tj = Jt[jl];
                    // C6_jk and its CN derivatives from the j row for type(k)
                    T c6jk = T(0), dk6 = T(0), dj6 = T(0);
                    const int yb = (tk * nref) * AW + jl;
                    for (int bb = 0; bb < nref; ++bb) {
                        const T yv = Jy[yb + bb * AW], dyv = Jdy[yb + bb * AW];
                        c6jk += yv * kw[bb];
                        dk6 += yv * kdw[bb];
                        dj6 += dyv * kw[bb];
                    }
                    const T ic6jk = c6jk > T(0) ? xrcp(c6jk) : T(0);
                    const T C9 = -fast_sqrt(Jf[F_C6 * AW + jl] * kc6 * c6jk);
                    const T abc = a * b * c;
                    const T iR = xrsqrt(abc);
                    const T iR2 = iR * iR, iR3 = iR2 * iR, iR5 = iR3 * iR2;
                    const T ic = xrcp(c);
                    const T u1 = a + b - c, u2 = a - b + c, u3 = -a + b + c;
                    const T Q = u1 * u2 * u3;
                    const T dQa = u2 * u3 + u1 * (u3 - u2);
                    const T dQb = u2 * u3 + u1 * (u2 - u3);
                    const T dQc = u1 * (u2 + u3) - u2 * u3;
                    const T A = iR3 + T(0.375) * Q * iR5;
                    const T yv6 = T(K6) * Jf[F_R0SQ * AW + jl] * kr0 * R0sq[tj * ne + tk] * ia * ib * ic;
                    const T x16 = yv6 * yv6 * yv6 * inv_cbrt(yv6);
                    const T f = xrcp(T(1) + T(6) * x16);
                    const T fx = T(16) * x16 * f * f;
                    const T P1 = T(-1.5) * iR3 - T(0.9375) * Q * iR5;
                    const T G2 = T(2) * C9 * (P1 * f + A * fx);
                    const T H2 = T(0.75) * C9 * iR5 * f;
                    const T tij = G2 * ia + H2 * dQa;
                    const T tik = G2 * ib + H2 * dQb;
                    const T tjk = G2 * ic + H2 * dQc;
                    // hC = e/2; the centre's share, its force, dE/dCN and the virial follow from
                    // the per-neighbour sums after the tiles (phase 4)
                    const T hC = T(0.5) * C9 * A * f;
                    // neighbour j (LDS slot jl: distinct across lanes at each step)
                    const T fjx = -tij * jx + tjk * vx, fjy = -tij * jy + tjk * vy, fjz = -tij * jz + tjk * vz;
                    Jacc[0 * AW + jl] += fjx;
                    Jacc[1 * AW + jl] += fjy;
                    Jacc[2 * AW + jl] += fjz;
                    Jacc[3 * AW + jl] += hC;
                    Jacc[4 * AW + jl] += hC * dj6 * ic6jk;
                    // neighbour k (registers): grad_k e = -t_ik d_k - t_jk v
                    kF0 -= tik * kx + tjk * vx;
                    kF1 -= tik * ky + tjk * vy;
                    kF2 -= tik * kz + tjk * vz;
                    kH += hC;
                    kG += hC * dk6 * ic6jk;
                    cnt += 1.0f;
                }
                if (kval) {
                    La[0 * cap + k] += (double)kF0;
                    La[1 * cap + k] += (double)kF1;
                    La[2 * cap + k] += (double)kF2;
                    La[3 * cap + k] += (double)kH;
                    La[4 * cap + k] += (double)kG;
                }
                wave_sync();
                if (jval)
                    for (int q = 0; q < 5; ++q) La[q * cap + pj] += (double)Jacc[q * AW + lane];
            }
        }

        // ---- phase 4.  Per neighbour slot p: F_p = sum of grad_p e, H_p = sum of e/2, Gc_p = sum of
        // (e/2) dlnC6_jk/dCN_p over the triangles it belongs to.  Then
        //   dE/dCN_p = Gc_p + (dlnC6_ip/dCN_p) H_p,  E/3 share of p = (2/3) H_p,
        //   centre: E = sum H_p (each triangle feeds two slots), dE/dCN_i = sum (dlnC6_ip/dCN_i) H_p,
        //   F_i = -sum F_p (translation invariance), virial = sum d_p (x) F_p
        //   (= -sum over edges t r r^T, since sum_v r_v (x) grad_v e = sum_edges t r r^T).
        Acc<double> aE, aG, aF[3], aW[9];
        aE.init();
        aG.init();
        for (int q = 0; q < 3; ++q) aF[q].init();
        for (int q = 0; q < 9; ++q) aW[q].init();
        for (int p = lane; p < nf; p += AW) {
            const int j = Lj[p];
            const double fx = La[0 * cap + p], fy = La[1 * cap + p], fz = La[2 * cap + p];
            const double H = La[3 * cap + p], Gc = La[4 * cap + p];
            const double dx = (double)LT[F_DX * cap + p], dy = (double)LT[F_DY * cap + p],
                         dz = (double)LT[F_DZ * cap + p];
            atomicAdd(&P.fa[3 * j + 0], fx);
            atomicAdd(&P.fa[3 * j + 1], fy);
            atomicAdd(&P.fa[3 * j + 2], fz);
            atomicAdd(&P.ga[j], Gc + (double)LT[F_DLJ * cap + p] * H);
            atomicAdd(&P.eas[j], H * (2.0 / 3.0));
            aE.add(H);
            aG.add((double)LT[F_DLI * cap + p] * H);
            aF[0].add(-fx);
            aF[1].add(-fy);
            aF[2].add(-fz);
            aW[0].add(dx * fx);
            aW[1].add(dy * fy);
            aW[2].add(dz * fz);
            aW[3].add(dx * fy);
            aW[4].add(dy * fx);
            aW[5].add(dx * fz);
            aW[6].add(dz * fx);
            aW[7].add(dy * fz);
            aW[8].add(dz * fy);
        }
        const double E = wave_total(aE), G = wave_total(aG);
        const double F[3] = {wave_total(aF[0]), wave_total(aF[1]), wave_total(aF[2])};
        double wv[9];
        for (int q = 0; q < 9; ++q) wv[q] = wave_total(aW[q]);
        const double W[6] = {wv[0], wv[1], wv[2], 0.5 * (wv[3] + wv[4]), 0.5 * (wv[5] + wv[6]),
                             0.5 * (wv[7] + wv[8])};
        cnt = wave_sumf(cnt);
        if (lane == 0) {
            atomicAdd(&P.fa[3 * i + 0], F[0]);
            atomicAdd(&P.fa[3 * i + 1], F[1]);
            atomicAdd(&P.fa[3 * i + 2], F[2]);
            atomicAdd(&P.ga[i], G);
            atomicAdd(&P.eas[i], E * (1.0 / 3.0));
            P.eat[i] = E;
            const double iv = C.S->inv_vol;
            for (int q = 0; q < 6; ++q) P.vir[6 * i + q] += W[q] * iv;
            P.npair[3 * i + 2] = cnt;
        }
    }
}

__global__ void k_fill_d(double *p, size_t n) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = 0.0;
}

// P.g += ga, P.f += fa (after the ATM kernel, before the chain rule).
template <class T>
__global__ void k_merge(Params<T> P) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P.n) return;
    P.g[p] += P.ga[p];
    for (int q = 0; q < 3; ++q) P.f[3 * p + q] += P.fa[3 * p + q];
}

}  // namespace

size_t atm_lds_bytes(int ne, int nref, size_t ts) {
    const size_t ynr = (size_t)ne * nref;
    return 9 * AW * ts + AW * sizeof(int) + 2 * ynr * AW * ts + 5 * AW * ts + 2 * ynr * ts + (size_t)ne * ne * ts;
}

template <class T>
void launch_atm(const Params<T> &P, int nblocks, hipStream_t s) {
    // fa (3N), ga, eas, eat (N each) are contiguous: zero them in one pass
    const size_t nz = (size_t)6 * P.n;
    if (nz) hipLaunchKernelGGL(k_fill_d, dim3((unsigned)((nz + 255) / 256)), dim3(256), 0, s, P.fa, nz);
    if (P.n == 0 || !(P.rc_32 > T(0))) return;
    const size_t lds = atm_lds_bytes(P.ne, P.nref, sizeof(T));
    if (P.nref == 5)
        hipLaunchKernelGGL((k_atm<T, 5>), dim3(nblocks), dim3(AW), lds, s, P);
    else
        hipLaunchKernelGGL((k_atm<T, 7>), dim3(nblocks), dim3(AW), lds, s, P);
    hipLaunchKernelGGL(k_merge<T>, dim3((P.n + 255) / 256), dim3(256), 0, s, P);
}

template void launch_atm<double>(const Params<double> &, int, hipStream_t);
template void launch_atm<float>(const Params<float> &, int, hipStream_t);

}  // namespace d3
