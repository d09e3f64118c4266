// HIP kernels of the D3 pipeline for gfx950 (CDNA4, wave64).
//
// Pipeline per evaluation (all on one stream, no host synchronisation in between):
//   bin sort   k_zero -> k_bin -> k_scan -> k_scatter -> k_canon -> k_gather
//   k_cn       coordination numbers + normalised CN weights per atom
//   k_2b       two-body energy, dE/dCN, force, virial
//   k_atm      Axilrod-Teller-Muto three-body term (d3_atm.hip)
//   k_chain    dE/dCN chain-rule force and virial
//   reduce / finalize (d3_reduce.hip)
// The formulas restate the reference's src/d3_kernel.cu; each kernel cites its counterpart.
// T = double is the plain FP64 path, T = float the FP32 path with compensated accumulation.
#include <hip/hip_runtime.h>

#include <utility>

#include <cfloat>

#include "d3_device.hpp"

namespace d3 {
namespace {

// (device helpers: d3_device.hpp)

// ---------------------------------------------------------------------------------------------
// Bin sort: counting sort of atoms by bin, then a canonical (ascending caller index) order
// inside every bin so that all later sums run in a fixed order.  Replaces the host counting
// sort of src/d3_buffer.cu:496-634 (wrap to [0,1) with the edge fix of :556-560).
// ---------------------------------------------------------------------------------------------
__global__ void k_fill_i32(int *p, int n, int v) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = v;
}

// Wrapped coordinate c of position r with lattice shifts fl, and the position relative to the
// centre of bin (b0, b1, b2): the expressions of the bin sort, shared with the regather of moved
// positions so that an unmoved atom lands on the same bits.
__device__ __forceinline__ double wrapped_coord(const double r[3], const double fl[3], const SysDev &S, int c) {
    return r[c] - fl[0] * S.H[0 * 3 + c] - fl[1] * S.H[1 * 3 + c] - fl[2] * S.H[2 * 3 + c];
}
__device__ __forceinline__ void local_coords(const double w[3], int key, const SysDev &S, double l[3]) {
    const int b = key - S.bin_base;
    const int b0 = b % S.nb[0], b1 = (b / S.nb[0]) % S.nb[1], b2 = b / (S.nb[0] * S.nb[1]);
    const double u0 = (b0 + 0.5) / S.nb[0], u1 = (b1 + 0.5) / S.nb[1], u2 = (b2 + 0.5) / S.nb[2];
#pragma unroll
    for (int c = 0; c < 3; ++c) l[c] = w[c] - (u0 * S.H[0 * 3 + c] + u1 * S.H[1 * 3 + c] + u2 * S.H[2 * 3 + c]);
}

template <class T>
__global__ void k_bin(Params<T> P) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.n) return;
    const SysDev &S = P.sys[P.sys_in[i]];
    const double r[3] = {P.pos_in[3 * i], P.pos_in[3 * i + 1], P.pos_in[3 * i + 2]};
    int b[3];
    double fl[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double s = r[0] * S.Hinv[0 * 3 + k] + r[1] * S.Hinv[1 * 3 + k] + r[2] * S.Hinv[2 * 3 + k];
        fl[k] = floor(s);
        const double u = s - fl[k];
        int bk = (int)(u * S.nb[k]);
        if (bk >= S.nb[k]) {  // u rounded to 1.0: the atom belongs to bin 0 of the next image
            bk = 0;
            fl[k] += 1.0;
        }
        if (bk < 0) bk = 0;
        b[k] = bk;
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) P.wrapped[3 * i + c] = wrapped_coord(r, fl, S, c);
    if (P.shift)
        for (int k = 0; k < 3; ++k) P.shift[3 * i + k] = fl[k];
    const int key = S.bin_base + b[0] + S.nb[0] * (b[1] + S.nb[1] * b[2]);
    P.key_in[i] = key;
    atomicAdd(&P.bin_count[key], 1);
}

// Exclusive scan of bin_count into bin_start[0..nbins] by one block.
__global__ void k_scan(const int *count, int *start, int *cursor, int nbins) {
    __shared__ int part[1024];
    __shared__ int carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < nbins; base += 1024) {
        const int t = base + threadIdx.x;
        const int v = t < nbins ? count[t] : 0;
        part[threadIdx.x] = v;
        __syncthreads();
        for (int o = 1; o < 1024; o <<= 1) {  // Hillis-Steele inclusive scan
            int add = threadIdx.x >= o ? part[threadIdx.x - o] : 0;
            __syncthreads();
            part[threadIdx.x] += add;
            __syncthreads();
        }
        if (t < nbins) {
            const int ex = carry + part[threadIdx.x] - v;
            start[t] = ex;
            cursor[t] = ex;
        }
        __syncthreads();
        if (threadIdx.x == 1023) carry += part[1023];
        __syncthreads();
    }
    if (threadIdx.x == 0) start[nbins] = carry;
}

template <class T>
__global__ void k_scatter(Params<T> P) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.n) return;
    const int slot = atomicAdd(&P.cursor[P.key_in[i]], 1);
    P.perm[slot] = i;
}

// One block per bin: rank of every entry among the bin's entries (ascending caller index).
__global__ void k_canon(const int *bin_start, const int *perm, int *perm2) {
    const int b = blockIdx.x;
    const int s = bin_start[b], e = bin_start[b + 1];
    for (int t = s + threadIdx.x; t < e; t += blockDim.x) {
        const int v = perm[t];
        int rank = 0;
        for (int u = s; u < e; ++u) rank += perm[u] < v;
        perm2[s + rank] = v;
    }
}

template <class T>
__global__ void k_gather(Params<T> P) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P.n) return;
    const int i = P.perm2[p];
    const int key = P.key_in[i];
    const int sid = P.sys_in[i];
    const SysDev &S = P.sys[sid];
    const double w[3] = {P.wrapped[3 * i], P.wrapped[3 * i + 1], P.wrapped[3 * i + 2]};
    double l[3];
    local_coords(w, key, S, l);
    P.lx[p] = (T)l[0];
    P.ly[p] = (T)l[1];
    P.lz[p] = (T)l[2];
    P.type_s[p] = P.type_in[i];
    P.orig_s[p] = i;
    P.bin_s[p] = key;
    P.sys_s[p] = sid;
}

// Moved positions of the same atoms into the kept order: the wrap shifts and bins of the last
// sort, the new caller positions (atoms within half the skin of where they were binned).
template <class T>
__global__ void k_regather(Params<T> P) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P.n) return;
    const int i = P.orig_s[p];
    const SysDev &S = P.sys[P.sys_s[p]];
    const double r[3] = {P.pos_in[3 * i], P.pos_in[3 * i + 1], P.pos_in[3 * i + 2]};
    const double fl[3] = {P.shift[3 * i], P.shift[3 * i + 1], P.shift[3 * i + 2]};
    double w[3], l[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) w[c] = wrapped_coord(r, fl, S, c);
    local_coords(w, P.bin_s[p], S, l);
    P.lx[p] = (T)l[0];
    P.ly[p] = (T)l[1];
    P.lz[p] = (T)l[2];
}

// ---------------------------------------------------------------------------------------------
// Pair walks.  A wave walks its centre's stencil entries [s0, s1); lanes test 64 atoms of a bin
// at a time against the cutoff (0 < r^2 <= rc^2, or < rc^2 when strict), the accepted pairs go to
// the wave's ring queue in LDS (ballot + mbcnt), and every 64 queued pairs are evaluated by the
// whole wave, one pair per lane (the remainder at the end by the lanes that have one).  The
// loads of a candidate are clamped to the bin (no lane-dependent branch) and addressed by 32-bit
// offsets; the kernels that choose it (AHEAD) issue them one chunk ahead.
// ---------------------------------------------------------------------------------------------
// NPAY: values per candidate loaded with its coordinates and queued with it (the evaluation reads
// them from LDS instead of gathering them after the pair was accepted).
template <class T, int NPAY = 0>
struct Queue {
    int j[QCAP];
    T dx[QCAP], dy[QCAP], dz[QCAP];
    T pay[NPAY][QCAP];
};
template <class T>
struct Queue<T, 0> {
    int j[QCAP];
    T dx[QCAP], dy[QCAP], dz[QCAP];
};

// AHEAD: the coordinates of the next chunk of candidates are loaded before the current chunk is
// tested (more registers; kernels choose it per precision by measurement).
// pay(jj, v): the NPAY payload values of candidate jj; eval(j, dx, dy, dz, v) evaluates a pair.
template <class T, bool STRICT, bool AHEAD, int NPAY, class Pay, class Eval>
__device__ __forceinline__ void walk_pairs(const Params<T> &P, const Centre<T> &C, int s0, int s1, T rc, T rc2,
                                           Queue<T, NPAY> &Q, Pay &&pay, Eval &&eval) {
    const int lane = lane_id();
    int qt = 0, qh = 0;  // pairs queued, pairs evaluated (ring positions modulo QCAP)
    // One chunk: up to 64 candidates of one bin (this lane's atom index j, the bin's end j1, the
    // centre's position plus the image vector, the candidate's coordinates and payload).
    struct Chunk {
        unsigned j, j1;
        T ox, oy, oz, x, y, z;
        T v[NPAY > 0 ? NPAY : 1];
    };
    auto pop = [&](int pos) __attribute__((always_inline)) {
        T v[NPAY > 0 ? NPAY : 1];
        if constexpr (NPAY > 0)
#pragma unroll
            for (int k = 0; k < NPAY; ++k) v[k] = Q.pay[k][pos];
        eval(Q.j[pos], Q.dx[pos], Q.dy[pos], Q.dz[pos], v);
    };
    auto test = [&](const Chunk &K) __attribute__((always_inline)) {
        const T dx = K.ox - K.x, dy = K.oy - K.y, dz = K.oz - K.z;
        const T r2 = dx * dx + dy * dy + dz * dz;
        // one ballot per compare: each folds into its v_cmp (a ballot of the combined condition
        // costs a select and a compare more per candidate step)
        const bool c0 = K.j < K.j1, c1 = STRICT ? r2 < rc2 : r2 <= rc2, c2 = r2 > T(0);
        const bool in = c0 && c1 && c2;
        const uint64_t m = ballot64(c0) & ballot64(c1) & ballot64(c2);
        if (m == 0) return;
        if (in) {
            const int pos = (qt + mbcnt64(m)) & (QCAP - 1);
            Q.j[pos] = (int)K.j;
            Q.dx[pos] = dx;
            Q.dy[pos] = dy;
            Q.dz[pos] = dz;
            if constexpr (NPAY > 0)
#pragma unroll
                for (int k = 0; k < NPAY; ++k) Q.pay[k][pos] = K.v[k];
        }
        qt += __popcll(m);
        if (qt - qh >= WAVE) {
            wave_sync();
            pop((qh + lane) & (QCAP - 1));
            qh += WAVE;
            wave_sync();  // the next pairs may go to the slots just read
        }
    };
    for (int sb = s0; sb < s1; sb += WAVE) {
        const auto Bn = C.bins(P, sb, s1, rc);
        if constexpr (!AHEAD) {
            const int nsb = s1 - sb < WAVE ? s1 - sb : WAVE;
            for (int sk = 0; sk < nsb; ++sk) {
                int j0, j1;
                Chunk K;
                Bn.get(sk, j0, j1, K.ox, K.oy, K.oz);
                K.j1 = (unsigned)j1;
                for (int jb = j0; jb < j1; jb += WAVE) {
                    K.j = (unsigned)(jb + lane);
                    const unsigned jj = min(K.j, K.j1 - 1);  // loads clamped to the bin
                    K.x = at32(P.lx, jj);
                    K.y = at32(P.ly, jj);
                    K.z = at32(P.lz, jj);
                    if constexpr (NPAY > 0) pay(jj, K.v);
                    test(K);
                }
            }
            continue;
        }
        // the group's chunks in stencil order: lane l's entry has nch of them, incl chunks end
        // with it (a scan over the lanes), so chunk c belongs to the first lane with incl > c
        const int nch = (Bn.j1 - Bn.j0 + WAVE - 1) / WAVE;
        int incl = nch;
#pragma unroll
        for (int o = 1; o < WAVE; o <<= 1) {
            const int v = __shfl_up(incl, o, WAVE);
            if (lane >= o) incl += v;
        }
        const int excl = incl - nch, total = __builtin_amdgcn_readlane(incl, WAVE - 1);
        auto fetch = [&](int c, Chunk &K) __attribute__((always_inline)) {
            const int k = (int)__builtin_ctzll(ballot64(incl > c));
            int j0, j1;
            Bn.get(k, j0, j1, K.ox, K.oy, K.oz);
            K.j = (unsigned)(j0 + (c - __builtin_amdgcn_readlane(excl, k)) * WAVE + lane);
            K.j1 = (unsigned)j1;
            const unsigned jj = min(K.j, K.j1 - 1);  // loads clamped to the bin: no lane branch
            K.x = at32(P.lx, jj);
            K.y = at32(P.ly, jj);
            K.z = at32(P.lz, jj);
            if constexpr (NPAY > 0) pay(jj, K.v);
        };
        // two chunks in flight: the next chunk's coordinates are loaded before the current one
        // is tested, so their latency overlaps the test and any evaluation it triggers (the loads
        // are issued on every path -- past the end they repeat the last chunk -- so that the
        // wait before a test covers only the older chunk)
        if (total == 0) continue;
        Chunk A, B;
        fetch(0, A);
        for (int c = 0; c < total; c += 2) {
            fetch(min(c + 1, total - 1), B);
            test(A);
            if (c + 1 >= total) break;
            fetch(min(c + 2, total - 1), A);
            test(B);
        }
    }
    wave_sync();
    if (lane < qt - qh) pop((qh + lane) & (QCAP - 1));
}

// no payload
template <class T>
struct NoPay {
    __device__ void operator()(unsigned, T *) const {}
};

// Reciprocal and reciprocal square root of the pair math: the FP32 hardware instructions
// (1 ulp), correctly rounded FP64.
__device__ __forceinline__ float prcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double prcp(double x) { return 1.0 / x; }
__device__ __forceinline__ float prsqrt(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ double prsqrt(double x) { return rsqrt(x); }

// Per-lane sums of N terms: plain partial sums over at most FOLD pairs, folded into the
// compensated totals (two-sum in FP32) -- one compensated add per FOLD pairs instead of per pair.
template <class T, int N>
struct AccN {
    static constexpr int FOLD = 16;
    Acc<T> t[N];
    T p[N];
    int np;
    __device__ void init() {
        for (int k = 0; k < N; ++k) t[k].init();
        clear();
    }
    __device__ void clear() {
        for (int k = 0; k < N; ++k) p[k] = T(0);
        np = 0;
    }
    __device__ void fold() {
        for (int k = 0; k < N; ++k) t[k].add(p[k]);
        clear();
    }
    __device__ void add(const T v[N]) {
        if constexpr (sizeof(T) == sizeof(float)) {
            for (int k = 0; k < N; ++k) p[k] += v[k];
            if (++np == FOLD) fold();
        } else {
            for (int k = 0; k < N; ++k) t[k].add(v[k]);
        }
    }
};

// ---------------------------------------------------------------------------------------------
// Coordination number, src/d3_kernel.cu:544-574 (calculate_CN) and 582-683: one wave per
// centre atom over the pairs of its CN stencil (walk_pairs).  CN_i = sum over
// 0 < r^2 <= rc_CN^2 of 1 / (1 + exp(-k1 ((Rcov_i + Rcov_j) / r - 1))).  The same wave then
// forms the centre's normalised Gaussian weights of the C6 interpolation:
//   w_a = exp(-k3 (CN - CNref_a)^2) / sum_b exp(-k3 (CN - CNref_b)^2), a valid iff CNref_a > -1
//   dw_a/dCN = w_a (l_a - sum_b w_b l_b), l_a = -2 k3 (CN - CNref_a)
// With CN references that depend only on (element, reference) -- checked at handle init -- the
// reference's 25-term Z/W of src/d3_kernel.cu:709-800 equals w_A^T C6ref w_B exactly.
// ---------------------------------------------------------------------------------------------
// ONE: a handle of one element type, Rcov_i + Rcov_j = 2 Rcov_i (no per-pair gathers).  FP32 at
// 6 waves per SIMD (78 VGPRs with the look-ahead walk, no spills).
template <class T, bool ONE>
__global__ __launch_bounds__(BLOCK, sizeof(T) == sizeof(float) ? 6 : 1) void k_cn(Params<T> P) {
    __shared__ Queue<T> Qs[WPB];
    const int lane = lane_id(), wv = threadIdx.x >> 6;
    const int i = blockIdx.x * WPB + wv;
    if (i >= P.n) return;
    Centre<T> C;
    C.load(P, i);
    const T rci = P.rcov[C.ti];
    AccN<T, 1> acc;
    acc.init();
    float cnt = 0.0f;
    const int s0 = C.S->st_off[CUT_CN], s1 = s0 + C.S->st_cnt[CUT_CN];
    walk_pairs<T, false, true, 0>(P, C, s0, s1, P.rc_cn, P.rc_cn2, Qs[wv], NoPay<T>{},
                                  [&](int j, T dx, T dy, T dz, const T *) {
        const T r2 = dx * dx + dy * dy + dz * dz;
        const T ir = prsqrt(r2);
        const T rr = ONE ? rci + rci : rci + P.rcov[at32(P.type_s, (unsigned)j)];
        const T ex = xexp(-T(K1) * (rr * ir - T(1)));
        T c[1] = {prcp(T(1) + ex)};
        if (P.cn_smooth) {
            T unused = T(0);
            smooth_cn_term(P.rc_cn, r2 * ir, ir, c[0], unused);
        }
        acc.add(c);
        cnt += 1.0f;
    });
    if constexpr (sizeof(T) == sizeof(float)) acc.fold();
    const double cnd = wave_total(acc.t[0]);
    cnt = wave_sumf(cnt);
    if (lane == 0) {
        const T cn = (T)cnd;
        P.cn[i] = cn;
        P.npair[3 * i + 1] = cnt;
        P.npair[3 * i + 2] = 0.0f;  // three-body count, set by the ATM kernel when it runs
        T ex[MAX_NREF], la[MAX_NREF];
        T mx = -T(FLT_MAX);
        for (int a = 0; a < P.nref; ++a) {
            const T ref = P.cnref[C.ti * P.nref + a];
            const T dcn = cn - ref;
            ex[a] = (ref > T(-1)) ? -T(K3) * dcn * dcn : -T(FLT_MAX);
            la[a] = -T(2.0 * K3) * dcn;
            mx = ex[a] > mx ? ex[a] : mx;
        }
        T sum = T(0);
        for (int a = 0; a < P.nref; ++a) {
            ex[a] = (ex[a] > -T(FLT_MAX)) ? xexp(ex[a] - mx) : T(0);
            sum += ex[a];
        }
        T mean = T(0);
        for (int a = 0; a < P.nref; ++a) {
            ex[a] /= sum;
            mean += ex[a] * la[a];
        }
        for (int a = 0; a < P.nref; ++a) {
            P.w[a * P.n + i] = ex[a];
            P.dw[a * P.n + i] = ex[a] * (la[a] - mean);
        }
    }
}

// ---------------------------------------------------------------------------------------------
// Two-body term, src/d3_kernel.cu:902-986 (calculate_twobody) and 997-1197 (two_body_kernel).
// One wave per centre; lanes test 64 candidates at a time, in-range pairs (0 < r^2 < rc^2,
// strict) are compacted through an LDS queue with ballot/mbcnt, and every 64 queued pairs are
// evaluated by the full wave.  Per pair (A = type i, B = type j):
//   C6 = w_i^T C6ref_AB w_j = y_i[B] . w_j with y_i[B] = C6ref_AB^T w_i (per-centre, in LDS)
//   dC6/dCN_i = dy_i[B] . w_j,  C8 = 3 C6 r2r4_A r2r4_B
//   E+_i += (s6 C6 f6 / r^6 + s8 C8 f8 / r^8) / 2,  dE/dCN_i += s6 f6 dC6 / r^6 + s8 f8 dC8 / r^8
//   phi = (1/r) d/dr (s6 C6 f6 r^-6 + s8 C8 f8 r^-8),  F_i += phi d,  vir += -phi d d^T / 2
// Damping (src/d3_kernel.cu:359-424): zero f_n = 1/(1 + 6 (s_rn R0/r)^a_n), a = 14, 16, R0 = r0ab;
// BJ f_n = r^n / (r^n + (a1 sqrt(3 r2r4_A r2r4_B) + a2)^n).
// ---------------------------------------------------------------------------------------------

// Per-lane sums of the two-body terms: plain partial sums over at most FOLD pairs, folded into
// the compensated totals (two-sum in FP32) — one compensated add per FOLD pairs instead of per
// pair.
template <class T>
struct Acc2b {
    static constexpr int FOLD = 16;
    Acc<T> e, g, f[3], v[6];
    T pe, pg, pf[3], pv[6];
    int np;
    __device__ void init() {
        e.init();
        g.init();
        for (int k = 0; k < 3; ++k) f[k].init();
        for (int k = 0; k < 6; ++k) v[k].init();
        clear();
    }
    __device__ void clear() {
        pe = pg = T(0);
        for (int k = 0; k < 3; ++k) pf[k] = T(0);
        for (int k = 0; k < 6; ++k) pv[k] = T(0);
        np = 0;
    }
    __device__ void fold() {
        e.add(pe);
        g.add(pg);
        for (int k = 0; k < 3; ++k) f[k].add(pf[k]);
        for (int k = 0; k < 6; ++k) v[k].add(pv[k]);
        clear();
    }
    __device__ void step() {
        if (++np == FOLD) fold();
    }
    // one pair's terms: FP32 into the partial sums, FP64 straight into the totals
    __device__ void add(T de, T dg, T f0, T f1, T f2, const T vv[6]) {
        if constexpr (sizeof(T) == sizeof(float)) {
            pe += de;
            pg += dg;
            pf[0] += f0;
            pf[1] += f1;
            pf[2] += f2;
            for (int k = 0; k < 6; ++k) pv[k] += vv[k];
            step();
        } else {
            e.add(de);
            g.add(dg);
            f[0].add(f0);
            f[1].add(f1);
            f[2].add(f2);
            for (int k = 0; k < 6; ++k) v[k].add(vv[k]);
        }
    }
};

// NR = the handle's reference count: the weight gathers of a pair are unrolled and issued
// together.  NR = 0: tables that do not factorise per element; C6 and dC6/dCN_i are interpolated
// per pair from the pair's reference entries (c6_pairwise) at the CNs of i (cni) and j.
// ONE: a handle of one element type -- the partner's type is 0 (not loaded) and the centre's C6
// row and damping constants are registers, not LDS.
// NPAY = NR: the partner's weights wq came with the candidate; else gathered here from the rows wr.
template <class T, int NR, bool ONE, int NPAY>
__device__ __forceinline__ void pair_2b(const Params<T> &P, const T *const *wr, const T *wq, int j, T dx, T dy,
                                        T dz, const T *y, const T *dy_, const T *rq, const T *rR, const T *rX,
                                        int ti, T cni, Acc2b<T> &A) {
    const int tj = ONE ? 0 : at32(P.type_s, (unsigned)j);
    // the partner's weights are gathered first; the damping factors below do not need them, so
    // their arithmetic covers the gathers' latency
    T wj[NR > 0 ? NR : 1];
    if constexpr (NR > 0) {
#pragma unroll
        for (int b = 0; b < NR; ++b) wj[b] = NPAY == NR ? wq[b] : at32(wr[b], (unsigned)j);
    }
    const T r2 = dx * dx + dy * dy + dz * dz;
    // e = E1 C6 is E+ of the whole pair (the centre's share is half of it: the sums are halved
    // once, in k_2b), g = E1 dC6/dCN_i, phi = P1 C6 = (1/r) dE+/dr; the virial sums collect
    // phi d_a d_b (times -1/2 there)
    T E1, P1;
    if (P.damping == 0) {
        const T q3 = T(3) * rq[tj], R = rR[tj];  // C8 = 3 q C6
        const T ir = prsqrt(r2);
        const T ir2 = ir * ir;
        const T ir6 = ir2 * ir2 * ir2, ir8 = ir6 * ir2;
        const T t = R * ir;
        const T b6 = P.p1 * t, b8 = P.p2 * t;
        const T b6_2 = b6 * b6, b6_4 = b6_2 * b6_2, b6_8 = b6_4 * b6_4, b6_14 = b6_8 * b6_4 * b6_2;
        const T b8_2 = b8 * b8, b8_4 = b8_2 * b8_2, b8_8 = b8_4 * b8_4, b8_16 = b8_8 * b8_8;
        const T f6 = finite_or_zero(prcp(T(1) + T(6) * b6_14)), f8 = finite_or_zero(prcp(T(1) + T(6) * b8_16));
        const T df6 = finite_or_zero(T(84) * f6 * f6 * b6_14 * ir), df8 = finite_or_zero(T(96) * f8 * f8 * b8_16 * ir);
        E1 = P.s6 * f6 * ir6 + P.s8 * q3 * f8 * ir8;
        P1 = (P.s6 * (df6 - T(6) * f6 * ir) * ir6 + P.s8 * q3 * (df8 - T(8) * f8 * ir) * ir8) * ir;
    } else {
        // BJ: f_n r^-n = 1 / (r^n + a^n), so with rq = 3 s8 r2r4_A r2r4_B, rR = a^6, rX = a^8,
        // d6 = 1 / (r^6 + a^6), d8 = 1 / (r^8 + a^8), t8 = rq d8:
        //   E+ = C6 (s6 d6 + t8),  dE+/dCN = dC6 (s6 d6 + t8),
        //   (1/r) dE+/dr = -C6 (6 s6 d6^2 r^4 + 8 t8 d8 r^6) = -C6 r^4 (8 t8 d8 r^2 + 6 s6 d6^2)
        // (no 1/r anywhere; a denominator below the smallest normal number -- it can only have
        // underflowed -- contributes nothing; tested before the reciprocal, so the select does
        // not wait for it)
        const T r4 = r2 * r2;
        const T n6 = r4 * r2 + rR[tj], n8 = r4 * r4 + rX[tj];
        const T tiny = sizeof(T) == sizeof(float) ? T(FLT_MIN) : T(DBL_MIN);
        const T d6 = n6 >= tiny ? prcp(n6) : T(0), d8 = n8 >= tiny ? prcp(n8) : T(0);
        const T t8 = rq[tj] * d8;
        E1 = P.s6 * d6 + t8;
        const T u8 = T(8) * t8 * d8, u6 = (T(6) * P.s6) * d6 * d6;
        P1 = -r4 * (u8 * r2 + u6);
    }
    T c6 = T(0), dc6 = T(0);
    if constexpr (NR == 0) {
        T dc6j;
        c6_pairwise(P.c6g, P.ne, P.nref, ti, tj, cni, at32(P.cn, (unsigned)j), c6, dc6, dc6j);
    } else {
#pragma unroll
        for (int b = 0; b < NR; ++b) {
            c6 += y[tj * NR + b] * wj[b];
            dc6 += dy_[tj * NR + b] * wj[b];
        }
    }
    const T e = c6 * E1, g = dc6 * E1, phi = c6 * P1;
    const T fx = phi * dx, fy = phi * dy, fz = phi * dz;
    const T vv[6] = {fx * dx, fy * dy, fz * dz, fx * dy, fx * dz, fy * dz};
    A.add(e, g, fx, fy, fz, vv);
}

// FP32 with up to 3 references: the partner's CN weights travel with the candidate (queue payload)
template <class T>
__host__ __device__ constexpr int twobody_npay(int nref) {
    return sizeof(T) == sizeof(float) && nref >= 1 && nref <= 3 ? nref : 0;
}
template <class T>
__host__ __device__ constexpr size_t twobody_queue_bytes(int nref) {
    return sizeof(Queue<T, 0>) + (size_t)twobody_npay<T>(nref) * QCAP * sizeof(T);
}
template <class T>
__host__ __device__ constexpr size_t twobody_wave_lds(int ne, int nref) {
    return ((twobody_queue_bytes<T>(nref) + (size_t)(2 * ne * nref + 3 * ne) * sizeof(T)) + 15) / 16 * 16;
}

// At least 5 waves per SIMD (96 VGPRs; the few values spilled live outside the pair loop): the
// gathers of a pair step wait less (Cu two-body FP32 -4.5 %, FP64 -3 %; 6 waves gain little
// more in FP32 and cost FP64 its pair loop).
template <class T, int NR, bool ONE>
__global__ __launch_bounds__(BLOCK, 5) void k_2b(Params<T> P) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int wv = threadIdx.x >> 6, lane = lane_id();
    const int w = blockIdx.x * WPB + wv, nparts = P.tb_parts, i = w / nparts, part = w - i * nparts;
    if (i >= P.n) return;
    const int ynr = P.ne * NR;
    constexpr int NPAY = twobody_npay<T>(NR);
    unsigned char *base = smem + wv * twobody_wave_lds<T>(P.ne, NR);
    Queue<T, NPAY> &Q = *reinterpret_cast<Queue<T, NPAY> *>(base);
    static_assert(sizeof(Queue<T, NPAY>) == twobody_queue_bytes<T>(NR), "queue layout");
    T *y = reinterpret_cast<T *>(base + sizeof(Queue<T, NPAY>));
    T *dyv = y + ynr, *rq = dyv + ynr, *rR = rq + P.ne, *rX = rR + P.ne;
    T y1[ONE ? NR : 1], dy1[ONE ? NR : 1], rq1[1], rR1[1], rX1[1];  // ONE: the same, in registers

    Centre<T> C;
    C.load(P, i);
    const T cni = NR == 0 ? P.cn[i] : T(0);
    if constexpr (ONE) {
        static_assert(NR > 0, "one element type: factorised rows only");
        T wi[NR], dwi[NR];
        for (int a = 0; a < NR; ++a) {
            wi[a] = P.w[a * P.n + i];
            dwi[a] = P.dw[a * P.n + i];
        }
#pragma unroll
        for (int bb = 0; bb < NR; ++bb) {
            T sy = T(0), sdy = T(0);
            for (int a = 0; a < NR; ++a) {
                const T c = P.c6[a * NR + bb];
                sy += c * wi[a];
                sdy += c * dwi[a];
            }
            y1[bb] = sy;
            dy1[bb] = sdy;
        }
        const T q = P.q4[0], R = P.rdamp[0];
        if (P.damping == 0) {
            rq1[0] = q;
            rR1[0] = R;
            rX1[0] = T(0);
        } else {
            const T a2 = R * R, a6 = a2 * a2 * a2;
            rq1[0] = T(3) * P.s8 * q;
            rR1[0] = a6;
            rX1[0] = a6 * a2;
        }
    } else if constexpr (NR > 0) {
        T wi[NR], dwi[NR];
        for (int a = 0; a < NR; ++a) {
            wi[a] = P.w[a * P.n + i];
            dwi[a] = P.dw[a * P.n + i];
        }
        for (int idx = lane; idx < ynr; idx += WAVE) {
            const int B = idx / NR, bb = idx % NR;
            T sy = T(0), sdy = T(0);
            for (int a = 0; a < NR; ++a) {
                const T c = P.c6[((C.ti * P.ne + B) * NR + a) * NR + bb];
                sy += c * wi[a];
                sdy += c * dwi[a];
            }
            y[idx] = sy;
            dyv[idx] = sdy;
        }
    }
    if constexpr (!ONE) {
        for (int B = lane; B < P.ne; B += WAVE) {
            const T q = P.q4[C.ti * P.ne + B], R = P.rdamp[C.ti * P.ne + B];
            if (P.damping == 0) {  // zero damping: q and R0
                rq[B] = q;
                rR[B] = R;
            } else {  // BJ: 3 s8 q, a^6, a^8
                const T a2 = R * R, a6 = a2 * a2 * a2;
                rq[B] = T(3) * P.s8 * q;
                rR[B] = a6;
                rX[B] = a6 * a2;
            }
        }
    }
    wave_sync();

    Acc2b<T> A;
    A.init();
    float cnt = 0.0f;
    // this wave's share of the stencil (all of it with one part per centre)
    const int sa = C.S->st_off[CUT_2B], sz = C.S->st_cnt[CUT_2B], per = (sz + nparts - 1) / nparts;
    const int s0 = sa + min(sz, part * per), s1 = sa + min(sz, (part + 1) * per);
    const T *wr[NR > 0 ? NR : 1];  // rows of the weights, one per reference
    for (int b = 0; b < NR; ++b) wr[b] = sbase(P.w + (size_t)b * P.n);
    auto pay = [&](unsigned jj, T *v) __attribute__((always_inline)) {
#pragma unroll
        for (int b = 0; b < NPAY; ++b) v[b] = at32(wr[b], jj);
    };
    walk_pairs<T, true, sizeof(T) == sizeof(float), NPAY>(P, C, s0, s1, xsqrt(P.rc_22), P.rc_22, Q, pay,
                                                           [&](int j, T dx, T dy, T dz, const T *wq) {
        if constexpr (ONE)
            pair_2b<T, NR, true, NPAY>(P, wr, wq, j, dx, dy, dz, y1, dy1, rq1, rR1, rX1, 0, cni, A);
        else
            pair_2b<T, NR, false, NPAY>(P, wr, wq, j, dx, dy, dz, y, dyv, rq, rR, rX, C.ti, cni, A);
        cnt += 1.0f;
    });
    A.fold();
    // the centre's half of each pair's energy and virial (scaling by 1/2 is exact)
    const double E = 0.5 * wave_total(A.e), G = wave_total(A.g);
    double F[3], V[6];
    for (int k = 0; k < 3; ++k) F[k] = wave_total(A.f[k]);
    for (int k = 0; k < 6; ++k) V[k] = -0.5 * wave_total(A.v[k]);
    cnt = wave_sumf(cnt);
    if (lane == 0) {
        const double iv = C.S->inv_vol;
        if (nparts == 1) {
            P.e[i] = E;
            P.g[i] = G;
            for (int k = 0; k < 3; ++k) P.f[3 * i + k] = F[k];
            for (int k = 0; k < 6; ++k) P.vir[6 * i + k] = V[k] * iv;
            P.npair[3 * i + 0] = cnt;
        } else {
            double *o = P.tb_part + ((size_t)part * P.n + i) * TB_PART_W;
            o[0] = E;
            o[1] = G;
            for (int k = 0; k < 3; ++k) o[2 + k] = F[k];
            for (int k = 0; k < 6; ++k) o[5 + k] = V[k] * iv;
            o[11] = cnt;
        }
    }
}

// Per-centre two-body sums from the parts, in part order (bitwise reproducible).
template <class T>
__global__ void k_2b_parts(Params<T> P) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.n) return;
    double acc[TB_PART_W];
    for (int k = 0; k < TB_PART_W; ++k) acc[k] = 0.0;
    for (int c = 0; c < P.tb_parts; ++c) {
        const double *o = P.tb_part + ((size_t)c * P.n + i) * TB_PART_W;
        for (int k = 0; k < TB_PART_W; ++k) acc[k] += o[k];
    }
    P.e[i] = acc[0];
    P.g[i] = acc[1];
    for (int k = 0; k < 3; ++k) P.f[3 * i + k] = acc[2 + k];
    for (int k = 0; k < 6; ++k) P.vir[6 * i + k] = acc[5 + k];
    P.npair[3 * i + 0] = (float)acc[11];
}

// ---------------------------------------------------------------------------------------------
// dE/dCN chain rule, src/d3_kernel.cu:1511-1556 and 1571-1698: one wave per centre over the CN
// pair set, F_i += (g_i + g_j) c'(r)/r d,  vir += -(g_i + g_j) c'(r)/r d d^T / 2,
// c'(r)/r = -k1 e (Rcov_i + Rcov_j) / ((1 + e)^2 r^3).
// ---------------------------------------------------------------------------------------------
template <class T>
constexpr int chain_min_waves() { return sizeof(T) == sizeof(float) ? 5 : 1; }  // FP32 -14 % on Cu

template <class T, bool ONE>
__global__ __launch_bounds__(BLOCK, chain_min_waves<T>()) void k_chain(Params<T> P) {
    __shared__ Queue<T> Qs[WPB];
    const int lane = lane_id(), wv = threadIdx.x >> 6;
    const int i = blockIdx.x * WPB + wv;
    if (i >= P.n) return;
    Centre<T> C;
    C.load(P, i);
    const T rci = P.rcov[C.ti];
    const T gi = (T)P.g[i];
    AccN<T, 9> A;  // F (3), virial (6)
    A.init();
    const int s0 = C.S->st_off[CUT_CN], s1 = s0 + C.S->st_cnt[CUT_CN];
    walk_pairs<T, false, sizeof(T) == sizeof(float), 0>(P, C, s0, s1, P.rc_cn, P.rc_cn2, Qs[wv], NoPay<T>{},
                                                         [&](int j, T dx, T dy, T dz, const T *) {
        const T gij = gi + (T)at32(P.g, (unsigned)j);  // issued first: its latency overlaps the rest
        const T r2 = dx * dx + dy * dy + dz * dz;
        const T ir = prsqrt(r2);
        const T rr = ONE ? rci + rci : rci + P.rcov[at32(P.type_s, (unsigned)j)];
        const T ex = xexp(-T(K1) * (rr * ir - T(1)));
        const T c = prcp(T(1) + ex);
        T dcn = -T(K1) * ex * rr * ir * ir * ir * c * c;
        if (P.cn_smooth) {
            T cs = c;
            smooth_cn_term(P.rc_cn, r2 * ir, ir, cs, dcn);
        }
        const T sc = finite_or_zero(gij * dcn);
        const T h = T(-0.5) * sc;
        const T hx = h * dx, hy = h * dy, hz = h * dz;
        const T v[9] = {sc * dx, sc * dy, sc * dz, hx * dx, hy * dy, hz * dz, hx * dy, hx * dz, hy * dz};
        A.add(v);
    });
    if constexpr (sizeof(T) == sizeof(float)) A.fold();
    double Fd[3], Vd[6];
    for (int k = 0; k < 3; ++k) Fd[k] = wave_total(A.t[k]);
    for (int k = 0; k < 6; ++k) Vd[k] = wave_total(A.t[3 + k]);
    if (lane == 0) {
        const double iv = C.S->inv_vol;
        for (int k = 0; k < 3; ++k) P.f[3 * i + k] += Fd[k];
        for (int k = 0; k < 6; ++k) P.vir[6 * i + k] += Vd[k] * iv;
    }
}

}  // namespace

template <class T>
void Launch<T>::bin_sort(const Params<T> &P, int nbins, int *bin_start_mut, hipStream_t s) {
    const int tb = 256;
    hipLaunchKernelGGL(k_fill_i32, dim3((nbins + tb - 1) / tb), dim3(tb), 0, s, P.bin_count, nbins, 0);
    if (P.n > 0) hipLaunchKernelGGL(k_bin<T>, dim3((P.n + tb - 1) / tb), dim3(tb), 0, s, P);
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, s, P.bin_count, bin_start_mut, P.cursor, nbins);
    if (P.n > 0) {
        hipLaunchKernelGGL(k_scatter<T>, dim3((P.n + tb - 1) / tb), dim3(tb), 0, s, P);
        hipLaunchKernelGGL(k_canon, dim3(nbins), dim3(tb), 0, s, P.bin_start, P.perm, P.perm2);
        hipLaunchKernelGGL(k_gather<T>, dim3((P.n + tb - 1) / tb), dim3(tb), 0, s, P);
    }
}

template <class T>
void Launch<T>::regather(const Params<T> &P, hipStream_t s) {
    if (P.n > 0) hipLaunchKernelGGL(k_regather<T>, dim3((P.n + 255) / 256), dim3(256), 0, s, P);
}

template <class T>
void Launch<T>::cn(const Params<T> &P, hipStream_t s) {
    if (P.n == 0) return;
    if (P.ne == 1)
        hipLaunchKernelGGL((k_cn<T, true>), dim3((P.n + WPB - 1) / WPB), dim3(BLOCK), 0, s, P);
    else
        hipLaunchKernelGGL((k_cn<T, false>), dim3((P.n + WPB - 1) / WPB), dim3(BLOCK), 0, s, P);
}

template <class T, bool ONE, int... NR>
const void *twobody_pick(int nref, std::integer_sequence<int, NR...>) {
    const void *fn = nullptr;
    ((fn = (nref == NR + 1) ? reinterpret_cast<const void *>(&k_2b<T, NR + 1, ONE>) : fn), ...);
    return fn;
}

template <class T>
void Launch<T>::twobody(const Params<T> &P, hipStream_t s) {
    if (P.n == 0) return;
    const size_t lds = WPB * twobody_wave_lds<T>(P.ne, P.c6_pairwise ? 0 : P.nref);
    const auto seq = std::make_integer_sequence<int, MAX_NREF>{};
    const void *fn = P.c6_pairwise ? reinterpret_cast<const void *>(&k_2b<T, 0, false>)
                     : P.ne == 1   ? twobody_pick<T, true>(P.nref, seq)
                                   : twobody_pick<T, false>(P.nref, seq);
    if (!fn) return;  // not reached: a handle's reference count is 1..MAX_NREF
    Params<T> arg = P;
    void *args[] = {&arg};
    const size_t waves = (size_t)P.n * P.tb_parts;
    (void)hipLaunchKernel(fn, dim3((unsigned)((waves + WPB - 1) / WPB)), dim3(BLOCK), args, lds, s);
    if (P.tb_parts > 1) hipLaunchKernelGGL(k_2b_parts<T>, dim3((P.n + 255) / 256), dim3(256), 0, s, P);
}

template <class T>
void Launch<T>::chain(const Params<T> &P, hipStream_t s) {
    if (P.n == 0) return;
    if (P.ne == 1)
        hipLaunchKernelGGL((k_chain<T, true>), dim3((P.n + WPB - 1) / WPB), dim3(BLOCK), 0, s, P);
    else
        hipLaunchKernelGGL((k_chain<T, false>), dim3((P.n + WPB - 1) / WPB), dim3(BLOCK), 0, s, P);
}

template struct Launch<double>;
template struct Launch<float>;

}  // namespace d3
