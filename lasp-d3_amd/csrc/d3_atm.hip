// Axilrod-Teller-Muto three-body term, replacing the reference's atm_kernel
// (src/d3_kernel.cu:1199-1509).  Every periodic triangle is evaluated exactly once.
//
// Enumeration.  Every vertex of a periodic triangle lies in a bin of the unbounded (unwrapped)
// bin grid; the anchor of a triangle is its vertex whose unwrapped bin comes first in
// lexicographic (z, y, x) order, and among vertices in the same unwrapped bin (distinct atoms,
// since one atom occupies one point per bin image) the one with the smallest atom index.  A centre
// i therefore keeps only its "forward" neighbours within rc3: atoms in the bins at stencil
// offsets d >_lex 0 (any atom, i itself at another image included), and atoms j > i of its own
// bin (d = 0); each triangle appears once, as a pair {j, k} of forward neighbours of its anchor
// with r_jk <= rc3.  The test is on integers (offsets, indices), so it is exact, and every centre
// keeps the upper half of its stencil (an anchor rule by atom index alone would give the centres
// in the first bin layers of the cell, whose lower neighbours wrap to the top layers, nearly
// whole spheres).  (The reference enumerates i >= j >= k over every image and rescales
// index-degenerate triplets by 1/2 or 1/6.)
//
// Work layout.  One wave per centre; persistent one-wave blocks take centres from an atomic
// counter.  Phase 1 compacts the forward neighbours (ballot/mbcnt) into per-block scratch;
// phase 2 turns them into records (C6_ij and its logarithmic CN derivatives, 1/a, a^-1/2,
// sqrt C6_ij, damping term); phase 3 visits the pairs in 64 x 64 tiles: lane l owns neighbour
// k = 64 kb + l in registers (FP32: two k tiles, kb and kb + 1), the j tile sits in LDS with its
// C6 rows y_j[B] = C6ref_{A_j B}^T w_j and dy_j[B] = C6ref^T dw_j, and at step q lane l takes
// j = l xor q (on a tile paired with itself, j = (l + q) mod 64); with two k tiles a step reads
// the j slot and writes its sums back once for two pairs.  Tile pairs whose bounding boxes are
// farther apart than rc3 are skipped.  Phase 4 turns the per-neighbour sums
// into forces, dE/dCN, energies and the virial and adds them to the per-atom arrays with FP64
// atomics.
//
// Per triangle (a = r_ij^2, b = r_ik^2, c = r_jk^2, d_j = r_i - r_j, d_k = r_i - r_k):
//   e = C9 A f,  C9 = -sqrt(C6_ij C6_ik C6_jk)
//   A = (abc)^(-3/2) + (3/8) Q (abc)^(-5/2),  Q = (a+b-c)(a-b+c)(-a+b+c)   [= (1 + 3 cos cos cos)/(r r r)^3]
//   f = 1 / (1 + 6 x^16),  x^6 = y = (4/3)^6 R0_ij^2 R0_ik^2 R0_jk^2 / (abc),  x^16 = y^(8/3)
//   t_xy = 2 de/d(r_xy^2) = 2 C9 f [(P1 + (8/3) A (1 - f)) / r_xy^2 + (3/8) iR5 dQ/d(r_xy^2)],
//     P1 = -(3/2)(abc)^(-3/2) - (15/16) Q (abc)^(-5/2),  (6 x^16 f = 1 - f)
//   grad_j e = -(t_ij + t_jk) d_j + t_jk d_k = -t_ij d_j + t_jk v,
//   grad_k e = -(t_ik + t_jk) d_k + t_jk d_j = -t_ik d_k - t_jk v,  v = d_k - d_j,
//   grad_i e = -(grad_j e + grad_k e)
//   de/dCN_x = (C9/2)(dlnC6_xy/dCN_x + dlnC6_xz/dCN_x) A f
//   each vertex's atom takes e/3 of the energy (the reference's split, src/d3_kernel.cu:1394-1399)
// FP32 evaluates x^16 as exp2((8/3) log2 y) with log2 y split into per-neighbour, per-type and
// per-pair parts; FP64 as y^3 y^(-1/3).
#include <hip/hip_runtime.h>

#include <type_traits>
#include <utility>

#include "d3_device.hpp"

namespace d3 {
namespace {

constexpr int AW = 64;                          // one wave per block
constexpr double K6 = 4096.0 / 729.0;           // (4/3)^6
constexpr double L2K6_83 = 6.6405999884615;     // (8/3) log2((4/3)^6)
constexpr float FAR = 1e18f;                    // coordinate of padding lanes (fails every cutoff)
// record fields (global scratch, one array of cap entries per field)
// (r^2, 1/r^2, 1/r, sqrt C6 and the damping term are recomputed where a tile is staged: they
// cost a few instructions per slot and tile, against twice the record bytes read back from HBM)
enum { F_DX, F_DY, F_DZ, F_DLI, F_DLJ, F_SC6 };  // offset, CN log-derivatives of C6_ip, sqrt C6_ip
static_assert(F_SC6 + 1 == ATM_NF, "ATM_NF record fields per neighbour slot");
// per-neighbour FP64 accumulators over the triangles {i, p, q} of the centre: A_p = sum t'_ip,
// B_p = sum t'_pq (d_q - d_p), H_p = sum e/2, Gc_p = sum (e/2) dlnC6_pq/dCN_p, with t' = (2/3) t,
// so that grad_p = (3/2) (B_p - A_p d_p)
enum { A_A, A_BX, A_BY, A_BZ, A_H, A_G, A_NF };
static_assert(A_NF == 6, "atm_acc_stride holds 6 doubles per neighbour slot");

template <class T>
struct alignas(2 * sizeof(T)) T2 {
    T x, y;
};

// y^(-1/3) in FP64: float guess and two Newton steps z <- z (4 - y z^3) / 3.
__device__ __forceinline__ double inv_cbrt(double y) {
    double z = (double)__builtin_amdgcn_exp2f(-(1.0f / 3.0f) * __builtin_amdgcn_logf((float)y));
    z = z * (4.0 - y * z * z * z) * (1.0 / 3.0);
    z = z * (4.0 - y * z * z * z) * (1.0 / 3.0);
    return z;
}
// Per-pair reciprocal and reciprocal square root: the FP32 hardware instructions (1 ulp),
// correctly rounded FP64.
__device__ __forceinline__ float prcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double prcp(double x) { return 1.0 / x; }
__device__ __forceinline__ float prsqrt(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ double prsqrt(double x) { return 1.0 / sqrt(x); }

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

// LDS carve-up, shared by the kernel and the host-side size query.
template <class T>
struct AtmLds {
    // Rows of 64 pair-sized slots (T2), all at fixed offsets before the C6 rows, so that a pair
    // step addresses every per-j value from one register (the slot's byte offset) plus immediates:
    //   J0..J2  j tile: (dx, dy) (dz, a) (1/a, a^-1/2)
    //   J3      (sqrt C6_ij, lam_j): without DR, or with pairwise C6 (the steps of the DR form
    //           with per-element references read neither)                  [0 or 1 row]
    //   Jp      j-side tile sums as pairs (A, Bx) (By, Bz) (H, G)        [3 rows]
    //   Jt      without DR: the j type x ne, one int per slot (and the type) [0 or 1 row]
    //   Jyd     per type B of k: with DR, one damping row (FP32: lam_j + (8/3) log2
    //           R0_{A_j B}^2, FP64: lam_j R0_{A_j B}^2; 0), then the type's reference count of
    //           rows (y, dy); then padding rows.  A pair step reads NR rows from the start of
    //           type(k)'s C6 rows: for a type with fewer references the extra rows are the next
    //           type's (finite values) or padding, met by k weights that are exactly zero.
    // The fixed rows (J0..Jt) are a static LDS array of the kernel, so their addresses are
    // link-time constants and a step's slot offset needs no base register; the rest is dynamic
    // LDS: Jyd, then one region that holds the centre's C6 rows Cy, Cdy [ne * nref] in phase 2
    // and, with KL, the k-side sums Khi, Klo [A_NF][64] in phase 3 (the pair terms read sqrt C6
    // from the neighbour records, not from Cy), then without DR the per-type-pair damping terms
    // Rt [ne * ne] (with DR the staging computes them from r0).  Fcc Cu, FP32, KL + DR: 6 fixed
    // and 4 C6 rows (static, one element type) + the sums of two k tiles = 11,264 bytes; DR alone
    // 5,120 bytes.
    T2<T> *J0, *J1, *J2, *J3, *Jp, *Jyd;
    int *Jt;
    T *Khi, *Klo;  // k-side sums over the j tiles (two-sum pair in FP32; KL only)
    T *Cy, *Cdy;   // centre rows (phase 2)
    T *Rt;         // per type pair: FP32 (8/3) log2 R0^2, FP64 R0^2 (without DR)
    static constexpr int TI = sizeof(T2<T>) / sizeof(int);  // ints per slot
    __host__ __device__ static constexpr bool has_j3(bool dr, bool pairwise) { return !dr || pairwise; }
    __host__ __device__ static constexpr int jp_row(bool j3) { return j3 ? 4 : 3; }
    __host__ __device__ static constexpr int jt_row(bool j3) { return jp_row(j3) + 3; }
    __host__ __device__ static constexpr int fixed_rows(bool dr, bool j3) { return jt_row(j3) + (dr ? 0 : 1); }
    __host__ __device__ static constexpr size_t fixed_bytes(bool dr, bool j3) {
        return (size_t)fixed_rows(dr, j3) * AW * sizeof(T2<T>);
    }
    __host__ __device__ static int jy_rows(int ne, int yrows, bool dr) { return yrows + (dr ? ne : 0); }
    // k tiles a pass of phase 3 holds: two in FP32 (each j slot read and written once for two
    // pairs), one in FP64 (two would need the registers of a wave of the SIMD's two or three)
    static constexpr int KTILES = sizeof(T) == 4 ? 2 : 1;
    __host__ __device__ static size_t region_bytes(int ne, int nref, bool kl) {
        // KL: (hi, lo) columns for each k tile of a pass
        const size_t sums = kl ? (size_t)2 * KTILES * A_NF * AW * sizeof(T) : 0, rows = (size_t)2 * ne * nref * sizeof(T);
        return sums > rows ? sums : rows;
    }
    // dynamic LDS bytes of a launch (nref 0: pairwise C6).  one: a single element type, whose
    // rows Jyd are part of the kernel's static LDS (their offsets then are immediates too)
    __host__ __device__ static size_t dyn_bytes(int ne, int nref, int yrows, bool kl, bool dr, bool one = false) {
        return (one ? 0 : (size_t)jy_rows(ne, yrows, dr) * AW * sizeof(T2<T>)) + region_bytes(ne, nref, kl) +
               (dr ? 0 : (size_t)ne * ne * sizeof(T));
    }
    // all LDS of a block (static + dynamic)
    __host__ __device__ static size_t bytes(int ne, int nref, int yrows, bool kl, bool dr) {
        return fixed_bytes(dr, has_j3(dr, nref == 0)) + dyn_bytes(ne, nref, yrows, kl, dr);
    }
    __device__ AtmLds(T2<T> *fixed, unsigned char *s, int ne, int nref, int yrows, bool kl, bool dr, bool one) {
        const bool j3 = has_j3(dr, nref == 0);
        J0 = fixed;
        J1 = J0 + AW;
        J2 = J0 + 2 * AW;
        J3 = j3 ? J0 + 3 * AW : nullptr;
        Jp = J0 + jp_row(j3) * AW;
        Jt = dr ? nullptr : reinterpret_cast<int *>(J0 + jt_row(j3) * AW);
        Jyd = one ? J0 + fixed_rows(dr, j3) * AW : reinterpret_cast<T2<T> *>(s);
        unsigned char *region = one ? s : reinterpret_cast<unsigned char *>(Jyd + (size_t)jy_rows(ne, yrows, dr) * AW);
        Khi = reinterpret_cast<T *>(region);
        Klo = Khi + A_NF * AW;
        Cy = reinterpret_cast<T *>(region);
        Cdy = Cy + (size_t)ne * nref;
        Rt = dr ? nullptr : reinterpret_cast<T *>(region + region_bytes(ne, nref, kl));
    }
};

// First forward stencil entry of a cell's three-body stencil: its middle entry, the offset d = 0
// (the stencil is symmetric under d -> -d and ordered lexicographically by (dz, dy, dx); the host
// checks that the middle entry is d = 0).
__device__ __forceinline__ int fwd_first(const SysDev *S) { return S->st_off[CUT_ATM] + (S->st_cnt[CUT_ATM] - 1) / 2; }

// Waves per SIMD the compiler must allow: FP32 with 3 or more references at 4 waves (128 VGPRs;
// with 5 references some spilling, which still leaves the water box 13 % and the six-element
// batch 4 % faster than at 3 waves); 1 or 2 references keep the default (diamond Si, 2
// references: 8 % faster than forced to 4 waves); FP64 with 3 references at 3 waves (167 VGPRs,
// one spilled pointer outside the loops).
template <class T, int NR>
constexpr int atm_min_waves() {
    return sizeof(T) == 4 ? (NR >= 3 ? 4 : 1) : (NR == 3 ? 3 : 1);
}

// KL: the k side's tile sums go to lane-private LDS (two-sum pairs) and reach the FP64
// accumulators once per k tile; otherwise they are added to them after every tile.  DR: the
// per-type-pair damping term is staged as an extra row per type of k (two dependent LDS reads and
// three VALU fewer per pair, ne rows more LDS).  The host picks them from the resident block
// counts (atm_choose_variant); the autotuner times all four.  ONE: a handle of one element type,
// whose rows of type(k) start at a fixed LDS offset (an immediate of the step's LDS reads instead
// of a per-lane row pointer added at every step).
template <class T, int NR, bool KL, bool DR, bool ONE>
__global__ __launch_bounds__(AW, (atm_min_waves<T, NR>())) void k_atm(Params<T> P) {
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr bool J3 = AtmLds<T>::has_j3(DR, NR == 0);  // the (sqrt C6_ij, lam_j) row is staged
    constexpr int JP = AtmLds<T>::jp_row(J3), JT = AtmLds<T>::jt_row(J3);  // rows of Jp, Jt
    // the fixed rows, and with a single element type its (damping,) C6 rows
    __shared__ T2<T> fixed[(AtmLds<T>::fixed_rows(DR, J3) + (ONE ? NR + (DR ? 1 : 0) : 0)) * AW];
    constexpr bool F32 = sizeof(T) == 4;
    static_assert(NR > 0 || !DR, "the pairwise-C6 form (NR = 0) stages no C6 or damping rows");
    constexpr int DRO = DR ? 1 : 0;  // C6 rows of a type start after its damping row
    const int lane = threadIdx.x;
    const int ne = P.ne, ynr = ne * NR;
    AtmLds<T> S(fixed, smem, ne, NR, P.atm_yrows, KL, DR, ONE);

    const int cap = P.atm_cap;
    int *Lj = P.atm_i32 + (size_t)blockIdx.x * cap * 2;
    int *Lt = Lj + cap;
    T *LT = P.atm_T + (size_t)blockIdx.x * atm_T_stride(cap);
    T *LB = LT + (size_t)ATM_NF * cap;  // per tile: min x, max x, min y, max y, min z, max z
    double *La = P.atm_acc + (size_t)blockIdx.x * atm_acc_stride(cap);
    const T rc32 = P.rc_32;

    // per-type-pair damping terms, FP32 (8/3) log2 R0^2, FP64 R0^2 (in LDS without DR)
    auto rt_of = [&](int idx) -> T {
        const T r0sq = P.r0[idx] * P.r0[idx];
        return F32 ? T((8.0f / 3.0f) * __builtin_amdgcn_logf((float)r0sq)) : r0sq;
    };
    if constexpr (!DR)
        for (int idx = lane; idx < ne * ne; idx += AW) S.Rt[idx] = rt_of(idx);
    // lane B (< ne <= 64): first C6 row of type B, read with readlane / a lane shuffle instead of
    // dependent loads in the tile loops
    const int roff_l = lane < ne ? P.atm_roff[lane] : 0;

    for (;;) {
        // Ticket: every lane takes part in the atomic (lane 0 adds 1, the others 0), and the value
        // is read back as a scalar.  No lane-dependent branch may sit in this loop: the compiler
        // may split a loop on such a branch into an inner loop for lanes 1..63 and an outer one
        // for lane 0, and lanes 1..63 would then run on without ever taking a ticket.
        const int ticket = __builtin_amdgcn_readfirstlane(atomicAdd(P.atm_next, lane == 0 ? 1 : 0));
        const int nch = P.atm_chunks, i = ticket / nch, chunk = ticket - i * nch;
        if (i >= P.n) break;
        Centre<T> C;
        C.load(P, i);
        const int ti = ONE ? 0 : C.ti;
        const T cni = NR == 0 ? P.cn[i] : T(0);  // pairwise C6 only
        __syncthreads();  // previous centre done with LDS and scratch
        if constexpr (NR > 0)
            for (int idx = lane; idx < ynr; idx += AW) {
                const int B = idx / NR, bb = idx % NR;
                T sy = T(0), sdy = T(0);
                for (int a = 0; a < NR; ++a) {
                    const T c = P.c6[((ti * ne + B) * NR + a) * NR + bb];
                    sy += c * P.w[a * P.n + i];
                    sdy += c * P.dw[a * P.n + i];
                }
                S.Cy[idx] = sy;
                S.Cdy[idx] = sdy;
            }
        __syncthreads();

        // ---- phase 1: forward neighbours within rc3 (1e-12 <= r^2 <= rc3^2, src/d3_kernel.cu:1268),
        // compacted in stencil-walk order: from the centre's forward list when the handle keeps one
        // (the same candidates in the same order, so the same compaction), else by the walk.
        int count = 0;
        auto take = [&](int j, T dx, T dy, T dz, bool in) {
            const uint64_t m = ballot64(in);
            if (m == 0) return;
            if (in) {
                const int pos = count + mbcnt64(m);
                if (pos < cap) {
                    Lj[pos] = j;
                    LT[F_DX * cap + pos] = dx;
                    LT[F_DY * cap + pos] = dy;
                    LT[F_DZ * cap + pos] = dz;
                }
            }
            count += __popcll(m);
        };
        if (P.atm_fl) {
            const FwdEntry *FL = P.atm_fl + (size_t)i * P.atm_fl_cap;
            const int nl = min(P.atm_fl_n[i], P.atm_fl_cap);
            for (int fb = 0; fb < nl; fb += AW) {
                const int f = fb + lane;
                int j = 0;
                bool in = false;
                T dx = T(0), dy = T(0), dz = T(0);
                if (f < nl) {
                    const FwdEntry e = FL[f];
                    j = e.j;
                    dx = (C.x + P.st_D[4 * e.s + 0]) - P.lx[j];
                    dy = (C.y + P.st_D[4 * e.s + 1]) - P.ly[j];
                    dz = (C.z + P.st_D[4 * e.s + 2]) - P.lz[j];
                    const T r2 = dx * dx + dy * dy + dz * dz;
                    in = r2 <= rc32 && r2 >= T(1e-12);
                }
                take(j, dx, dy, dz, in);
            }
        } else {
            // forward neighbours: the stencil entries after the centre's own bin (offset d = 0, the
            // middle entry of the symmetric stencil, whose entries run in lexicographic (dz, dy,
            // dx) order), and the atoms of the own bin after i
            const int s0 = fwd_first(C.S), s1 = C.S->st_off[CUT_ATM] + C.S->st_cnt[CUT_ATM];
            const T rc3 = xsqrt(rc32);
            for (int sb = s0; sb < s1; sb += AW) {
                const auto Bn = C.bins(P, sb, s1, rc3);
                const int nsb = s1 - sb < AW ? s1 - sb : AW;
                for (int sk = 0; sk < nsb; ++sk) {
                    int j0, j1;
                    T ox, oy, oz;
                    Bn.get(sk, j0, j1, ox, oy, oz);
                    if (sb + sk == s0) j0 = i + 1;  // own bin, no image shift
                    for (int jb = j0; jb < j1; jb += AW) {
                        const int j = jb + lane;
                        bool in = false;
                        T dx = T(0), dy = T(0), dz = T(0);
                        if (j < j1) {
                            dx = ox - P.lx[j];
                            dy = oy - P.ly[j];
                            dz = oz - P.lz[j];
                            const T r2 = dx * dx + dy * dy + dz * dz;
                            in = r2 <= rc32 && r2 >= T(1e-12);
                        }
                        take(j, dx, dy, dz, in);
                    }
                }
            }
        }
        // Over capacity: flag it and run the centre with no neighbours (its results are zeros);
        // the host retries with a larger capacity.  (No `continue`: one back-edge keeps the
        // loop in the simple shape described at the ticket.)
        if (count > cap && lane == 0) atomicOr(P.status, 2u);
        const int nf = count > cap ? 0 : count;
        const int ntile = (nf + AW - 1) / AW;
        // this work item's k tiles: chunk c of nch takes [ntile sqrt(c / nch), ntile sqrt((c + 1) / nch)),
        // about the same number of tile pairs each (tile kb pairs with jt = 0..kb), with the bounds
        // rounded to even tiles: phase 3 takes k tiles in pairs (2m, 2m + 1) whatever the chunks,
        // so every FP32 tile sum, and hence every result, is the same for any number of work items.
        // It touches the slots of tiles 0..kb_hi-1 only: phases 2 and 4 stop there (the other
        // slots' sums are zero).
        auto even_tile = [&](int c) { return c <= 0 ? 0 : c >= nch ? ntile : min(ntile, 2 * (int)(0.5f * ntile * sqrtf((float)c / nch) + 0.5f)); };
        const int kb_lo = even_tile(chunk), kb_hi = even_tile(chunk + 1);
        const int np = min(nf, kb_hi * AW);
        __syncthreads();  // phase-1 slots were written by arbitrary lanes

        // ---- phase 2: records and each tile's bounding box (float, padding lanes excluded); zero
        // the accumulators.  Slot p is always touched by lane p % 64, here and below.
        for (int t = 0; t < kb_hi; ++t) {
            const int p = t * AW + lane;
            float bb[6] = {3e38f, -3e38f, 3e38f, -3e38f, 3e38f, -3e38f};
            if (p < nf) {
                const int j = Lj[p], tj = ONE ? 0 : P.type_s[j];
                T c6 = T(0), dci = T(0), dcj = T(0);
                if constexpr (NR == 0) {
                    c6_pairwise(P.c6g, ne, P.nref, ti, tj, cni, P.cn[j], c6, dci, dcj);
                } else {
                    for (int b = 0; b < NR; ++b) {
                        const T wj = P.w[b * P.n + j], dwj = P.dw[b * P.n + j];
                        c6 += S.Cy[tj * NR + b] * wj;
                        dci += S.Cdy[tj * NR + b] * wj;
                        dcj += S.Cy[tj * NR + b] * dwj;
                    }
                }
                const T ic6 = c6 > T(0) ? T(1) / c6 : T(0);
                Lt[p] = tj;
                LT[F_DLI * cap + p] = dci * ic6;
                LT[F_DLJ * cap + p] = dcj * ic6;
                LT[F_SC6 * cap + p] = xsqrt(c6);
                // the accumulators take FP64 atomic adds in phase 3 (no returned value, so the wave
                // never waits on them) and are read back with atomic loads (L2, not a stale L1 line)
                for (int q = 0; q < A_NF; ++q)
                    __hip_atomic_store(&La[q * cap + p], 0.0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                const float x = (float)LT[F_DX * cap + p], y = (float)LT[F_DY * cap + p], z = (float)LT[F_DZ * cap + p];
                bb[0] = bb[1] = x;
                bb[2] = bb[3] = y;
                bb[4] = bb[5] = z;
            }
            // all lanes: one box per tile, computed once here instead of at every tile pair
            const float xlo = wmin(bb[0]), xhi = wmax(bb[1]), ylo = wmin(bb[2]), yhi = wmax(bb[3]),
                        zlo = wmin(bb[4]), zhi = wmax(bb[5]);
            if (lane == 0) {
                LB[8 * t + 0] = T(xlo);
                LB[8 * t + 1] = T(xhi);
                LB[8 * t + 2] = T(ylo);
                LB[8 * t + 3] = T(yhi);
                LB[8 * t + 4] = T(zlo);
                LB[8 * t + 5] = T(zhi);
            }
        }
        __syncthreads();  // the boxes were written by lane 0

        // ---- phase 3: unordered pairs {j, k} of forward neighbours.  Per-pair terms are summed in T
        // over one 64 x 64 tile (at most 64 terms per sum); tile sums go to the FP64 per-neighbour
        // accumulators La, which lane l alone touches for its slots, as atomic adds: one lane's
        // adds to one slot reach L2 in issue order, so the sums stay bitwise repeatable, and
        // they cost no read latency (a read-modify-write here took 7.6 % of the kernel on Cu).
        unsigned cnt = 0;  // pairs within rc3 (at most 2^31 per centre)
        const unsigned rc32b = __builtin_bit_cast(unsigned, (float)rc32);
        // Pair (i, p) terms of a forward neighbour p from its offset, type and weights, the same
        // expressions for the k and the j side: r^2, 1/r^2, 1/r, sqrt C6_ip, damping term
        // (FP32: (8/3) log2(R0^2 / r^2), FP64: R0^2 / r^2).
        // FP32 takes the hardware reciprocal, reciprocal square root and log2 (1 ulp, as the pair
        // steps do), FP64 the correctly rounded forms.
        auto pair_terms = [&](T dx, T dy, T dz, int tp, int slot, T &a, T &ia, T &ra, T &sc6, T &lam) {
            a = dx * dx + dy * dy + dz * dz;
            ia = prcp(a);
            ra = prsqrt(a);
            sc6 = at32(LT + F_SC6 * cap, (unsigned)slot);  // from phase 2
            const T r0 = P.r0[ti * ne + tp];
            lam = F32 ? T((8.0f / 3.0f) * __builtin_amdgcn_logf((float)(r0 * r0 * ia))) : r0 * r0 * ia;
        };
        // One k tile: lane l owns neighbour k = 64 kt + l, its pair terms with the centre and its
        // tile sums (per reference b for the CN term: Gb[b] = sum (e / 2 c6) y_b(j), contracted
        // with k's dw_b once per tile, one FMA per pair fewer than forming dC6_jk/dCN_k per pair).
        // (Only what the steps read stays in registers through them: the slot index, validity and
        // LDS columns of a tile are recomputed from its index where the sums are folded.)
        struct KTile {
            T x, y, z, b2, ib, rb, sc6h, lam, fac, cn;
            T w[NR > 0 ? NR : 1], dw[NR > 0 ? NR : 1];
            int t;
            const T2<T> *Yk;  // rows of type(k) in the j tile: its damping row (DR), then its C6 rows
            float lo[3], hi[3];  // bounding box of the tile's valid lanes (wave-uniform)
            T A, Bx, By, Bz, H, G, Gb[NR > 0 ? NR : 1];
        };
        // kl: which of the two k tiles of a pass (LDS columns of its k-side sums with KL)
        auto load_k = [&](KTile &K, int kt, int kl) __attribute__((always_inline)) {
            const int k = kt * AW + lane;
            // padding lanes sit far away, so every pair with them fails the cutoff test
            K.x = K.y = K.z = T(-FAR);
            K.b2 = K.ib = K.rb = T(1);
            K.sc6h = K.lam = K.fac = K.cn = T(0);
            K.t = 0;
            for (int a = 0; a < NR; ++a) K.w[a] = K.dw[a] = T(0);
            if (k < nf) {
                const unsigned jk = (unsigned)at32(Lj, (unsigned)k);
                K.x = at32(LT + F_DX * cap, (unsigned)k);
                K.y = at32(LT + F_DY * cap, (unsigned)k);
                K.z = at32(LT + F_DZ * cap, (unsigned)k);
                K.t = ONE ? 0 : at32(Lt, (unsigned)k);  // one element type: every table index is uniform
                for (int a = 0; a < NR; ++a) {
                    K.w[a] = at32(P.w + (size_t)a * P.n, jk);
                    K.dw[a] = at32(P.dw + (size_t)a * P.n, jk);
                }
                if constexpr (NR == 0) K.cn = at32(P.cn, jk);
                T sc6, lam;
                pair_terms(K.x, K.y, K.z, K.t, k, K.b2, K.ib, K.rb, sc6, lam);
                K.sc6h = T(-0.5) * sc6;  // -sqrt(C6_ik) / 2: C9 / 2 below (pairwise C6 only)
                K.lam = F32 ? lam + T(L2K6_83) : T(K6) * lam;
                // The pair steps form c6'' = sum_b y'_b(j) kw'_b = (C6_ik / 4) C6_ij C6_jk from
                // j rows staged times C6_ij and k weights times C6_ik / 4, so that C9 / 2 =
                // -sqrt(c6'') and the log-derivative ratios (dC6_jk/dCN) / C6_jk are unchanged;
                // dw (used once per tile) takes the same factor.
                const T sk = T(0.25) * sc6 * sc6;
                for (int a = 0; a < NR; ++a) {
                    K.w[a] *= sk;
                    K.dw[a] *= sk;
                }
                // FP32: 6 x^16 = exp2(lam_jk - (8/3) log2 r_jk^2) * 6 exp2(lam_k)
                if constexpr (F32) K.fac = T(6) * T(__builtin_amdgcn_exp2f((float)K.lam));
            }
            for (int c = 0; c < 3; ++c) {
                K.lo[c] = uni((float)LB[8 * kt + 2 * c]);
                K.hi[c] = uni((float)LB[8 * kt + 2 * c + 1]);
            }
            K.Yk = ONE ? S.Jyd : S.Jyd + (__shfl(roff_l, K.t, AW) + (DR ? K.t : 0)) * AW;
            // k-side sums over all j tiles: lane-private LDS columns, folded to FP64 once per tile
            if constexpr (KL) {
                T *Khi = S.Khi + kl * 2 * A_NF * AW, *Klo = Khi + A_NF * AW;
                for (int q = 0; q < A_NF; ++q) Khi[q * AW + lane] = Klo[q * AW + lane] = T(0);
            }
        };
        // the bounding boxes of a k tile and of the j tile (lo, hi) are within rc3 of each other
        // (skip only when clearly apart: a 1e-3 relative margin covers float rounding)
        auto near = [&](const KTile &K, const float *jlo, const float *jhi) __attribute__((always_inline)) {
            float g2 = 0.f;
            for (int c = 0; c < 3; ++c) {
                const float g = fmaxf(0.f, fmaxf(jlo[c] - K.hi[c], K.lo[c] - jhi[c]));
                g2 += g * g;
            }
            return g2 <= (float)rc32 * 1.001f;
        };
        auto zero_k = [&](KTile &K) __attribute__((always_inline)) {
            K.A = K.Bx = K.By = K.Bz = K.H = K.G = T(0);
            for (int bb = 0; bb < NR; ++bb) K.Gb[bb] = T(0);
        };
        // fold the tile sums of neighbour k (exact two-sum in FP32, plain in FP64)
        auto fold_k = [&](KTile &K, int kt, int kl) __attribute__((always_inline)) {
            if constexpr (NR > 0)
                for (int bb = 0; bb < NR; ++bb) K.G += K.Gb[bb] * K.dw[bb];
            const T v[A_NF] = {K.A, K.Bx, K.By, K.Bz, K.H, K.G};
            const int k = kt * AW + lane;
            if constexpr (KL) {
                T *Khi = S.Khi + kl * 2 * A_NF * AW, *Klo = Khi + A_NF * AW;
                for (int q = 0; q < A_NF; ++q) {
                    const T hi = Khi[q * AW + lane], s = hi + v[q];
                    if constexpr (F32) {
                        const T bp = s - hi;
                        Klo[q * AW + lane] += (hi - (s - bp)) + (v[q] - bp);
                    }
                    Khi[q * AW + lane] = s;
                }
            } else if (k < nf) {
                for (int q = 0; q < A_NF; ++q) unsafeAtomicAdd(ptr32(La + q * cap, (unsigned)k), (double)v[q]);
            }
        };
        auto close_k = [&](int kt, int kl) __attribute__((always_inline)) {
            const int k = kt * AW + lane;
            if constexpr (KL)
                if (k < nf) {
                    const T *Khi = S.Khi + kl * 2 * A_NF * AW, *Klo = Khi + A_NF * AW;
                    for (int q = 0; q < A_NF; ++q)
                        unsafeAtomicAdd(ptr32(La + q * cap, (unsigned)k), (double)Khi[q * AW + lane] + (double)Klo[q * AW + lane]);
                }
        };

        // slot offsets: a j slot as a byte offset into any row of 64 slots, every row (the C6
        // rows of type(k) too) that offset plus a base
        constexpr int SB = (int)sizeof(T2<T>);
        // fixed rows by index in the static array (J0..J2 = rows 0-2, then J3, Jp and Jt as
        // AtmLds lays them out): the row offsets are immediates of the LDS instructions
        auto at = [&](int row, int jb) -> T2<T> & {
            return *reinterpret_cast<T2<T> *>(reinterpret_cast<unsigned char *>(fixed) + row * AW * SB + jb);
        };
        auto cat = [](const T2<T> *row, int jb) -> const T2<T> & {
            return *reinterpret_cast<const T2<T> *>(reinterpret_cast<const unsigned char *>(row) + jb);
        };
        // the lane's own slot offset, opaque to the optimiser: (l SB) xor (q SB) would otherwise
        // be rewritten as (l xor q) SB, two instructions per step instead of one
        int lsb = lane * SB;
        asm volatile("" : "+v"(lsb));

        // Cutoff test of the pair {j, k} at slot offset jb: v = r_j - r_k, c = r_jk^2.
        auto test = [&](const KTile &K, const T2<T> &j0, const T2<T> &j1, T &vx, T &vy, T &vz, T &c, bool &ok,
                        uint64_t &m) __attribute__((always_inline)) {
            vx = K.x - j0.x;
            vy = K.y - j0.y;
            vz = K.z - j1.x;
            if constexpr (F32) {
                // 1e-12 <= r_jk^2 <= rc3^2 as one unsigned compare of the bits of c = r_jk^2 - 1e-12
                // (a negative c has its sign bit set); the 1e-12 is far below the rounding error of
                // any r_jk^2 past the floor, so c stands for r_jk^2
                c = fmaf(vz, vz, fmaf(vy, vy, fmaf(vx, vx, -1e-12f)));
                ok = __builtin_bit_cast(unsigned, (float)c) <= rc32b;
                m = __builtin_amdgcn_ballot_w64(ok);
            } else {
                c = vx * vx + vy * vy + vz * vz;
                const bool okc = c <= rc32, okz = c >= T(1e-12);
                ok = okc && okz;
                // one ballot per compare: the ballot of the combined flag would be re-derived from
                // a VGPR copy of it
                m = __builtin_amdgcn_ballot_w64(okc) & __builtin_amdgcn_ballot_w64(okz);
            }
        };
        // The damping term of the pair's types (with lam_j) and the C6 rows y_b(j) of type(k) at
        // slot jb (with per-element references).
        auto rows = [&](const KTile &K, int jb, T &lamjk, T &sc6j, T *yx, T *yy) __attribute__((always_inline)) {
            sc6j = T(0);  // sqrt C6_ij (pairwise C6 only)
            if constexpr (DR) {
                lamjk = cat(K.Yk, jb).x;
                if constexpr (NR == 0) sc6j = at(3, jb).x;
            } else {
                const T2<T> j3 = at(3, jb);
                const T rt = S.Rt[*reinterpret_cast<const int *>(&at(JT, jb)) + K.t];
                lamjk = F32 ? j3.y + rt : j3.y * rt;
                sc6j = j3.x;
            }
            if constexpr (NR > 0)
#pragma unroll
                for (int bb = 0; bb < NR; ++bb) {
                    const T2<T> yd = cat(K.Yk + (DRO + bb) * AW, jb);
                    yx[bb] = yd.x;
                    yy[bb] = yd.y;
                }
        };
        // One pair {j, k}: its terms added to k's tile sums and to the j-side sums p0..p2 (FP32:
        // registers the caller reads from and writes back to the j slot; FP64: the slot itself,
        // by LDS atomics).
        auto body = [&](KTile &K, int jb, T a, const T2<T> &j2, T vx, T vy, T vz, T c, T lamjk, T sc6j, const T *yx,
                        const T *yy, T2<T> &p0, T2<T> &p1, T2<T> &p2) __attribute__((always_inline)) {
            const T b = K.b2;
            // C6 and dC6/dCN_j of the pair {j, k} from the j row for type(k) (the sums start from
            // their first product: an FMA onto +0 is not a multiply in IEEE arithmetic).  With
            // per-element references the rows and weights carry C6_ij and C6_ik / 4: c6 = (C6_ik
            // / 4) C6_ij C6_jk, and C9 / 2 = -sqrt(c6).
            T c6 = T(0), dk6 = T(0), dj6 = T(0);
            if constexpr (NR == 0) {
                const int jl = jb / SB;
                c6_pairwise(P.c6g, ne, P.nref, S.Jt[jl * S.TI + 1], K.t, S.Jyd[jl].x, K.cn, c6, dj6, dk6);
            } else {
#pragma unroll
                for (int bb = 0; bb < NR; ++bb) {
                    c6 = bb == 0 ? yx[0] * K.w[0] : c6 + yx[bb] * K.w[bb];
                    dj6 = bb == 0 ? yy[0] * K.w[0] : dj6 + yy[bb] * K.w[bb];
                }
            }
            const T rs = prsqrt(c6);
            const T rcq = prsqrt(c);
            const T ic = rcq * rcq;
            const T iR = (j2.y * K.rb) * rcq;  // (abc)^(-1/2)
            const T iR2 = iR * iR, iR3 = iR2 * iR;
            // Q = x y z, x = a + b - c, y = c - d, z = c + d with d = b - a, so yz = c^2 - d^2:
            // dQ/da = yz + 2 x d, dQ/db = yz - 2 x d, dQ/dc = 2 x c - yz
            const T d = b - a, x = (a + b) - c;
            const T yz = c * c - d * d;
            const T Q = x * yz;
            const T z = Q * iR2;  // Q / (abc): A = iR3 (1 + (3/8) z)
            const T x2 = x + x;
            T dQa = T(0), dQb = T(0), dQc = T(0);
            if constexpr (F32) {
                dQa = yz + x2 * d;
                dQb = yz - x2 * d;
                dQc = x2 * c - yz;
            }
            T f;
            if constexpr (F32) {
                // 6 x^16 = exp2(lam_jk - (8/3) log2 c) * 6 exp2(lam_k)
                const T e = __builtin_amdgcn_exp2f(lamjk - (8.0f / 3.0f) * __builtin_amdgcn_logf(c));
                f = prcp(T(1) + e * K.fac);
            } else {
                const T y = lamjk * K.lam * ic;
                f = prcp(T(1) + T(6) * (y * y * y * inv_cbrt(y)));
            }
            // rf3 = (C9 / 2) f iR3 / c6 (per-element references: C9 / 2 = -sqrt(c6) = -c6 rs);
            // every term below is a multiple of it
            const T rf3 = (NR == 0 ? (sc6j * K.sc6h) * rs : -rs) * f * iR3;
            const T Ap = T(1) + T(0.375) * z;  // A / iR3
            const T cf3 = c6 * rf3;            // (C9 / 2) f iR3
            const T hC = cf3 * Ap;             // e / 2
            const T hc6 = rf3 * Ap;            // e / (2 c6)
            // t' = (2/3) t_xy: the 3/2 of H2 is taken out of every t and restored where the
            // per-neighbour sums are read (phase 4): G2' = cfh (P4 + (32/3) A (1 - f)) / 1.5
            // = cf3 ((28/9 - (64/9) f) + z (1/6 - (8/3) f)), P4 = 4 P1, cfh = (C9 / 2) f
            const T G2 = cf3 * ((T(28.0 / 9.0) - T(64.0 / 9.0) * f) + z * (T(1.0 / 6.0) - T(8.0 / 3.0) * f));
            const T H2 = cf3 * iR2;  // cfh iR5
            T tij, tik, tjk;
            if constexpr (F32) {
                tij = G2 * j2.x + H2 * dQa;
                tik = G2 * K.ib + H2 * dQb;
                tjk = G2 * ic + H2 * dQc;
            } else {
                // H2 dQ/da = Hy + Hx d, H2 dQ/db = Hy - Hx d, H2 dQ/dc = Hx c - Hy: two products
                // shared by the three edges, 9 operations instead of 10 (FP64 -1.5 % on Cu; the
                // FP32 kernel, issue-bound on a different mix, ran 1.9 % slower with it)
                const T Hy = H2 * yz, Hx = H2 * x2;
                tij = Hx * d + (G2 * j2.x + Hy);
                tik = (G2 * K.ib + Hy) - Hx * d;
                tjk = Hx * c + (G2 * ic - Hy);
            }
            // Per-neighbour sums (F_p = (3/2) (B_p - A_p d_p), phase 4) with the {j, k} edge along
            // v = d_k - d_j: grad_j e = -t_ij d_j + t_jk v, grad_k e = -t_ik d_k - t_jk v.
            // Neighbour j: FP32 plain read-modify-write of its slot (distinct across lanes at each
            // step; LDS float atomics measured ~6x slower on gfx950); FP64: LDS atomics (no
            // per-step barrier, measured ~4% faster than the barrier form).
            if constexpr (F32) {
                p0.x += tij;
                p0.y += tjk * vx;
                p1.x += tjk * vy;
                p1.y += tjk * vz;
                p2.x += hC;
                p2.y += hc6 * dj6;
            } else {
                T2<T> &r0 = at(JP, jb), &r1 = at(JP + 1, jb), &r2 = at(JP + 2, jb);
                atomicAdd(&r0.x, tij);
                atomicAdd(&r0.y, tjk * vx);
                atomicAdd(&r1.x, tjk * vy);
                atomicAdd(&r1.y, tjk * vz);
                atomicAdd(&r2.x, hC);
                atomicAdd(&r2.y, hc6 * dj6);
            }
            // neighbour k: registers
            K.A += tik;
            K.Bx -= tjk * vx;
            K.By -= tjk * vy;
            K.Bz -= tjk * vz;
            K.H += hC;
            if constexpr (NR == 0) {
                K.G += hc6 * dk6;
            } else {
#pragma unroll
                for (int bb = 0; bb < NR; ++bb) K.Gb[bb] += hc6 * yx[bb];
            }
        };
        // The steps of one k tile against the staged j tile.  On the diagonal tile pair (the j tile
        // is the k tile) every unordered pair {j, k} comes up at two steps, q and 64 - q, of the
        // rotation that pairs lane l with slot (l + q) mod 64; the terms are symmetric in j and k,
        // so steps 1..31 and the lower half of the lanes at step 32 take each pair once (32 steps,
        // not 64).  The other tile pairs pair lane l with slot l xor q, one instruction per step
        // (a permutation of the slots at every step, and every (j, k) pair comes up once over the
        // 64 steps), and need no order test.
        auto steps1 = [&](KTile &K, auto diag) __attribute__((always_inline)) {
            constexpr bool DIAG = decltype(diag)::value;
            constexpr int Q0 = DIAG ? 1 : 0, Q1 = DIAG ? AW / 2 + 1 : AW;
            unsigned tc = 0;  // pairs within rc3 in this tile pair (a scalar register)
#pragma unroll 1
            for (int q = Q0; q < Q1; ++q) {
                // FP32: lane l updates a j slot that another lane updated one step earlier
                if constexpr (F32) wave_sync();
                const int jb = DIAG ? ((lane * SB + q * SB) & (AW * SB - 1)) : (lsb ^ (q * SB));
                __builtin_assume(jb >= 0 && jb < AW * SB);
                const T2<T> j0 = at(0, jb), j1 = at(1, jb);
                T vx, vy, vz, c;
                bool ok;
                uint64_t m;
                test(K, j0, j1, vx, vy, vz, c, ok, m);
                if constexpr (DIAG) {
                    if (q == AW / 2) {  // lanes l and l + 32 share the pair
                        m &= 0xffffffffull;
                        ok = ok && lane < AW / 2;
                    }
                }
                if (m == 0) continue;  // uniform
                tc += (unsigned)__popcll(m);
                if (!ok) continue;
                const T2<T> j2 = at(2, jb);
                T lamjk, sc6j, yx[NR > 0 ? NR : 1], yy[NR > 0 ? NR : 1];
                rows(K, jb, lamjk, sc6j, yx, yy);
                if constexpr (F32) {
                    T2<T> &r0 = at(JP, jb), &r1 = at(JP + 1, jb), &r2 = at(JP + 2, jb);
                    T2<T> p0 = r0, p1 = r1, p2 = r2;
                    body(K, jb, j1.y, j2, vx, vy, vz, c, lamjk, sc6j, yx, yy, p0, p1, p2);
                    r0 = p0;
                    r1 = p1;
                    r2 = p2;
                } else {
                    T2<T> p0, p1, p2;
                    body(K, jb, j1.y, j2, vx, vy, vz, c, lamjk, sc6j, yx, yy, p0, p1, p2);
                }
            }
            cnt += tc;
        };
        // Two k tiles against the staged j tile (off-diagonal for both): the j slot is read, and
        // its sums written back, once for two pairs.
        auto steps2 = [&](KTile &K1, KTile &K2) __attribute__((always_inline)) {
            unsigned tc = 0;
#pragma unroll 1
            for (int q = 0; q < AW; ++q) {
                if constexpr (F32) wave_sync();
                const int jb = lsb ^ (q * SB);
                __builtin_assume(jb >= 0 && jb < AW * SB);
                const T2<T> j0 = at(0, jb), j1 = at(1, jb);
                T vx1, vy1, vz1, c1, vx2, vy2, vz2, c2;
                bool ok1, ok2;
                uint64_t m1, m2;
                test(K1, j0, j1, vx1, vy1, vz1, c1, ok1, m1);
                test(K2, j0, j1, vx2, vy2, vz2, c2, ok2, m2);
                if ((m1 | m2) == 0) continue;  // uniform
                tc += (unsigned)(__popcll(m1) + __popcll(m2));
                const T2<T> j2 = at(2, jb);
                T lamjk, sc6j, yx[NR > 0 ? NR : 1], yy[NR > 0 ? NR : 1];
                // one element type: the rows do not depend on k
                if constexpr (ONE) rows(K1, jb, lamjk, sc6j, yx, yy);
                T2<T> p0, p1, p2;
                if constexpr (F32) {
                    p0 = at(JP, jb);
                    p1 = at(JP + 1, jb);
                    p2 = at(JP + 2, jb);
                }
                if (ok1) {
                    if constexpr (!ONE) rows(K1, jb, lamjk, sc6j, yx, yy);
                    body(K1, jb, j1.y, j2, vx1, vy1, vz1, c1, lamjk, sc6j, yx, yy, p0, p1, p2);
                }
                if (ok2) {
                    if constexpr (!ONE) rows(K2, jb, lamjk, sc6j, yx, yy);
                    body(K2, jb, j1.y, j2, vx2, vy2, vz2, c2, lamjk, sc6j, yx, yy, p0, p1, p2);
                }
                if constexpr (F32) {
                    at(JP, jb) = p0;
                    at(JP + 1, jb) = p1;
                    at(JP + 2, jb) = p2;
                }
            }
            cnt += tc;
        };

        // FP32: k tiles two at a time (K1 = tile kb, K2 = tile kb + 1 when this work item has it):
        // each j tile is staged once for both, and on the tiles below kb both run in the same
        // steps.  FP64: one at a time.
        constexpr int KT = AtmLds<T>::KTILES;
        for (int kb = kb_lo; kb < kb_hi; kb += KT) {
            const bool two = KT == 2 && kb + 1 < kb_hi;
            KTile K1, K2;
            load_k(K1, kb, 0);
            if (two) load_k(K2, kb + 1, 1);
            const int jt_end = two ? kb + 1 : kb;
            for (int jt = 0; jt <= jt_end; ++jt) {
                const int pj = jt * AW + lane;
                const bool jval = pj < nf;
                // bounding box of the j tile (phase 2)
                float jlo[3], jhi[3];
                for (int c = 0; c < 3; ++c) {
                    jlo[c] = uni((float)LB[8 * jt + 2 * c]);
                    jhi[c] = uni((float)LB[8 * jt + 2 * c + 1]);
                }
                const bool u1 = jt <= kb && near(K1, jlo, jhi);
                const bool u2 = two && near(K2, jlo, jhi);
                if (!u1 && !u2) continue;

                // stage the j tile (padding lanes: far coordinates, zero rows)
                wave_sync();
                {
                    T jxs = T(FAR), jys = T(FAR), jzs = T(FAR);
                    T a = T(1), ia = T(1), ra = T(1), sc6 = T(0), lam = T(0);
                    int tj = 0;
                    T wj[NR > 0 ? NR : 1], dwj[NR > 0 ? NR : 1];
                    T jcn = T(0);  // pairwise C6 only
                    for (int q = 0; q < NR; ++q) wj[q] = dwj[q] = T(0);
                    if (jval) {
                        jxs = at32(LT + F_DX * cap, (unsigned)pj);
                        jys = at32(LT + F_DY * cap, (unsigned)pj);
                        jzs = at32(LT + F_DZ * cap, (unsigned)pj);
                        const unsigned jj = (unsigned)at32(Lj, (unsigned)pj);
                        tj = ONE ? 0 : at32(Lt, (unsigned)pj);
                        for (int q = 0; q < NR; ++q) {
                            wj[q] = at32(P.w + (size_t)q * P.n, jj);
                            dwj[q] = at32(P.dw + (size_t)q * P.n, jj);
                        }
                        if constexpr (NR == 0) jcn = at32(P.cn, jj);
                        pair_terms(jxs, jys, jzs, tj, pj, a, ia, ra, sc6, lam);
                    }
                    S.J0[lane] = T2<T>{jxs, jys};
                    S.J1[lane] = T2<T>{jzs, a};
                    S.J2[lane] = T2<T>{ia, ra};
                    if constexpr (J3) S.J3[lane] = T2<T>{sc6, lam};
                    if constexpr (!DR) S.Jt[lane * S.TI] = tj * ne;  // row of the per-type-pair table
                    if constexpr (NR == 0) {
                        // pairwise C6: the j type and CN, one row
                        S.Jt[lane * S.TI + 1] = tj;
                        S.Jyd[lane] = T2<T>{jcn, T(0)};
                    }
                    // Types in ascending order, NR rows each: the rows past a type's own count
                    // (zero: its C6 entries there are 0) land on the next type's rows, which that
                    // type then writes over; the last type's fill the padding.  Every row a pair
                    // step can read is thus rewritten at each staging.
                    for (int B = 0; B < ne; ++B) {
                        const int base = __builtin_amdgcn_readlane(roff_l, B) + (DR ? B : 0);
                        if constexpr (DR) {
                            const T rt = rt_of(tj * ne + B);
                            T lr;
                            {
                                // rt rounded on its own, as the per-type-pair table of the variant
                                // without DR holds it (a fused lam + (8/3) log2 R0^2 would differ
                                // in the last bit, and with it every result of that variant)
#pragma clang fp contract(off)
                                lr = F32 ? lam + rt : lam * rt;
                            }
                            S.Jyd[base * AW + lane] = T2<T>{lr, T(0)};
                        }
                        // rows times C6_ij = sc6^2 (see the k weights)
                        const T c6ij = sc6 * sc6;
#pragma unroll
                        for (int bb = 0; bb < NR; ++bb) {
                            T sy = T(0), sdy = T(0);
                            for (int q = 0; q < NR; ++q) {
                                const T c = P.c6[((tj * ne + B) * NR + q) * NR + bb];
                                sy += c * wj[q];
                                sdy += c * dwj[q];
                            }
                            S.Jyd[(base + DRO + bb) * AW + lane] = T2<T>{sy * c6ij, sdy * c6ij};
                        }
                    }
                    for (int q = 0; q < 3; ++q) S.Jp[q * AW + lane] = T2<T>{T(0), T(0)};
                }
                wave_sync();

                if (jt < kb) {
                    if (u1 && u2) {
                        zero_k(K1);
                        zero_k(K2);
                        steps2(K1, K2);
                        fold_k(K1, kb, 0);
                        fold_k(K2, kb + 1, 1);
                    } else if (u1) {
                        zero_k(K1);
                        steps1(K1, std::false_type{});
                        fold_k(K1, kb, 0);
                    } else {
                        zero_k(K2);
                        steps1(K2, std::false_type{});
                        fold_k(K2, kb + 1, 1);
                    }
                } else if (jt == kb) {
                    if (u1) {
                        zero_k(K1);
                        steps1(K1, std::true_type{});
                        fold_k(K1, kb, 0);
                    }
                    if (u2) {
                        zero_k(K2);
                        steps1(K2, std::false_type{});
                        fold_k(K2, kb + 1, 1);
                    }
                } else {  // jt == kb + 1: K2's own tile
                    zero_k(K2);
                    steps1(K2, std::true_type{});
                    fold_k(K2, kb + 1, 1);
                }
                wave_sync();
                if (jval)
                    for (int q = 0; q < 3; ++q) {
                        const T2<T> pq = S.Jp[q * AW + lane];
                        unsafeAtomicAdd(ptr32(La + (2 * q) * cap, (unsigned)pj), (double)pq.x);
                        unsafeAtomicAdd(ptr32(La + (2 * q + 1) * cap, (unsigned)pj), (double)pq.y);
                    }
            }
            close_k(kb, 0);
            if (two) close_k(kb + 1, 1);
        }

        // ---- phase 4.  Per neighbour slot p: F_p = grad_p E = (3/2) (B_p - d_p A_p), H_p = sum of e/2,
        // Gc_p = sum of (e/2) dlnC6_pq/dCN_p over the triangles it belongs to.  Then
        //   dE/dCN_p = Gc_p + (dlnC6_ip/dCN_p) H_p,  E/3 share of p = (2/3) H_p,
        //   centre: E = sum H_p (each triangle feeds two slots), dE/dCN_i = sum (dlnC6_ip/dCN_i) H_p,
        //   F_i = -sum F_p (translation invariance), virial = sum d_p (x) F_p
        //   (= -sum over edges t r r^T, since sum_v r_v (x) grad_v e = sum_edges t r r^T).
        Acc<double> aE, aG, aF[3], aW[9];
        aE.init();
        aG.init();
        for (int q = 0; q < 3; ++q) aF[q].init();
        for (int q = 0; q < 9; ++q) aW[q].init();
        for (int p = lane; p < np; p += AW) {
            const int j = Lj[p];
            const double dx = (double)LT[F_DX * cap + p], dy = (double)LT[F_DY * cap + p],
                         dz = (double)LT[F_DZ * cap + p];
            double acc[A_NF];
            for (int q = 0; q < A_NF; ++q)
                acc[q] = __hip_atomic_load(&La[q * cap + p], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            // the pair steps sum t' = (2/3) t
            const double Ap = acc[A_A];
            const double fx = 1.5 * (acc[A_BX] - dx * Ap), fy = 1.5 * (acc[A_BY] - dy * Ap),
                         fz = 1.5 * (acc[A_BZ] - dz * Ap);
            const double H = acc[A_H], Gc = acc[A_G];
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
        if (lane == 0) {
            atomicAdd(&P.fa[3 * i + 0], F[0]);
            atomicAdd(&P.fa[3 * i + 1], F[1]);
            atomicAdd(&P.fa[3 * i + 2], F[2]);
            atomicAdd(&P.ga[i], G);
            atomicAdd(&P.eas[i], E * (1.0 / 3.0));
            const double iv = C.S->inv_vol;
            if (nch == 1) {
                P.eat[i] = E;
                for (int q = 0; q < 6; ++q) P.vir[6 * i + q] += W[q] * iv;
                P.npair[3 * i + 2] = (float)cnt;
            } else {
                P.atm_eatc[(size_t)chunk * P.n + i] = E;
                for (int q = 0; q < 6; ++q) atomicAdd(&P.vir[6 * i + q], W[q] * iv);
                atomicAdd(&P.npair[3 * i + 2], (float)cnt);  // whole numbers below 2^24: exact
            }
        }
    }
}

// Forward-neighbour list of each centre (one wave per centre): the three-body stencil walk of
// phase 1 with the list cutoff, keeping (j, stencil entry) of every forward neighbour within it in
// walk order.  With a skin the list stays valid while no atom moves more than half the skin.
template <class T>
__global__ __launch_bounds__(256) void k_atm_fwd(Params<T> P) {
    const int lane = threadIdx.x & (AW - 1);
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= P.n) return;  // whole wave
    Centre<T> C;
    C.load(P, i);
    const int cap = P.atm_fl_cap;
    FwdEntry *FL = P.atm_fl + (size_t)i * cap;
    const T rcf = P.rc_fl2;
    int count = 0;
    const int s0 = fwd_first(C.S), s1 = C.S->st_off[CUT_ATM] + C.S->st_cnt[CUT_ATM];
    const T rcfl = xsqrt(rcf);
    for (int sb = s0; sb < s1; sb += AW) {
        const auto Bn = C.bins(P, sb, s1, rcfl);
        const int nsb = s1 - sb < AW ? s1 - sb : AW;
        for (int sk = 0; sk < nsb; ++sk) {
            int j0, j1;
            T ox, oy, oz;
            Bn.get(sk, j0, j1, ox, oy, oz);
            if (sb + sk == s0) j0 = i + 1;  // own bin, no image shift
            for (int jb = j0; jb < j1; jb += AW) {
                const int j = jb + lane;
                bool in = false;
                if (j < j1) {
                    const T dx = ox - P.lx[j], dy = oy - P.ly[j], dz = oz - P.lz[j];
                    in = dx * dx + dy * dy + dz * dz <= rcf;
                }
                const uint64_t m = ballot64(in);
                if (m == 0) continue;
                if (in) {
                    const int pos = count + mbcnt64(m);
                    if (pos < cap) FL[pos] = FwdEntry{j, sb + sk};
                }
                count += __popcll(m);
            }
        }
    }
    if (lane == 0) {
        P.atm_fl_n[i] = count;
        if (count > cap) atomicOr(P.status, 2u);  // the host grows the capacity and runs again
    }
}

// Zeroes fa | ga | eas | eat (6 N doubles) and the centre counter.
__global__ void k_fill_d(double *p, size_t n, int *counter) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = 0.0;
    if (t == 0) *counter = 0;
}

// Per-centre three-body energy from the work items' shares, in chunk order (deterministic).
template <class T>
__global__ void k_eat_chunks(Params<T> P) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.n) return;
    double e = 0.0;
    for (int c = 0; c < P.atm_chunks; ++c) e += P.atm_eatc[(size_t)c * P.n + i];
    P.eat[i] = e;
}

// P.g += ga, P.f += fa (after the ATM kernel, before the chain rule).
template <class T>
__global__ void k_merge(Params<T> P) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P.n) return;
    P.g[p] += P.ga[p];
    for (int q = 0; q < 3; ++q) P.f[3 * p + q] += P.fa[3 * p + q];
}

// One instantiation per reference count (1..MAX_NREF): the C6 dot products and the LDS rows
// are sized by the references the handle's elements actually have.  var: bit 0 = KL, bit 1 = DR;
// one: the handle has a single element type.  nref 0: the pairwise-C6 form for tables that do
// not factorise (without DR).
template <class T, bool KL, bool DR, bool ONE, int... NR>
const void *atm_pick(int nref, std::integer_sequence<int, NR...>) {
    const void *fn = nullptr;
    ((fn = (nref == NR + 1) ? reinterpret_cast<const void *>(&k_atm<T, NR + 1, KL, DR, ONE>) : fn), ...);
    return fn;
}

template <class T, bool ONE>
const void *atm_kernel_of(int nref, int var) {
    using S = std::make_integer_sequence<int, MAX_NREF>;
    switch (var & 3) {
        case 0: return atm_pick<T, false, false, ONE>(nref, S{});
        case 1: return atm_pick<T, true, false, ONE>(nref, S{});
        case 2: return atm_pick<T, false, true, ONE>(nref, S{});
        default: return atm_pick<T, true, true, ONE>(nref, S{});
    }
}

template <class T>
const void *atm_kernel(int ne, int nref, int var) {
    if (nref == 0)
        return (var & ATM_DR) ? nullptr
                              : (var & ATM_KL) ? reinterpret_cast<const void *>(&k_atm<T, 0, true, false, false>)
                                               : reinterpret_cast<const void *>(&k_atm<T, 0, false, false, false>);
    return ne == 1 ? atm_kernel_of<T, true>(nref, var) : atm_kernel_of<T, false>(nref, var);
}

// dynamic LDS of a launch (the fixed rows are the kernel's static LDS)
size_t atm_dyn_lds_bytes(int ne, int nref, int yrows, size_t ts, int var) {
    const bool kl = var & ATM_KL, dr = var & ATM_DR;
    const bool one = ne == 1 && nref > 0;  // atm_kernel's choice
    return ts == sizeof(double) ? AtmLds<double>::dyn_bytes(ne, nref, yrows, kl, dr, one)
                                : AtmLds<float>::dyn_bytes(ne, nref, yrows, kl, dr, one);
}

int blocks_per_cu(int ne, int nref, int yrows, size_t ts, int var) {
    const size_t lds = atm_dyn_lds_bytes(ne, nref, yrows, ts, var);
    const void *fn = ts == sizeof(double) ? atm_kernel<double>(ne, nref, var) : atm_kernel<float>(ne, nref, var);
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, AW, lds) != hipSuccess) per_cu = 0;
    return per_cu;
}

}  // namespace

size_t atm_lds_bytes(int ne, int nref, int yrows, size_t ts, int var) {
    const bool kl = var & ATM_KL, dr = var & ATM_DR;
    return ts == sizeof(double) ? AtmLds<double>::bytes(ne, nref, yrows, kl, dr)
                                : AtmLds<float>::bytes(ne, nref, yrows, kl, dr);
}

int atm_choose_variant(int ne, int nref, int yrows, size_t ts) {
    // The first of KL + DR, DR, KL, neither (most work saved first: DR per pair, KL per tile)
    // that keeps at least 90 % of the most resident blocks any of them reaches; at 10 blocks per
    // CU or fewer (2.5 waves per SIMD) the first that keeps all of them.  Measured: fcc Cu (FP32,
    // 3 references) runs KL + DR at 18 blocks per CU as fast as DR at 20, with a third less HBM
    // traffic; the six-element batch runs neither at 10 blocks 3 % faster than DR at 9 and 22 %
    // faster than KL at 7.
    const int order[4] = {ATM_KL | ATM_DR, ATM_DR, ATM_KL, 0};
    int blocks[4], most = 0;
    for (int k = 0; k < 4; ++k) {
        blocks[k] = (nref == 0 && (order[k] & ATM_DR)) || atm_lds_bytes(ne, nref, yrows, ts, order[k]) > 65536
                        ? 0
                        : blocks_per_cu(ne, nref, yrows, ts, order[k]);
        most = blocks[k] > most ? blocks[k] : most;
    }
    const bool low = most <= 10;
    for (int k = 0; k < 4; ++k)
        if (blocks[k] > 0 && (low ? blocks[k] == most : 10 * blocks[k] >= 9 * most)) return order[k];
    return 0;
}

int atm_resident_blocks(int ne, int nref, int yrows, size_t ts, int var) {
    const int per_cu = blocks_per_cu(ne, nref, yrows, ts, var);
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        cus = 0;
    return per_cu > 0 && cus > 0 ? per_cu * cus : 4096;
}

template <class T>
void launch_atm(const Params<T> &P, int nblocks, hipStream_t s) {
    // fa (3N), ga, eas, eat (N each) are contiguous: zero them (and the centre counter) in one pass
    const size_t nz = (size_t)6 * P.n;
    if (nz) hipLaunchKernelGGL(k_fill_d, dim3((unsigned)((nz + 255) / 256)), dim3(256), 0, s, P.fa, nz, P.atm_next);
    if (P.n == 0 || !(P.rc_32 > T(0))) return;
    const int knr = P.c6_pairwise ? 0 : P.nref;
    const size_t lds = atm_dyn_lds_bytes(P.ne, knr, P.atm_yrows, sizeof(T), P.atm_var);
    const void *fn = atm_kernel<T>(P.ne, knr, P.atm_var);
    if (!fn) return;  // not reached: 1..MAX_NREF references, or pairwise C6 without DR
    Params<T> arg = P;
    void *args[] = {&arg};
    (void)hipLaunchKernel(fn, dim3(nblocks), dim3(AW), args, lds, s);
    if (P.atm_chunks > 1) hipLaunchKernelGGL(k_eat_chunks<T>, dim3((P.n + 255) / 256), dim3(256), 0, s, P);
    hipLaunchKernelGGL(k_merge<T>, dim3((P.n + 255) / 256), dim3(256), 0, s, P);
}

template <class T>
void launch_atm_lists(const Params<T> &P, hipStream_t s) {
    if (P.n == 0 || !P.atm_fl || !(P.rc_32 > T(0))) return;
    hipLaunchKernelGGL(k_atm_fwd<T>, dim3((P.n + 3) / 4), dim3(256), 0, s, P);
}
template void launch_atm_lists<double>(const Params<double> &, hipStream_t);
template void launch_atm_lists<float>(const Params<float> &, hipStream_t);

template void launch_atm<double>(const Params<double> &, int, hipStream_t);
template void launch_atm<float>(const Params<float> &, int, hipStream_t);

template <class T>
void launch_atm_merge(const Params<T> &P, hipStream_t s) {
    if (P.n) hipLaunchKernelGGL(k_merge<T>, dim3((P.n + 255) / 256), dim3(256), 0, s, P);
}
template void launch_atm_merge<double>(const Params<double> &, hipStream_t);
template void launch_atm_merge<float>(const Params<float> &, hipStream_t);

}  // namespace d3
