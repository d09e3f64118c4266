// Device-side helpers shared by the HIP kernel files: wave64 primitives, math overloads for
// float/double, the compensated accumulator and the centre-atom geometry of a stencil scan.
#pragma once
#include <hip/hip_runtime.h>

#include <cfloat>

#include "d3_engine.hpp"

namespace d3 {
namespace {

constexpr int WAVE = 64;
constexpr int BLOCK = 256;
constexpr int WPB = BLOCK / WAVE;
constexpr int QCAP = 2 * WAVE;  // per-wave compaction queue (entries)

__device__ __forceinline__ int lane_id() { return threadIdx.x & (WAVE - 1); }
__device__ __forceinline__ uint64_t ballot64(bool p) { return __builtin_amdgcn_ballot_w64(p); }
__device__ __forceinline__ int mbcnt64(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}
// Orders LDS traffic between the lanes of one wave (the queue is private to the wave).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
// Lane k's value of v, as a wave-uniform value (k uniform).
__device__ __forceinline__ float readlane_t(float v, int k) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), k));
}
__device__ __forceinline__ double readlane_t(double v, int k) {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, k);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), k);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// Element i of a per-atom array through a 32-bit byte offset: the load takes the array base in
// scalar registers and one 32-bit offset register instead of 64-bit address arithmetic per lane.
// Handles hold at most 2^28 atoms (max_length), so every offset of an array of 8-byte elements
// indexed by atom fits.
template <class U>
__device__ __forceinline__ U at32(const U *base, unsigned i) {
    return *reinterpret_cast<const U *>(reinterpret_cast<const char *>(base) + i * (unsigned)sizeof(U));
}

// Address of element i of such an array, for stores and atomics (the same 32-bit offset form).
template <class U>
__device__ __forceinline__ U *ptr32(U *base, unsigned i) {
    return reinterpret_cast<U *>(reinterpret_cast<char *>(base) + i * (unsigned)sizeof(U));
}
// A wave-uniform pointer made opaque to the optimiser (it stays in SGPRs as computed), so that
// at32(p, j) keeps the scalar-base + 32-bit-offset form instead of being re-associated into 64-bit
// vector address arithmetic when p is itself a sum (a row of a row-major array).
template <class U>
__device__ __forceinline__ const U *sbase(const U *p) {
    auto g = (const __attribute__((address_space(1))) U *)p;  // global memory
    asm volatile("" : "+s"(g));
    return (const U *)g;
}

__device__ __forceinline__ int wrapi(int x, int n) {
    int r = x % n;
    return r < 0 ? r + n : r;
}

__device__ __forceinline__ double xsqrt(double x) { return sqrt(x); }
__device__ __forceinline__ float xsqrt(float x) { return sqrtf(x); }
__device__ __forceinline__ double xexp(double x) { return exp(x); }
__device__ __forceinline__ float xexp(float x) { return expf(x); }
__device__ __forceinline__ double xcbrt(double x) { return cbrt(x); }
__device__ __forceinline__ float xcbrt(float x) { return cbrtf(x); }
__device__ __forceinline__ double xrsqrt(double x) { return rsqrt(x); }
__device__ __forceinline__ float xrsqrt(float x) { return rsqrtf(x); }
__device__ __forceinline__ double xtanh(double x) { return tanh(x); }
__device__ __forceinline__ float xtanh(float x) { return tanhf(x); }

// Smooth CN cutoff, src/d3_kernel.cu:550-561 (SMOOTH_CUTOFF_CN): the CN term c times
// s = tanh^3(rc - r); c'/r (dcn) becomes (c s)'/r = (c'/r) s - 3 tanh^2 (1 - tanh^2) c / r.
template <class T>
__device__ __forceinline__ void smooth_cn_term(T rc, T r, T ir, T &c, T &dcn) {
    const T th = xtanh(rc - r), th2 = th * th;
    dcn = dcn * th2 * th - T(3) * th2 * (T(1) - th2) * c * ir;
    c *= th2 * th;
}
template <class T>
__device__ __forceinline__ T finite_or_zero(T x) {
    return isfinite(x) ? x : T(0);
}

// C6 of one pair by the Gaussian-weighted interpolation over the pair's own reference entries,
// src/d3_kernel.cu:709-800 (calculate_c6ab_dual), for parameter tables whose CN references do
// not factorise per element.  g: the handle's pair table, entry (A, B, a, b) = {C6ref, CN_A ref,
// CN_B ref} at 3 (((A ne + B) nr + a) nr + b); an entry counts iff both CN refs > -1.  Exponents
// are shifted by their maximum (C6 = Z / W is unchanged; the shift keeps W from underflowing far
// from every reference), and the C6 references by that entry's c0: C6 = c0 + Z'/W with Z' =
// sum (C6ref - c0) L, dC6/dCN = (Z'_1 W - Z' W_1) / W^2 — the same values, but when one entry
// carries nearly all the weight (a CN far from the others) the derivative is no longer the
// difference of two nearly equal products, which in FP32 left noise of ~1e-7 C6 per unit CN.
// Returns C6, dC6/dCN_A, dC6/dCN_B; non-finite results become 0.
template <class T>
__device__ void c6_pairwise(const T *g, int ne, int nr, int A, int B, T cna, T cnb, T &c6, T &d1, T &d2) {
    const T *e = g + (size_t)3 * ((A * ne + B) * nr * nr);
    const int m = nr * nr;
    T mx = T(-1e30), c0 = T(0);
    for (int q = 0; q < m; ++q) {
        const T r1 = e[3 * q + 1], r2 = e[3 * q + 2];
        if (r1 > T(-1) && r2 > T(-1)) {
            const T u = cna - r1, v = cnb - r2;
            const T x = T(-K3) * (u * u + v * v);
            if (x > mx) {
                mx = x;
                c0 = e[3 * q];
            }
        }
    }
    T Z = T(0), W = T(0), cz1 = T(0), cz2 = T(0), dw1 = T(0), dw2 = T(0);
    for (int q = 0; q < m; ++q) {
        const T cref = e[3 * q] - c0, r1 = e[3 * q + 1], r2 = e[3 * q + 2];
        if (r1 > T(-1) && r2 > T(-1)) {
            const T u = cna - r1, v = cnb - r2;
            const T L = xexp(T(-K3) * (u * u + v * v) - mx);
            const T l1 = T(-2 * K3) * u * L, l2 = T(-2 * K3) * v * L;
            Z += cref * L;
            W += L;
            cz1 += cref * l1;
            cz2 += cref * l2;
            dw1 += l1;
            dw2 += l2;
        }
    }
    const T W2 = W * W;
    d1 = finite_or_zero(W2 > T(0) ? (cz1 * W - Z * dw1) / W2 : T(0));
    d2 = finite_or_zero(W2 > T(0) ? (cz2 * W - Z * dw2) / W2 : T(0));
    c6 = finite_or_zero(W > T(0) ? c0 + Z / W : T(0));
}

// Per-lane accumulator: plain sum in FP64; two-sum compensated (hi, lo) in FP32.
template <class T>
struct Acc {
    T s;
    __device__ void init() { s = T(0); }
    __device__ void add(T x) { s += x; }
};
template <>
struct Acc<float> {
    float hi, lo;
    __device__ void init() { hi = lo = 0.0f; }
    __device__ void add(float x) {
        float s = hi + x;
        float bp = s - hi;
        lo += (hi - (s - bp)) + (x - bp);
        hi = s;
    }
};

// Wave-wide total of an accumulator, returned in double in every lane (xor butterfly over the
// 64 lanes; the FP32 variant keeps the exact two-sum error of every step).
__device__ __forceinline__ double wave_total(const Acc<double> &a) {
    double v = a.s;
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
    return v;
}
__device__ __forceinline__ double wave_total(const Acc<float> &a) {
    float hi = a.hi, lo = a.lo;
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) {
        float oh = __shfl_xor(hi, o, WAVE), ol = __shfl_xor(lo, o, WAVE);
        float s = hi + oh;
        float bp = s - hi;
        float err = (hi - (s - bp)) + (oh - bp);
        hi = s;
        lo = lo + ol + err;
    }
    return (double)hi + (double)lo;
}
// Value of one lane's accumulator.
__device__ __forceinline__ double acc_value(const Acc<double> &a) { return a.s; }
__device__ __forceinline__ double acc_value(const Acc<float> &a) { return (double)a.hi + (double)a.lo; }

__device__ __forceinline__ float wave_sumf(float v) {
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
    return v;
}

// Geometry of one centre atom: its cell, bin coordinates and local position.
template <class T>
struct Centre {
    int i, sid, ti, bx, by, bz, nx, ny, nz, bin_base;
    T x, y, z;
    const SysDev *S;
    __device__ void load(const Params<T> &P, int i_) {
        i = i_;
        sid = P.sys_s[i];
        S = &P.sys[sid];
        nx = S->nb[0];
        ny = S->nb[1];
        nz = S->nb[2];
        bin_base = S->bin_base;
        const int b = P.bin_s[i] - bin_base;
        bx = b % nx;
        by = (b / nx) % ny;
        bz = b / (nx * ny);
        x = P.lx[i];
        y = P.ly[i];
        z = P.lz[i];
        ti = P.type_s[i];
    }
    // Atom range of the bin at stencil entry s, and the centre's position plus the image vector.
    __device__ void bin(const Params<T> &P, int s, int &j0, int &j1, T &ox, T &oy, T &oz) const {
        const StEntry d = P.st_d[s];
        const int cx = wrapi(bx + d.dx, nx), cy = wrapi(by + d.dy, ny), cz = wrapi(bz + d.dz, nz);
        const int nb = bin_base + cx + nx * (cy + ny * cz);
        j0 = P.bin_start[nb];
        j1 = P.bin_start[nb + 1];
        ox = x + P.st_D[4 * s + 0];
        oy = y + P.st_D[4 * s + 1];
        oz = z + P.st_D[4 * s + 2];
    }
    // Up to 64 stencil entries at once: lane l holds entry s + l, so the wave issues one batch of
    // vector loads (and the wrap's integer divisions) for 64 entries instead of a chain of
    // dependent scalar loads per entry; get(k) reads entry s + k back as wave-uniform values.
    struct Bins {
        int j0, j1;
        T ox, oy, oz;
        __device__ void get(int k, int &a, int &b, T &px, T &py, T &pz) const {
            a = __builtin_amdgcn_readlane(j0, k);
            b = __builtin_amdgcn_readlane(j1, k);
            px = readlane_t(ox, k);
            py = readlane_t(oy, k);
            pz = readlane_t(oz, k);
        }
    };
    // rc: the cutoff the caller tests pairs against.  A stencil entry whose bin no atom within rc
    // of this centre can occupy comes back empty (j1 = j0): the stencil holds every bin that some
    // point of the centre's bin can reach, this test the bins the centre atom itself can.  Lower
    // bound of |r_ij| for j in the bin at image offset O = r_i + D from the bin centre: |O . n_k|
    // - hh_k along each lattice-plane normal n_k (in quadrature for an orthorhombic cell, the
    // largest otherwise), and |O| - hdiag; pad covers atoms that moved out of their bin within
    // the skin, and rounding.
    __device__ Bins bins(const Params<T> &P, int s, int s_end, T rc) const {
        Bins r{0, 0, T(0), T(0), T(0)};
        const int t = s + lane_id();
        if (t < s_end) {
            bin(P, t, r.j0, r.j1, r.ox, r.oy, r.oz);
            const double ox = (double)r.ox, oy = (double)r.oy, oz = (double)r.oz;
            double q2 = 0.0, qm = 0.0;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const double p = fabs(ox * S->nrm[3 * k] + oy * S->nrm[3 * k + 1] + oz * S->nrm[3 * k + 2]) - S->hh[k];
                const double g = p > 0.0 ? p : 0.0;
                q2 += g * g;
                qm = fmax(qm, g);
            }
            const double sph = sqrt(ox * ox + oy * oy + oz * oz) - S->hdiag;
            const double lb = S->ortho ? sqrt(q2) : fmax(qm, sph);
            if (lb > (double)rc + S->pad) r.j1 = r.j0;
        }
        return r;
    }
};

}  // namespace
}  // namespace d3
