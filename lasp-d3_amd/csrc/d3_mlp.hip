// Per-atom network of the learned three-body stand-in on MFMA (bf16 operands, fp32 accumulate):
// weight packing, inference (energies and input gradients), and training (energy-loss gradient,
// per-row backward, weight gradients as split-K MFMA contractions; force-loss gradient through a
// tangent pass; Adam).  Layout constants in d3_sur.hpp.
//
// One wave owns a tile of 16 rows (atoms).  v_mfma_f32_16x16x16_bf16 operand layout (lane l):
//   A (16 x 16, M x K): A[l % 16][4 (l / 16) + q],  B (K x N): B[4 (l / 16) + q][l % 16],
//   C/D (M x N):        D[4 (l / 16) + q][l % 16],  q = 0..3.
// A layer's output arrives in the D layout (lane: rows 4 (l/16) + q of one column); the next layer
// needs it in the A layout (lane: one row, 4 consecutive columns), so activations pass through a
// per-wave LDS tile [16][136] (bf16, rows padded to spread banks).  The D layout of h1 / dz2 / dz1
// is at the same time the A / B fragment layout of the weight-gradient products H1^T dZ2 and
// X'^T dZ1 (K = rows), which is how training stores them.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <vector>

#include "d3_sur.hpp"

// Fragment scratch of the training gradient for callers that cannot allocate
// (d3x_sur_train_grad_noalloc): STATIC_TRAIN_ROWS rows, 16 rows per tile of 24 x 64 fragments.
// One copy per device, held until the process ends; it keeps the last call's contents, which
// need no clearing (k_train writes every fragment of its tiles before k_wgrad reads them).  Host
// code never uses this name as an address: static_train_scratch() gives the device's address.
static __device__ __align__(256) unsigned char scratch[(size_t)d3::sur::STATIC_TRAIN_ROWS * d3::sur::TRAIN_ROW_BYTES];

namespace d3 {
namespace sur {
namespace {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int TP = 136;  // LDS tile row pitch (bf16)

__device__ __forceinline__ unsigned short bf16(float f) {  // round to nearest even
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float silu(float z) { return z / (1.0f + __expf(-z)); }
__device__ __forceinline__ float dsilu(float z) {
    const float s = 1.0f / (1.0f + __expf(-z));
    return s * (1.0f + z * (1.0f - s));
}
__device__ __forceinline__ float d2silu(float z) {
    const float s = 1.0f / (1.0f + __expf(-z));
    return s * (1.0f - s) * (2.0f + z * (1.0f - 2.0f * s));
}
__device__ __forceinline__ v4f mfma(v4s a, v4s b, v4f c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ v4s pack4(float a, float b, float c, float d) {
    v4s r;
    r[0] = (short)bf16(a);
    r[1] = (short)bf16(b);
    r[2] = (short)bf16(c);
    r[3] = (short)bf16(d);
    return r;
}
// Orders the per-wave LDS tile between lanes.
__device__ __forceinline__ void wsync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ------------------------------------------------------------------------------- packing

__global__ void k_pack(const float *__restrict__ P, v4s *__restrict__ PK) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= NCLASS * PK_CLASS) return;
    const int c = t / PK_CLASS, u = t % PK_CLASS;
    const float *W1 = P + HEAD + c * CLASS_SIZE + O_W1, *W2 = P + HEAD + c * CLASS_SIZE + O_W2;
    const int l = u % 64, lr = l % 16, lg = l / 16;
    float v[4];
    if (u < PK_W2) {  // W1 [64][128] as B: (kk, nb) of 4 x 8
        const int kk = u / (8 * 64), nb = (u / 64) % 8;
        for (int q = 0; q < 4; ++q) v[q] = W1[(16 * kk + 4 * lg + q) * NHID + 16 * nb + lr];
    } else if (u < PK_W2T) {  // W2 [128][128] as B
        const int w = u - PK_W2, kk = w / (8 * 64), nb = (w / 64) % 8;
        for (int q = 0; q < 4; ++q) v[q] = W2[(16 * kk + 4 * lg + q) * NHID + 16 * nb + lr];
    } else if (u < PK_W1T) {  // W2^T as B: B[k][n] = W2[n][k]
        const int w = u - PK_W2T, kk = w / (8 * 64), nb = (w / 64) % 8;
        for (int q = 0; q < 4; ++q) v[q] = W2[(16 * nb + lr) * NHID + 16 * kk + 4 * lg + q];
    } else {  // W1^T [128][64] as B: (kk, nb) of 8 x 4
        const int w = u - PK_W1T, kk = w / (4 * 64), nb = (w / 64) % 4;
        for (int q = 0; q < 4; ++q) v[q] = W1[(16 * nb + lr) * NHID + 16 * kk + 4 * lg + q];
    }
    PK[t] = pack4(v[0], v[1], v[2], v[3]);
}

// ------------------------------------------------------------------------------- one class, one tile

// Forward and backward of class c on the wave's 16-row tile.  A1: x' fragments (A layout).
// dyin[q]: weight of row 4 lg + q in the backward (1 for input gradients, dL/dE_row for
// training); E[q]: output energy of row 4 lg + q (every lane of the group), dx: D-layout input
// gradient (rows 4 lg + q, features 16 nb + lr), without istd.  With `store` (training) the D
// fragments of h1, dz2, dz1 are written to fr[0..23] (8 each) and the column sums of dz2, dz1 and
// of dy h2 are returned in cs (per lane: sum over the wave's rows for column 16 nb + lr).
struct TileOut {
    float E[4];
    v4f dx[4];
};

template <bool TRAIN>
__device__ __forceinline__ void tile_pass(const float *__restrict__ pc, const v4s *__restrict__ pk, const v4s A1[4],
                                          unsigned short (*T)[TP], int lane, const float dyin[4], TileOut &o,
                                          v4s *frag, float *cs_b1, float *cs_b2, float *cs_w3, float *cs_b3) {
    const int lr = lane & 15, lg = lane >> 4;
    const float scale = pc[O_TAIL + 1], shift = pc[O_TAIL + 2], b3 = pc[O_TAIL];
    // layer 1
    v4f z1[8];
#pragma unroll
    for (int nb = 0; nb < 8; ++nb) {
        v4f acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) acc = mfma(A1[kk], pk[PK_W1 + (kk * 8 + nb) * 64 + lane], acc);
        const float b = pc[O_B1 + 16 * nb + lr];
        for (int q = 0; q < 4; ++q) acc[q] += b;
        z1[nb] = acc;
    }
    wsync();
#pragma unroll
    for (int nb = 0; nb < 8; ++nb) {
        const float h0 = silu(z1[nb][0]), h1 = silu(z1[nb][1]), h2 = silu(z1[nb][2]), h3 = silu(z1[nb][3]);
        T[4 * lg + 0][16 * nb + lr] = bf16(h0);
        T[4 * lg + 1][16 * nb + lr] = bf16(h1);
        T[4 * lg + 2][16 * nb + lr] = bf16(h2);
        T[4 * lg + 3][16 * nb + lr] = bf16(h3);
        if (TRAIN) frag[nb * 64] = pack4(h0, h1, h2, h3);
    }
    wsync();
    v4s A[8];
#pragma unroll
    for (int kk = 0; kk < 8; ++kk) A[kk] = *reinterpret_cast<const v4s *>(&T[lr][16 * kk + 4 * lg]);
    // layer 2
    v4f z2[8];
#pragma unroll
    for (int nb = 0; nb < 8; ++nb) {
        v4f acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kk = 0; kk < 8; ++kk) acc = mfma(A[kk], pk[PK_W2 + (kk * 8 + nb) * 64 + lane], acc);
        const float b = pc[O_B2 + 16 * nb + lr];
        for (int q = 0; q < 4; ++q) acc[q] += b;
        z2[nb] = acc;
    }
    // output: y = w3 . h2 + b3 per row, reduced over the 16 lanes of a group
    float s[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int nb = 0; nb < 8; ++nb) {
        const float w = pc[O_W3 + 16 * nb + lr];
        for (int q = 0; q < 4; ++q) s[q] += w * silu(z2[nb][q]);
    }
    for (int q = 0; q < 4; ++q) {
        float v = s[q];
        v += __shfl_xor(v, 1, 64);
        v += __shfl_xor(v, 2, 64);
        v += __shfl_xor(v, 4, 64);
        v += __shfl_xor(v, 8, 64);
        o.E[q] = scale * (v + b3) + shift;
    }
    float dy[4];
    for (int q = 0; q < 4; ++q) dy[q] = scale * dyin[q];
    // backward: dz2 = dy w3 silu'(z2)
    wsync();
#pragma unroll
    for (int nb = 0; nb < 8; ++nb) {
        const float w = pc[O_W3 + 16 * nb + lr];
        float d[4];
        for (int q = 0; q < 4; ++q) d[q] = dy[q] * w * dsilu(z2[nb][q]);
        for (int q = 0; q < 4; ++q) T[4 * lg + q][16 * nb + lr] = bf16(d[q]);
        if (TRAIN) {
            frag[(8 + nb) * 64] = pack4(d[0], d[1], d[2], d[3]);
            float cb2 = d[0] + d[1] + d[2] + d[3];
            float cw3 = 0.f;
            for (int q = 0; q < 4; ++q) cw3 += dy[q] * silu(z2[nb][q]);
            cs_b2[nb] = cb2;
            cs_w3[nb] = cw3;
        }
    }
    if (TRAIN) *cs_b3 = dy[0] + dy[1] + dy[2] + dy[3];
    wsync();
#pragma unroll
    for (int kk = 0; kk < 8; ++kk) A[kk] = *reinterpret_cast<const v4s *>(&T[lr][16 * kk + 4 * lg]);
    v4f dh1[8];
#pragma unroll
    for (int nb = 0; nb < 8; ++nb) {
        v4f acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kk = 0; kk < 8; ++kk) acc = mfma(A[kk], pk[PK_W2T + (kk * 8 + nb) * 64 + lane], acc);
        dh1[nb] = acc;
    }
    wsync();
#pragma unroll
    for (int nb = 0; nb < 8; ++nb) {
        float d[4];
        for (int q = 0; q < 4; ++q) d[q] = dh1[nb][q] * dsilu(z1[nb][q]);
        for (int q = 0; q < 4; ++q) T[4 * lg + q][16 * nb + lr] = bf16(d[q]);
        if (TRAIN) {
            frag[(16 + nb) * 64] = pack4(d[0], d[1], d[2], d[3]);
            cs_b1[nb] = d[0] + d[1] + d[2] + d[3];
        }
    }
    wsync();
#pragma unroll
    for (int kk = 0; kk < 8; ++kk) A[kk] = *reinterpret_cast<const v4s *>(&T[lr][16 * kk + 4 * lg]);
#pragma unroll
    for (int nb = 0; nb < 4; ++nb) {
        v4f acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kk = 0; kk < 8; ++kk) acc = mfma(A[kk], pk[PK_W1T + (kk * 4 + nb) * 64 + lane], acc);
        o.dx[nb] = acc;
    }
}

// x' fragments of the tile (A layout: lane holds row r0 + l % 16, features 16 kk + 4 (l / 16) + q).
__device__ __forceinline__ void load_x(const float *__restrict__ P, const float *__restrict__ X, int n, int r0,
                                       int lane, v4s A1[4]) {
    const int row = r0 + (lane & 15), lg = lane >> 4;
    for (int kk = 0; kk < 4; ++kk) {
        const int k0 = 16 * kk + 4 * lg;
        float v[4] = {0.f, 0.f, 0.f, 0.f};
        if (row < n) {
            const float4 x = *reinterpret_cast<const float4 *>(X + (size_t)row * NFEAT + k0);
            v[0] = (x.x - P[k0 + 0]) * P[NFEAT + k0 + 0];
            v[1] = (x.y - P[k0 + 1]) * P[NFEAT + k0 + 1];
            v[2] = (x.z - P[k0 + 2]) * P[NFEAT + k0 + 2];
            v[3] = (x.w - P[k0 + 3]) * P[NFEAT + k0 + 3];
        }
        A1[kk] = pack4(v[0], v[1], v[2], v[3]);
    }
}

// u' = u istd fragments of the tile (A layout, as load_x).
__device__ __forceinline__ void load_u(const float *__restrict__ P, const float *__restrict__ U, int n, int r0,
                                       int lane, v4s U1[4]) {
    const int row = r0 + (lane & 15), lg = lane >> 4;
    for (int kk = 0; kk < 4; ++kk) {
        const int k0 = 16 * kk + 4 * lg;
        float v[4] = {0.f, 0.f, 0.f, 0.f};
        if (row < n) {
            const float4 u = *reinterpret_cast<const float4 *>(U + (size_t)row * NFEAT + k0);
            v[0] = u.x * P[NFEAT + k0 + 0];
            v[1] = u.y * P[NFEAT + k0 + 1];
            v[2] = u.z * P[NFEAT + k0 + 2];
            v[3] = u.w * P[NFEAT + k0 + 3];
        }
        U1[kk] = pack4(v[0], v[1], v[2], v[3]);
    }
}

// Tangent pass of class c on the wave's 16-row tile: the network's derivative along u per row,
// D = dE/dx . u = scale w3 . (silu'(z2) t2), t2 = W2^T (silu'(z1) t1), t1 = W1^T u', and the
// reverse pass of sum_rows wt D:
//   dt2 = g w3 silu'(z2),  dz2 = g w3 silu''(z2) t2  (g = scale wt of the row),
//   dt1 = (W2 dt2) silu'(z1),  dz1 = (W2 dz2) silu'(z1) + (W2 dt2) silu''(z1) t1,
//   dW2 = H1^T dZ2 + H1'^T dT2,  dW1 = X'^T dZ1 + U'^T dT1,  db2 = sum dz2,  db1 = sum dz1,
//   dw3 = sum g silu'(z2) t2 (b3 does not enter D).
// The D fragments of h1, h1' = silu'(z1) t1, dz2, dt2, dz1, dt1 go to fr[0..47] (8 each).
__device__ __forceinline__ void tile_tangent(const float *__restrict__ pc, const v4s *__restrict__ pk,
                                             const v4s A1[4], const v4s U1[4], unsigned short (*T)[TP], int lane,
                                             const float wt[4], float Dq[4], v4s *fr, float *cs_b1, float *cs_b2,
                                             float *cs_w3) {
    const int lr = lane & 15, lg = lane >> 4;
    const float scale = pc[O_TAIL + 1];
    v4f z1[8], t1[8];
#pragma unroll
    for (int nb = 0; nb < 8; ++nb) {
        v4f a = {0.f, 0.f, 0.f, 0.f}, t = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
            const v4s b = pk[PK_W1 + (kk * 8 + nb) * 64 + lane];
            a = mfma(A1[kk], b, a);
            t = mfma(U1[kk], b, t);
        }
        const float bb = pc[O_B1 + 16 * nb + lr];
        for (int q = 0; q < 4; ++q) a[q] += bb;
        z1[nb] = a;
        t1[nb] = t;
    }
    v4s A[8], At[8];
    wsync();
#pragma unroll
    for (int nb = 0; nb < 8; ++nb) {
        float h[4];
        for (int q = 0; q < 4; ++q) {
            h[q] = silu(z1[nb][q]);
            T[4 * lg + q][16 * nb + lr] = bf16(h[q]);
        }
        fr[nb * 64] = pack4(h[0], h[1], h[2], h[3]);
    }
    wsync();
#pragma unroll
    for (int kk = 0; kk < 8; ++kk) A[kk] = *reinterpret_cast<const v4s *>(&T[lr][16 * kk + 4 * lg]);
    wsync();
#pragma unroll
    for (int nb = 0; nb < 8; ++nb) {
        float h[4];
        for (int q = 0; q < 4; ++q) {
            h[q] = dsilu(z1[nb][q]) * t1[nb][q];
            T[4 * lg + q][16 * nb + lr] = bf16(h[q]);
        }
        fr[(8 + nb) * 64] = pack4(h[0], h[1], h[2], h[3]);
    }
    wsync();
#pragma unroll
    for (int kk = 0; kk < 8; ++kk) At[kk] = *reinterpret_cast<const v4s *>(&T[lr][16 * kk + 4 * lg]);
    // layer 2, primal and tangent
    v4f z2[8], t2[8];
#pragma unroll
    for (int nb = 0; nb < 8; ++nb) {
        v4f a = {0.f, 0.f, 0.f, 0.f}, t = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kk = 0; kk < 8; ++kk) {
            const v4s b = pk[PK_W2 + (kk * 8 + nb) * 64 + lane];
            a = mfma(A[kk], b, a);
            t = mfma(At[kk], b, t);
        }
        const float bb = pc[O_B2 + 16 * nb + lr];
        for (int q = 0; q < 4; ++q) a[q] += bb;
        z2[nb] = a;
        t2[nb] = t;
    }
    float g[4], s[4] = {0.f, 0.f, 0.f, 0.f};
    for (int q = 0; q < 4; ++q) g[q] = scale * wt[q];
    wsync();
    // D, dt2 (to LDS and fragments), dz2 (kept in z2 for the second product), bias / w3 sums
#pragma unroll
    for (int nb = 0; nb < 8; ++nb) {
        const float w = pc[O_W3 + 16 * nb + lr];
        float dt[4], dz[4], cw3 = 0.f;
        for (int q = 0; q < 4; ++q) {
            const float d1 = dsilu(z2[nb][q]), ht = d1 * t2[nb][q];
            s[q] += w * ht;
            cw3 += g[q] * ht;
            dt[q] = g[q] * w * d1;
            dz[q] = g[q] * w * d2silu(z2[nb][q]) * t2[nb][q];
            T[4 * lg + q][16 * nb + lr] = bf16(dt[q]);
        }
        fr[(16 + nb) * 64] = pack4(dz[0], dz[1], dz[2], dz[3]);
        fr[(24 + nb) * 64] = pack4(dt[0], dt[1], dt[2], dt[3]);
        cs_b2[nb] = dz[0] + dz[1] + dz[2] + dz[3];
        cs_w3[nb] = cw3;
        z2[nb] = v4f{dz[0], dz[1], dz[2], dz[3]};
    }
    for (int q = 0; q < 4; ++q) {
        float v = s[q];
        v += __shfl_xor(v, 1, 64);
        v += __shfl_xor(v, 2, 64);
        v += __shfl_xor(v, 4, 64);
        v += __shfl_xor(v, 8, 64);
        Dq[q] = scale * v;
    }
    wsync();
#pragma unroll
    for (int kk = 0; kk < 8; ++kk) A[kk] = *reinterpret_cast<const v4s *>(&T[lr][16 * kk + 4 * lg]);
    v4f dh1t[8];
#pragma unroll
    for (int nb = 0; nb < 8; ++nb) {
        v4f a = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kk = 0; kk < 8; ++kk) a = mfma(A[kk], pk[PK_W2T + (kk * 8 + nb) * 64 + lane], a);
        dh1t[nb] = a;
    }
    wsync();
#pragma unroll
    for (int nb = 0; nb < 8; ++nb)
        for (int q = 0; q < 4; ++q) T[4 * lg + q][16 * nb + lr] = bf16(z2[nb][q]);
    wsync();
#pragma unroll
    for (int kk = 0; kk < 8; ++kk) A[kk] = *reinterpret_cast<const v4s *>(&T[lr][16 * kk + 4 * lg]);
#pragma unroll
    for (int nb = 0; nb < 8; ++nb) {
        v4f a = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kk = 0; kk < 8; ++kk) a = mfma(A[kk], pk[PK_W2T + (kk * 8 + nb) * 64 + lane], a);
        float dz[4], dt[4];
        for (int q = 0; q < 4; ++q) {
            const float d1 = dsilu(z1[nb][q]);
            dt[q] = dh1t[nb][q] * d1;
            dz[q] = a[q] * d1 + dh1t[nb][q] * d2silu(z1[nb][q]) * t1[nb][q];
        }
        fr[(32 + nb) * 64] = pack4(dz[0], dz[1], dz[2], dz[3]);
        fr[(40 + nb) * 64] = pack4(dt[0], dt[1], dt[2], dt[3]);
        cs_b1[nb] = dz[0] + dz[1] + dz[2] + dz[3];
    }
}

// ------------------------------------------------------------------------------- inference

// E[row] (eV) and G[row][64] = dE/dx for rows 0..n-1, each row with its class cls[row].
__global__ __launch_bounds__(256) void k_infer(const float *__restrict__ P, const v4s *__restrict__ PK, int n,
                                               const float *__restrict__ X, const int *__restrict__ cls,
                                               float *__restrict__ E, float *__restrict__ G) {
    __shared__ __align__(16) unsigned short Ts[4][16][TP];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int r0 = (blockIdx.x * 4 + w) * 16;
    if (r0 >= n) return;  // whole wave; no block barrier follows
    const int lr = lane & 15, lg = lane >> 4;
    v4s A1[4];
    load_x(P, X, n, r0, lane, A1);
    const int ca = r0 + lr < n ? cls[r0 + lr] : -1;
    int cc[4];
    for (int q = 0; q < 4; ++q) cc[q] = r0 + 4 * lg + q < n ? cls[r0 + 4 * lg + q] : -1;
    float Eo[4] = {0.f, 0.f, 0.f, 0.f};
    v4f Go[4];
    for (int nb = 0; nb < 4; ++nb) Go[nb] = v4f{0.f, 0.f, 0.f, 0.f};
    const float one[4] = {1.f, 1.f, 1.f, 1.f};
    for (int c = 0; c < NCLASS; ++c) {
        if (__ballot(ca == c) == 0) continue;  // uniform
        TileOut o;
        tile_pass<false>(P + HEAD + c * CLASS_SIZE, PK + c * PK_CLASS, A1, Ts[w], lane, one, o, nullptr, nullptr,
                         nullptr, nullptr, nullptr);
        for (int q = 0; q < 4; ++q)
            if (cc[q] == c) {
                Eo[q] = o.E[q];
                for (int nb = 0; nb < 4; ++nb) Go[nb][q] = o.dx[nb][q] * P[NFEAT + 16 * nb + lr];
            }
    }
    for (int q = 0; q < 4; ++q) {
        const int row = r0 + 4 * lg + q;
        if (row >= n) continue;
        if (lr == 0 && E) E[row] = Eo[q];
        if (G)
            for (int nb = 0; nb < 4; ++nb) G[(size_t)row * NFEAT + 16 * nb + lr] = Go[nb][q];
    }
}

// ------------------------------------------------------------------------------- training

// Rows 0..n-1, all of class c: E, loss += sum 0.5 w (E - y)^2, per-row backward of
// dL/dE = w (E - y); stores the h1 / dz2 / dz1 fragments of every tile in `frag` (24 x 64 v4s per
// tile) and adds the bias and w3 gradients to grad.
__global__ __launch_bounds__(256) void k_train(const float *__restrict__ P, const v4s *__restrict__ PK, int c, int n,
                                               const float *__restrict__ X, const float *__restrict__ Y,
                                               const float *__restrict__ Wt, float *__restrict__ E,
                                               v4s *__restrict__ frag, float *__restrict__ grad,
                                               float *__restrict__ loss) {
    __shared__ __align__(16) unsigned short Ts[4][16][TP];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int tile = blockIdx.x * 4 + w, r0 = tile * 16;
    if (r0 >= n) return;
    const int lr = lane & 15, lg = lane >> 4;
    const float *pc = P + HEAD + c * CLASS_SIZE;
    const v4s *pk = PK + c * PK_CLASS;
    v4s A1[4];
    load_x(P, X, n, r0, lane, A1);
    // forward only first: E of the rows, then dL/dE (rows past n get weight 0)
    TileOut o;
    float dl[4];
    {
        // a forward pass is part of tile_pass; run it once with unit dy to get E, then the training
        // pass with the loss gradient (the forward is recomputed: cheap next to the data movement)
        const float one[4] = {1.f, 1.f, 1.f, 1.f};
        tile_pass<false>(pc, pk, A1, Ts[w], lane, one, o, nullptr, nullptr, nullptr, nullptr, nullptr);
        float lsum = 0.f;
        for (int q = 0; q < 4; ++q) {
            const int row = r0 + 4 * lg + q;
            const bool v = row < n;
            const float wt = v ? Wt[row] : 0.f, y = v ? Y[row] : 0.f;
            const float r = o.E[q] - y;
            dl[q] = wt * r;
            lsum += 0.5f * wt * r * r;
            if (v && lr == 0) E[row] = o.E[q];
        }
        // rows are replicated over the 16 lanes of a group: count each once
        lsum = lr == 0 ? lsum : 0.f;
        for (int off = 32; off > 0; off >>= 1) lsum += __shfl_xor(lsum, off, 64);
        if (lane == 0) atomicAdd(loss, lsum);
    }
    float cs_b1[8], cs_b2[8], cs_w3[8], cs_b3;
    v4s *fr = frag + (size_t)tile * 24 * 64 + lane;
    tile_pass<true>(pc, pk, A1, Ts[w], lane, dl, o, fr, cs_b1, cs_b2, cs_w3, &cs_b3);
    // column sums over the wave's rows: lanes of equal l % 16 hold the same column
    float *g = grad + HEAD + c * CLASS_SIZE;
    for (int nb = 0; nb < 8; ++nb) {
        float a = cs_b1[nb], b = cs_b2[nb], d = cs_w3[nb];
        a += __shfl_xor(a, 16, 64);
        a += __shfl_xor(a, 32, 64);
        b += __shfl_xor(b, 16, 64);
        b += __shfl_xor(b, 32, 64);
        d += __shfl_xor(d, 16, 64);
        d += __shfl_xor(d, 32, 64);
        if (lg == 0) {
            atomicAdd(&g[O_B1 + 16 * nb + lr], a);
            atomicAdd(&g[O_B2 + 16 * nb + lr], b);
            atomicAdd(&g[O_W3 + 16 * nb + lr], d);
        }
    }
    float b3 = lr == 0 ? cs_b3 : 0.f;
    for (int off = 32; off > 0; off >>= 1) b3 += __shfl_xor(b3, off, 64);
    if (lane == 0) atomicAdd(&g[O_TAIL], b3);
}

// Weight gradients of class c: dW2 = H1^T dZ2 (64 blocks of 16 x 16) and dW1 = X'^T dZ1 (32 blocks)
// over tiles [chunk * tpc, (chunk + 1) * tpc); blockIdx.x = output block, blockIdx.y = chunk; one
// wave per block.
__global__ __launch_bounds__(64) void k_wgrad(const float *__restrict__ P, int c, int n, const float *__restrict__ X,
                                              const v4s *__restrict__ frag, int tpc, float *__restrict__ grad) {
    const int lane = threadIdx.x, lr = lane & 15, lg = lane >> 4;
    const int blk = blockIdx.x, ntile = (n + 15) / 16;
    const int t0 = blockIdx.y * tpc, t1 = min(ntile, t0 + tpc);
    v4f acc = {0.f, 0.f, 0.f, 0.f};
    float *g = grad + HEAD + c * CLASS_SIZE;
    if (blk < 64) {  // dW2 (kb, nb)
        const int kb = blk / 8, nb = blk % 8;
        for (int t = t0; t < t1; ++t) {
            const v4s *f = frag + (size_t)t * 24 * 64;
            acc = mfma(f[kb * 64 + lane], f[(8 + nb) * 64 + lane], acc);
        }
        for (int q = 0; q < 4; ++q) atomicAdd(&g[O_W2 + (16 * kb + 4 * lg + q) * NHID + 16 * nb + lr], acc[q]);
    } else {  // dW1 (fb, nb)
        const int b = blk - 64, fb = b / 8, nb = b % 8;
        const int f = 16 * fb + lr;
        const float mu = P[f], is = P[NFEAT + f];
        for (int t = t0; t < t1; ++t) {
            float v[4];
            for (int q = 0; q < 4; ++q) {
                const int row = 16 * t + 4 * lg + q;
                v[q] = row < n ? (X[(size_t)row * NFEAT + f] - mu) * is : 0.f;
            }
            const v4s *fr = frag + (size_t)t * 24 * 64;
            acc = mfma(pack4(v[0], v[1], v[2], v[3]), fr[(16 + nb) * 64 + lane], acc);
        }
        for (int q = 0; q < 4; ++q) atomicAdd(&g[O_W1 + (16 * fb + 4 * lg + q) * NHID + 16 * nb + lr], acc[q]);
    }
}

// Force loss, rows 0..n-1 all of class c: D[row] = dE/dx . u (tangent pass) and the bias / w3
// parts of d(sum wt D)/dparams; the 48 fragments of each tile go to `frag` for k_wgrad_tan.
__global__ __launch_bounds__(256) void k_train_tan(const float *__restrict__ P, const v4s *__restrict__ PK, int c,
                                                   int n, const float *__restrict__ X, const float *__restrict__ U,
                                                   const float *__restrict__ Wt, float *__restrict__ D,
                                                   v4s *__restrict__ frag, float *__restrict__ grad) {
    __shared__ __align__(16) unsigned short Ts[4][16][TP];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int tile = blockIdx.x * 4 + w, r0 = tile * 16;
    if (r0 >= n) return;  // whole wave; no block barrier follows
    const int lr = lane & 15, lg = lane >> 4;
    v4s A1[4], U1[4];
    load_x(P, X, n, r0, lane, A1);
    load_u(P, U, n, r0, lane, U1);
    float wt[4], Dq[4];
    for (int q = 0; q < 4; ++q) {  // rows past n get weight 0: every gradient term of theirs is 0
        const int row = r0 + 4 * lg + q;
        wt[q] = row < n ? Wt[row] : 0.f;
    }
    float cs_b1[8], cs_b2[8], cs_w3[8];
    tile_tangent(P + HEAD + c * CLASS_SIZE, PK + c * PK_CLASS, A1, U1, Ts[w], lane, wt, Dq,
                 frag + (size_t)tile * 48 * 64 + lane, cs_b1, cs_b2, cs_w3);
    for (int q = 0; q < 4; ++q) {
        const int row = r0 + 4 * lg + q;
        if (row < n && lr == 0) D[row] = Dq[q];
    }
    float *g = grad + HEAD + c * CLASS_SIZE;
    for (int nb = 0; nb < 8; ++nb) {
        float a = cs_b1[nb], b = cs_b2[nb], d = cs_w3[nb];
        a += __shfl_xor(a, 16, 64);
        a += __shfl_xor(a, 32, 64);
        b += __shfl_xor(b, 16, 64);
        b += __shfl_xor(b, 32, 64);
        d += __shfl_xor(d, 16, 64);
        d += __shfl_xor(d, 32, 64);
        if (lg == 0) {
            atomicAdd(&g[O_B1 + 16 * nb + lr], a);
            atomicAdd(&g[O_B2 + 16 * nb + lr], b);
            atomicAdd(&g[O_W3 + 16 * nb + lr], d);
        }
    }
}

// Weight gradients of the force loss, as k_wgrad with two products per block:
// dW2 = H1^T dZ2 + H1'^T dT2 (64 blocks), dW1 = X'^T dZ1 + U'^T dT1 (32 blocks).
__global__ __launch_bounds__(64) void k_wgrad_tan(const float *__restrict__ P, int c, int n,
                                                  const float *__restrict__ X, const float *__restrict__ U,
                                                  const v4s *__restrict__ frag, int tpc, float *__restrict__ grad) {
    const int lane = threadIdx.x, lr = lane & 15, lg = lane >> 4;
    const int blk = blockIdx.x, ntile = (n + 15) / 16;
    const int t0 = blockIdx.y * tpc, t1 = min(ntile, t0 + tpc);
    v4f acc = {0.f, 0.f, 0.f, 0.f};
    float *g = grad + HEAD + c * CLASS_SIZE;
    if (blk < 64) {
        const int kb = blk / 8, nb = blk % 8;
        for (int t = t0; t < t1; ++t) {
            const v4s *f = frag + (size_t)t * 48 * 64;
            acc = mfma(f[kb * 64 + lane], f[(16 + nb) * 64 + lane], acc);
            acc = mfma(f[(8 + kb) * 64 + lane], f[(24 + nb) * 64 + lane], acc);
        }
        for (int q = 0; q < 4; ++q) atomicAdd(&g[O_W2 + (16 * kb + 4 * lg + q) * NHID + 16 * nb + lr], acc[q]);
    } else {
        const int b = blk - 64, fb = b / 8, nb = b % 8;
        const int f = 16 * fb + lr;
        const float mu = P[f], is = P[NFEAT + f];
        for (int t = t0; t < t1; ++t) {
            float v[4], u[4];
            for (int q = 0; q < 4; ++q) {
                const int row = 16 * t + 4 * lg + q;
                v[q] = row < n ? (X[(size_t)row * NFEAT + f] - mu) * is : 0.f;
                u[q] = row < n ? U[(size_t)row * NFEAT + f] * is : 0.f;
            }
            const v4s *fr = frag + (size_t)t * 48 * 64;
            acc = mfma(pack4(v[0], v[1], v[2], v[3]), fr[(32 + nb) * 64 + lane], acc);
            acc = mfma(pack4(u[0], u[1], u[2], u[3]), fr[(40 + nb) * 64 + lane], acc);
        }
        for (int q = 0; q < 4; ++q) atomicAdd(&g[O_W1 + (16 * fb + 4 * lg + q) * NHID + 16 * nb + lr], acc[q]);
    }
}

// Tied training: parameter k of class 0's block stands for its copies in the other classes and,
// for a radial row of W1 (row 1 + 21 c + m), for the rows of the same radial centre m in the
// other neighbour-class blocks; the gradient of every copy becomes the sum over the copies.
__global__ void k_tie(float *__restrict__ grad, int blocks) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    // owners: the W1 rows (with blocks only CN's and the first block's), then every entry from b1 on
    const int nw1 = (blocks ? 1 + NRBF : NFEAT) * NHID, nrest = CLASS_SIZE - O_B1;
    if (t >= nw1 + nrest) return;
    const int k = t < nw1 ? t : O_B1 + (t - nw1);
    const int row = t < nw1 ? t / NHID : -1, col = t < nw1 ? t % NHID : 0;
    const int nb = (blocks && row >= 1) ? 3 : 1;
    float s = 0.f;
    for (int c = 0; c < NCLASS; ++c)
        for (int b = 0; b < nb; ++b) {
            const int kk = b == 0 ? k : O_W1 + (row + NRBF * b) * NHID + col;
            s += grad[HEAD + c * CLASS_SIZE + kk];
        }
    for (int c = 0; c < NCLASS; ++c)
        for (int b = 0; b < nb; ++b) {
            const int kk = b == 0 ? k : O_W1 + (row + NRBF * b) * NHID + col;
            grad[HEAD + c * CLASS_SIZE + kk] = s;
        }
}

// Adam on the trainable entries (W1, b1, W2, b2, w3, b3 of every class); clears the gradient.
__global__ void k_adam(float *__restrict__ P, float *__restrict__ grad, float *__restrict__ m, float *__restrict__ v,
                       float lr, float b1, float b2, float eps, float c1, float c2) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= NPARAMS) return;
    const float g = grad[t];
    grad[t] = 0.f;
    if (t < HEAD || (t - HEAD) % CLASS_SIZE > O_TAIL) return;
    const float mm = b1 * m[t] + (1.f - b1) * g, vv = b2 * v[t] + (1.f - b2) * g * g;
    m[t] = mm;
    v[t] = vv;
    P[t] -= lr * (mm * c1) / (sqrtf(vv * c2) + eps);
}

}  // namespace

void launch_pack(const float *P, void *PK, hipStream_t s) {
    const int tot = NCLASS * PK_CLASS;
    hipLaunchKernelGGL(k_pack, dim3((tot + 255) / 256), dim3(256), 0, s, P, (v4s *)PK);
}

void launch_infer(const float *P, const void *PK, int n, const float *X, const int *cls, float *E, float *G,
                  hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_infer, dim3((n + 63) / 64), dim3(256), 0, s, P, (const v4s *)PK, n, X, cls, E, G);
}

void launch_train(const float *P, const void *PK, int c, int n, const float *X, const float *Y, const float *Wt,
                  float *E, void *work, float *grad, float *loss, hipStream_t s) {
    if (n <= 0) return;
    const int ntile = (n + 15) / 16;
    hipLaunchKernelGGL(k_train, dim3((ntile + 3) / 4), dim3(256), 0, s, P, (const v4s *)PK, c, n, X, Y, Wt, E,
                       (v4s *)work, grad, loss);
    const int tpc = 64;
    hipLaunchKernelGGL(k_wgrad, dim3(96, (ntile + tpc - 1) / tpc), dim3(64), 0, s, P, c, n, X, (const v4s *)work,
                       tpc, grad);
}

namespace {
std::mutex g_scratch_mtx;
std::vector<void *> g_scratch_ptr;  // address of `scratch` per device (nullptr: not looked up yet)
}  // namespace

hipError_t static_train_scratch(void **ptr) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    std::lock_guard<std::mutex> lock(g_scratch_mtx);
    if ((size_t)dev >= g_scratch_ptr.size()) g_scratch_ptr.resize((size_t)dev + 1, nullptr);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));  // with the device current
        if (e != hipSuccess) return e;
        g_scratch_ptr[dev] = p;
    }
    *ptr = g_scratch_ptr[dev];
    return hipSuccess;
}

void launch_train_force(const float *P, const void *PK, int c, int n, const float *X, const float *U, const float *Wt,
                        float *D, void *work, float *grad, hipStream_t s) {
    if (n <= 0) return;
    const int ntile = (n + 15) / 16;
    hipLaunchKernelGGL(k_train_tan, dim3((ntile + 3) / 4), dim3(256), 0, s, P, (const v4s *)PK, c, n, X, U, Wt, D,
                       (v4s *)work, grad);
    const int tpc = 64;
    hipLaunchKernelGGL(k_wgrad_tan, dim3(96, (ntile + tpc - 1) / tpc), dim3(64), 0, s, P, c, n, X, U,
                       (const v4s *)work, tpc, grad);
}

void launch_tie(float *grad, int blocks, hipStream_t s) {
    const int n = (blocks ? 1 + NRBF : NFEAT) * NHID + (CLASS_SIZE - O_B1);
    hipLaunchKernelGGL(k_tie, dim3((n + 255) / 256), dim3(256), 0, s, grad, blocks);
}

void launch_adam(float *P, float *grad, float *m, float *v, float lr, float b1, float b2, float eps, int step,
                 hipStream_t s) {
    const float c1 = 1.f / (1.f - powf(b1, (float)step)), c2 = 1.f / (1.f - powf(b2, (float)step));
    hipLaunchKernelGGL(k_adam, dim3((NPARAMS + 255) / 256), dim3(256), 0, s, P, grad, m, v, lr, b1, b2, eps, c1, c2);
}

}  // namespace sur
}  // namespace d3
