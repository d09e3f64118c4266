// Deterministic per-cell reduction of per-atom energy, virial and pair counts (FP64), and the
// scatter of per-atom results back to the caller's atom order with unit conversion.
// Replaces the host FP64 energy sum and unit conversion of src/d3.cu:182-219; the stress is
// reduced in a fixed order instead of the reference's float atomicAdd (src/d3_kernel.cu:1193).
#include <hip/hip_runtime.h>

#include "d3_engine.hpp"

namespace d3 {
namespace {

constexpr int RB = 256;

__device__ double block_sum_d(double v, double *red) {
    red[threadIdx.x] = v;
    __syncthreads();
    for (int o = RB / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
}

__global__ __launch_bounds__(RB) void k_reduce_chunks(const double *e, const double *eat, const double *vir,
                                                       const float *npair, const int *cb, const int *ce,
                                                       double *partial) {
    __shared__ double red[RB];
    const int c = blockIdx.x;
    double acc[RED_W];
    for (int k = 0; k < RED_W; ++k) acc[k] = 0.0;
    for (int p = cb[c] + threadIdx.x; p < ce[c]; p += RB) {
        acc[0] += e[p] + eat[p];  // two-body half-pair energies + anchored ATM triangle energies
        for (int k = 0; k < 6; ++k) acc[1 + k] += vir[6 * p + k];
        for (int k = 0; k < 3; ++k) acc[7 + k] += (double)npair[3 * p + k];
    }
    for (int k = 0; k < RED_W; ++k) {
        const double s = block_sum_d(acc[k], red);
        if (threadIdx.x == 0) partial[(size_t)c * RED_W + k] = s;
    }
}

__global__ __launch_bounds__(RB) void k_reduce_sys(const double *partial, const int *sys_chunk_off, double *out) {
    __shared__ double red[RB];
    const int s = blockIdx.x;
    double acc[RED_W];
    for (int k = 0; k < RED_W; ++k) acc[k] = 0.0;
    for (int c = sys_chunk_off[s] + threadIdx.x; c < sys_chunk_off[s + 1]; c += RB)
        for (int k = 0; k < RED_W; ++k) acc[k] += partial[(size_t)c * RED_W + k];
    for (int k = 0; k < RED_W; ++k) {
        const double v = block_sum_d(acc[k], red);
        if (threadIdx.x == 0) out[(size_t)s * RED_W + k] = v;
    }
}

template <class T>
__global__ void k_finalize(Params<T> P, double *out_f, double *out_e, double *out_cn, double *out_g) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P.n) return;
    const int o = P.orig_s[p];
    const double fs = HARTREE_EV / BOHR_A;  // Hartree/bohr -> eV/Angstrom, src/d3.cu:207-212
    for (int k = 0; k < 3; ++k) out_f[3 * o + k] = P.f[3 * p + k] * fs;
    out_e[o] = -HARTREE_EV * (P.e[p] + P.eas[p]);  // E = -Hartree * E+, src/d3.cu:206
    out_cn[o] = (double)P.cn[p];
    out_g[o] = P.g[p];
}

__global__ void k_fill_u32(unsigned *p, int n, unsigned v) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = v;
}

}  // namespace

void launch_fill_u32(unsigned *p, int n, unsigned v, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_fill_u32, dim3((n + 255) / 256), dim3(256), 0, s, p, n, v);
}

void launch_reduce(const double *e, const double *eat, const double *vir, const float *npair, const int *chunk_begin,
                   const int *chunk_end, const int * /*chunk_sys*/, int nchunks, const int *sys_chunk_off, int nsys,
                   double *partial, double *out_sys, hipStream_t s) {
    if (nchunks > 0)
        hipLaunchKernelGGL(k_reduce_chunks, dim3(nchunks), dim3(RB), 0, s, e, eat, vir, npair, chunk_begin,
                           chunk_end, partial);
    hipLaunchKernelGGL(k_reduce_sys, dim3(nsys), dim3(RB), 0, s, partial, sys_chunk_off, out_sys);
}

template <class T>
void launch_finalize(const Params<T> &P, double *out_f, double *out_e, double *out_cn, double *out_g,
                     hipStream_t s) {
    if (P.n == 0) return;
    hipLaunchKernelGGL(k_finalize<T>, dim3((P.n + 255) / 256), dim3(256), 0, s, P, out_f, out_e, out_cn, out_g);
}

template void launch_finalize<double>(const Params<double> &, double *, double *, double *, double *, hipStream_t);
template void launch_finalize<float>(const Params<float> &, double *, double *, double *, double *, hipStream_t);

}  // namespace d3
