// Learned stand-in for the three-body term: constants and parameter layout shared by the host
// engine and the HIP kernels (d3_mlp.hip, d3_sur.hip).  No counterpart in the reference, which
// has only the exact ATM kernel (src/d3_kernel.cu:1199-1509).
//
// Per atom i: features x_i[64] = {CN_i, 63 Gaussian radial sums} (21 centres 1..8 A, sigma 0.5 A,
// over neighbours of each of 3 element classes, smooth cosine cutoff at 9.5 A, each neighbour
// weighted by sqrt(C6_ij) C6_jj^(1/4) times the edge damping h_ij), then the network
// of the atom's own class: x' = (x - mean) * istd, h1 = silu(W1^T x' + b1), h2 = silu(W2^T h1 + b2),
// E_i = scale (w3 . h2 + b3) + shift (eV).
//
// Parameter vector (fp32): mean[64] | istd[64] | per class: W1[64][128] | b1[128] | W2[128][128] |
// b2[128] | w3[128] | b3, scale, shift, pad.
//
// Packed weights for the MFMA kernels (bf16, in 16x16x16 B-operand fragments of 64 lanes x 4
// values): per class W1 [4 kk][8 nb], W2 [8][8], W2^T [8][8], W1^T [8 kk][4 nb]; the fragment
// (kk, nb) of a K x N matrix B holds in lane l the values B[16 kk + 4 (l / 16) + q][16 nb + l % 16],
// q = 0..3.
#pragma once
#include <hip/hip_runtime_api.h>

#include <cstddef>

namespace d3 {
namespace sur {

constexpr int NFEAT = 64, NHID = 128, NRBF = 21, NCLASS = 3;
constexpr double BOHR_A = 0.52917726;
constexpr double SIGMA = 0.5 / BOHR_A;      // bohr
constexpr double MU0 = 1.0 / BOHR_A;        // first centre (bohr)
constexpr double DMU = 0.35 / BOHR_A;       // centre spacing (7 A / 20)
constexpr double RCF = 9.5 / BOHR_A;        // feature cutoff (bohr)

constexpr int HEAD = 2 * NFEAT;
constexpr int O_W1 = 0, O_B1 = NFEAT * NHID, O_W2 = O_B1 + NHID, O_B2 = O_W2 + NHID * NHID, O_W3 = O_B2 + NHID,
              O_TAIL = O_W3 + NHID, CLASS_SIZE = O_TAIL + 4;
constexpr int NPARAMS = HEAD + NCLASS * CLASS_SIZE;

// packed fragments (units of 4 bf16 = 8 bytes)
constexpr int PK_W1 = 0, PK_W2 = 4 * 8 * 64, PK_W2T = PK_W2 + 8 * 8 * 64, PK_W1T = PK_W2T + 8 * 8 * 64,
              PK_CLASS = PK_W1T + 8 * 4 * 64;
constexpr size_t PACKED_BYTES = (size_t)NCLASS * PK_CLASS * 8;

// element class of atomic number Z: 0 = H, He; 1 = Li..Ne; 2 = the rest
inline int element_class(int Z) { return Z <= 2 ? 0 : (Z <= 10 ? 1 : 2); }

// Training scratch per row of a batch (bytes): fragments of h1, dz2, dz1 (bf16) for the energy
// loss; of h1, h1', dz2, dt2, dz1, dt1 for the force loss (tangent pass).
constexpr size_t TRAIN_ROW_BYTES = 3 * NHID * 2, TRAIN_FORCE_ROW_BYTES = 6 * NHID * 2;

// Rows of one call of the training gradient with static scratch (d3x_sur_train_grad_noalloc):
// the fragment scratch is a file-scope device array of STATIC_TRAIN_ROWS * TRAIN_ROW_BYTES
// (12 MiB), one copy per device for the life of the process.
constexpr int STATIC_TRAIN_ROWS = 16384;
// Device address of that array on the current device (looked up once per device).
hipError_t static_train_scratch(void **ptr);

// Launchers (d3_mlp.hip).  P: fp32 parameters, PK: packed bf16 fragments (PACKED_BYTES).
void launch_pack(const float *P, void *PK, hipStream_t s);
void launch_infer(const float *P, const void *PK, int n, const float *X, const int *cls, float *E, float *G,
                  hipStream_t s);
void launch_train(const float *P, const void *PK, int c, int n, const float *X, const float *Y, const float *Wt,
                  float *E, void *work, float *grad, float *loss, hipStream_t s);
void launch_train_force(const float *P, const void *PK, int c, int n, const float *X, const float *U, const float *Wt,
                        float *D, void *work, float *grad, hipStream_t s);
void launch_adam(float *P, float *grad, float *m, float *v, float lr, float b1, float b2, float eps, int step,
                 hipStream_t s);
// Tied training: every class's gradient entry := the sum over the classes; with blocks, the
// radial rows of W1 summed over the three neighbour-class blocks as well.
void launch_tie(float *grad, int blocks, hipStream_t s);

}  // namespace sur
}  // namespace d3
