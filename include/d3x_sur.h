/* Learned stand-in for the ATM three-body term — C API.
 *
 * No counterpart in the reference, whose only three-body path is the exact ATM kernel
 * (src/d3_kernel.cu:1199-1509, reached through compute_dispersion_energy_from_handle in
 * src/d3.h:114-128).  The stand-in is a per-atom network: features x_i[64] = {CN_i, 63 Gaussian
 * radial sums over neighbours grouped in 3 element classes (21 centres 1..8 A, sigma 0.5 A,
 * smooth cutoff 9.5 A)}, each neighbour weighted by sqrt(C6_ij) C6_jj^(1/4) (D3 C6 at the atoms'
 * CNs) times (1 + 6 ((4/3) R0_ij / r_ij)^16)^(-1/2) (R0 from r0ab); per class an MLP
 * 64 -> 128 -> 128 -> 1 (SiLU) whose two hidden layers and input-gradient backward run on bf16
 * MFMA with fp32 accumulation; E = sum_i E_i, forces and stress by the hand-written backward
 * through the features, their weights and CN.
 *
 * Parameter vector (D3X_SUR_NPARAMS floats): mean[64] | istd[64] | per class c = 0, 1, 2:
 * W1[64][128] | b1[128] | W2[128][128] | b2[128] | w3[128] | b3, scale, shift, pad;
 * E_i = scale (w3 . silu(W2^T silu(W1^T ((x - mean) istd) + b1) + b2) + b3) + shift  [eV].
 * Classes: 0 = H, He; 1 = Li..Ne; 2 = heavier elements.
 */
#ifndef D3X_SUR_H
#define D3X_SUR_H

#include <stddef.h>
#include <stdint.h>

#include "d3.h"

#ifdef __cplusplus
extern "C" {
#endif

#define D3X_SUR_NPARAMS 75020
#define D3X_SUR_NFEAT 64

/* Three-body mode of a handle.  EXACT: the ATM kernel (default).  SURROGATE: the network; every
 * judge_every-th evaluation (and the first after atoms or cell change) also runs the exact kernel
 * and, when |E_net - E_exact| / atoms > tol_ev_per_atom (or the force criterion of
 * d3x_sur_set_force_tol fails), returns the exact result and keeps the handle on the exact kernel
 * from then on (the fallback; d3x_sur_state reports it).
 * judge_every = 0 judges only the first evaluation after a geometry change.
 * EXACT_FEATURES: exact kernel, and the network's features are computed too (d3x_sur_features),
 * for building training sets. */
#define D3X_ATM_EXACT 0
#define D3X_ATM_SURROGATE 1
#define D3X_ATM_EXACT_FEATURES 2

/* Network parameters for a handle (host array of D3X_SUR_NPARAMS floats, or a file written by
 * lasp-d3_amd/surrogate/train.py: 8 bytes "D3SUR001", uint32 count, count floats). */
int d3x_sur_load(d3_handle_t *handle, const float *params);
int d3x_sur_load_file(d3_handle_t *handle, const char *path);
int d3x_set_atm_mode(d3_handle_t *handle, int mode, int judge_every, double tol_ev_per_atom);
/* mode in use (after fallbacks), fallback count, |E_net - E_exact| / atoms of the last judged
 * evaluation (eV; negative before any judgement).  Any pointer may be NULL. */
int d3x_sur_state(d3_handle_t *handle, int *mode, int *fallbacks, double *last_dev_ev_per_atom);
/* Force criterion of the judge: a judged evaluation also falls back when, in any cell, the RMS
 * over the cell's force components of F_net - F_exact exceeds tol_rms_ev_per_A (eV/A; default
 * 5e-3; <= 0 judges energies only).  d3x_sur_last_force_dev: that RMS deviation of the last judged
 * evaluation (largest over cells; negative before any judgement or for a NULL handle). */
int d3x_sur_set_force_tol(d3_handle_t *handle, double tol_rms_ev_per_A);
double d3x_sur_last_force_dev(d3_handle_t *handle);
/* Features of the last evaluation (SURROGATE or EXACT_FEATURES), caller order, atoms x 64. */
int d3x_sur_features(d3_handle_t *handle, float *features);
/* Push-forward of a direction field through the features, for force matching: U[i] =
 * sum_j (dx_i / dr_j) . dir[j] (caller order; dir atoms x 3 per Angstrom of position step, U atoms
 * x 64), at the geometry of the last evaluation (SURROGATE or EXACT_FEATURES; a later set_atoms or
 * set_cell invalidates it).  With dir = F_net - F_ref (eV/A), U is the input of
 * d3x_sur_train_force_grad. */
int d3x_sur_tangent(d3_handle_t *handle, const double *dir, float *U);

/* Network kernels on device buffers (pointers from hipMalloc or a framework's allocator, stream
 * = hipStream_t or NULL for the default stream).  They enqueue and return; errors of the launch
 * are returned as D3X status bits. */
size_t d3x_sur_packed_bytes(void);
/* bytes of scratch for d3x_sur_train_grad or d3x_sur_train_force_grad on `rows` rows */
size_t d3x_sur_train_bytes(uint64_t rows);
/* bf16 MFMA fragments of every weight matrix from the fp32 parameters */
int d3x_sur_pack(const float *params, void *packed, void *stream);
/* per-row energies E[n] (eV) and input gradients G[n][64] (either may be NULL) of rows X[n][64]
 * with classes cls[n] */
int d3x_sur_infer(const float *params, const void *packed, uint64_t n, const float *X, const int32_t *cls,
                  float *E, float *G, void *stream);
/* one training gradient on n rows of class cls: E[n] (out), loss += sum 0.5 w (E - y)^2, and
 * grad (D3X_SUR_NPARAMS floats, accumulated) += dloss/dparams */
int d3x_sur_train_grad(const float *params, const void *packed, int cls, uint64_t n, const float *X, const float *y,
                       const float *w, float *E, void *work, float *grad, float *loss, void *stream);
/* No counterpart in the reference: d3x_sur_train_grad for callers that cannot allocate device
 * memory (code under HIP graph capture, training loops that forbid allocation).  Same arguments
 * and results without the workspace: the kernels' scratch is a static device array of the
 * library (12 MiB per device, held until the process ends), so n is at most
 * D3X_SUR_NOALLOC_MAX_ROWS (larger n returns D3X_ERR_INPUT without launching).  Calls share that
 * array: calls that overlap on one device (other streams, other threads) race on it. */
#define D3X_SUR_NOALLOC_MAX_ROWS 16384
int d3x_sur_train_grad_noalloc(const float *params, const void *packed, int cls, uint64_t n, const float *X,
                               const float *y, const float *w, float *E, float *grad, float *loss, void *stream);
/* one force-matching gradient on n rows of class cls: D[n] (out) = dE/dx . U of each row (the
 * network's derivative along U, eV per unit of U) and grad += d(sum_rows w D)/dparams.  Since
 * F_net = -J^T dE/dx, the force loss L = sum |F_net - F_ref|^2 has dL/dparams = -2 sum_rows
 * d(dE/dx . U)/dparams with U = J (F_net - F_ref) (d3x_sur_tangent): pass w = -2 lambda for
 * lambda L.  D summed over a cell's rows is -sum F_net . (F_net - F_ref). */
int d3x_sur_train_force_grad(const float *params, const void *packed, int cls, uint64_t n, const float *X,
                             const float *U, const float *w, float *D, void *work, float *grad, void *stream);
/* Tied training (one network for every element class): each trainable entry's gradient becomes
 * the sum over the three class blocks; blocks != 0 also sums the radial rows of W1 over the three
 * neighbour-class blocks (rows 1 + 21 c + m for c = 0, 1, 2), so that the network sees the
 * neighbour sums of all classes through one set of weights.  Identical copies stay identical
 * under d3x_sur_adam. */
int d3x_sur_tie_grad(float *grad, int blocks, void *stream);
/* Adam step (bias-corrected, step >= 1) on the trainable entries; clears grad */
int d3x_sur_adam(float *params, float *grad, float *m, float *v, float lr, float beta1, float beta2, float eps,
                 int step, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* D3X_SUR_H */
