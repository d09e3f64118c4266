/* D3 dispersion C API — extensions beside the reference's d3.h.
 *
 * These entry points have no counterpart in the reference (src/d3.h); they add what the float
 * API cannot carry: FP64 inputs/outputs (a float total energy resolves only ~6e-8 relative),
 * a precision mode, custom functional parameters (the reference has no setter for CUSTOM,
 * src/d3_buffer.cu:283), per-atom diagnostics, distinct status codes (the reference uses 1 for
 * every failure, src/d3.cu:134,224), batched evaluation of many cells, device-resident timing
 * entry points, and parameter-table selection.  Units are those of d3.h (Angstrom in, cutoffs
 * in bohr, eV / eV/Angstrom / eV/Angstrom^3 out) unless stated.
 */
#ifndef D3X_H
#define D3X_H

#include <stdint.h>

#include "d3.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Arithmetic of the pair and triplet kernels.  FP64: plain double everywhere.  FP32: float
 * pair/triplet math with compensated (two-sum) per-atom accumulation; per-atom results, the
 * total energy and the virial are reduced in double. */
typedef enum { D3X_FP64 = 0, D3X_FP32 = 1 } d3x_precision_t;

/* Status bits returned by d3x_* calls and d3x_last_status (0 = success). */
#define D3X_ERR_GENERIC 1u   /* any failure (also what compute_..._status returns)        */
#define D3X_ERR_OVERFLOW 2u  /* a neighbour-list capacity was exceeded                    */
#define D3X_ERR_INPUT 4u     /* bad argument: element, length, cell, cutoff              */
#define D3X_ERR_DEVICE 8u    /* HIP runtime error                                        */
#define D3X_ERR_STATE 16u    /* call order: atoms or cell not set                        */

/* Handle with every option.  atm_cutoff <= 0 disables the three-body term.  damping: 0 zero,
 * 1 BJ; functional: d3_types.h values; precision: d3x_precision_t. */
d3_handle_t *d3x_init(const uint16_t *elements, uint64_t num_elements, uint64_t max_length, double cutoff,
                      double cn_cutoff, double atm_cutoff, int damping, int functional, int precision);

/* params = {s6, s8_zero, s8_bj, sr6, sr8, a1, a2} (the reference's functional_params_t order,
 * src/d3_types.h:36-44).  Applies to the handle whatever functional it was created with. */
int d3x_set_functional_params(d3_handle_t *handle, const double params[7]);

/* FP64 variants of set_atoms / set_cell / compute.  coords: length x 3 Angstrom. */
int d3x_set_atoms_f64(d3_handle_t *handle, const double *coords, const uint16_t *elements, uint64_t length);
int d3x_set_cell_f64(d3_handle_t *handle, const double cell[9]);
int d3x_compute_f64(d3_handle_t *handle, double *energy, double *forces, double *stress);

/* Per-atom results of the last compute, caller's atom order; any pointer may be NULL.
 * cn: coordination numbers; dedcn: dE+/dCN in Hartree (two-body + three-body, the input of the
 * chain-rule pass); atom_energy: eV. */
int d3x_get_atom_data(d3_handle_t *handle, double *cn, double *dedcn, double *atom_energy);

/* Device-resident evaluation for timing: runs the whole pipeline on the handle's stream from
 * the coordinates already on the device and leaves the results on the device (no host
 * transfer, no synchronisation).  d3x_fetch_results copies them out (synchronising). */
int d3x_compute_device(d3_handle_t *handle);
int d3x_fetch_results(d3_handle_t *handle, double *energy, double *forces, double *stress);

/* Work counts of the last evaluation, 3 per cell: two-body pairs within the cutoff (each
 * unique pair counted twice, once per atom), CN pairs (same convention), and three-body
 * triangles with every edge within the ATM cutoff (each counted once). */
int d3x_get_counts(d3_handle_t *handle, double *counts);

/* Per-phase GPU timing with HIP events on the handle's stream.  d3x_set_timing(h, 1) enables
 * and resets; d3x_phase_times fills ms[6] with the summed milliseconds of the phases bin sort,
 * CN, two-body, three-body, chain rule, reduction over the timed evaluations and returns their
 * number (negative on error).  Timing synchronises once per evaluation. */
int d3x_set_timing(d3_handle_t *handle, int enable);
int d3x_phase_times(d3_handle_t *handle, double ms[6]);

/* hipGraph mode (enable != 0): the first evaluation captures the whole launch sequence (bin
 * sort, CN, two-body, three-body, chain rule, reductions) into a graph on the handle's stream and
 * later evaluations replay it; any change of atom count, cell geometry, parameters or scratch
 * sizes re-captures.  D3_GRAPH=1 in the environment enables it for handles made by the
 * reference-compatible init functions. */
int d3x_set_graph(d3_handle_t *handle, int enable);

/* Tuning parameters: bin width of the cell lists (bohr; default half the CN cutoff, 3/8 of it on
 * cells less than 6 CN cutoffs across) and the
 * number of persistent blocks of the three-body kernel (default: as many as the device holds at
 * once).  Values <= 0 keep the current setting.  d3x_autotune times full evaluations of the
 * handle's current input over a small candidate set -- bin widths, the three-body kernel's
 * variants (k-side sums in LDS or not, damping rows staged or not) and block counts -- (atoms and
 * cell must be set), keeps the fastest, stores it in the JSON cache at cache_path (NULL:
 * $D3_TUNE_CACHE, if set) keyed by precision, cutoffs, density bucket and element count, and
 * returns the best time in ms (negative on error); with D3_TUNE_LOG=1 in the environment it
 * prints one line per candidate on stderr.  Handles consult $D3_TUNE_CACHE for their input's key
 * unless tuning was set explicitly.  Results do not depend on these parameters beyond summation
 * order.  d3x_get_atm_variant: the three-body variant in use (bit 0: k-side sums in LDS, bit 1:
 * damping rows), chosen by occupancy unless the autotuner or the cache set it; -1 on error. */
int d3x_set_tuning(d3_handle_t *handle, double bin_width_bohr, int atm_blocks);
int d3x_get_tuning(d3_handle_t *handle, double *bin_width_bohr, int *atm_blocks);
int d3x_get_atm_variant(d3_handle_t *handle);

/* C6 interpolation of the handle: 0 = through per-atom Gaussian weights of per-element CN
 * references (w_A^T C6ref_AB w_B, the D3 tables), 1 = per pair over the pair's own entries of
 * c6ab[A][B][a][b] (src/d3_kernel.cu:709-800), used when the loaded table's CN references do not
 * factorise per element or when D3_C6_PAIRWISE=1 is in the environment at init.  Both give the
 * reference's C6; the per-pair form costs nref^2 exponentials per pair and per triangle edge, and
 * the three-body network (include/d3x_sur.h) needs the factorised form.  -1 on error.  No
 * counterpart in the reference, which always interpolates per pair. */
int d3x_get_c6_pairwise(d3_handle_t *handle);

/* Smooth CN cutoff (the reference's SMOOTH_CUTOFF_CN build option, src/d3_kernel.cu:550-561,
 * off in its default build): on != 0 multiplies every CN term by tanh^3(rc_CN - r) (bohr), with
 * its derivative in the chain-rule forces and stress.  D3_SMOOTH_CN=1 in the environment sets it
 * for handles made by the reference-compatible init functions. */
int d3x_set_cn_smooth(d3_handle_t *handle, int on);

/* Neighbour state across evaluations, for MD loops that call set_atoms with slightly moved atoms
 * (no counterpart in the reference, which sorts its cell list again at every compute,
 * src/d3.cu:141-143).  skin_bohr > 0: every stencil covers its cutoff plus the skin, and an
 * evaluation after set_atoms with the same atoms (count and elements) in the same cell keeps the
 * bin order and the three-body forward-neighbour lists of the last rebuild while no atom has moved
 * more than skin_bohr / 2 from its position then (compared on the host at set_atoms); otherwise it
 * rebuilds them.  Results equal those of a rebuild up to summation order.  0 (default) rebuilds
 * at every evaluation.  Batches always rebuild.  D3_SKIN=<bohr> in the environment sets it for
 * handles made by the reference-compatible init functions.
 * d3x_set_neighbour_lists: -1 (default) = three-body forward lists when a skin is set or centres
 * are split into more three-body work items than four per resident block, 0 = never (phase 1
 * walks the stencil), 1 = always; lists need atoms x capacity x 8 bytes and are not used beyond
 * 4 GiB.
 * d3x_neighbour_stats: evaluations so far that rebuilt / reused the neighbour state. */
int d3x_set_skin(d3_handle_t *handle, double skin_bohr);
int d3x_set_neighbour_lists(d3_handle_t *handle, int mode);
int d3x_neighbour_stats(d3_handle_t *handle, int *rebuilds, int *reuses);

/* Forward-neighbour capacity per centre of the three-body scratch (0 before the first
 * evaluation).  A centre with more forward neighbours within the ATM cutoff makes the
 * evaluation grow it (x4) and run again; the result is unaffected. */
int d3x_get_atm_capacity(d3_handle_t *handle);
double d3x_autotune(d3_handle_t *handle, const char *cache_path);

/* The HIP stream (hipStream_t) the handle launches on; d3x_set_stream replaces it (NULL =
 * the handle's own stream). */
void *d3x_get_stream(d3_handle_t *handle);
int d3x_set_stream(d3_handle_t *handle, void *stream);

/* Batched evaluation of n_struct independent periodic cells in one pass of launches.
 * atom_counts[s] atoms each, concatenated coords (Angstrom) and elements, cells (n_struct x 9).
 * Outputs: energies[n_struct], forces (concatenated, atoms x 3), stresses (n_struct x 9).
 * The total atom count must not exceed the handle's max_length. */
int d3x_compute_batch_f64(d3_handle_t *handle, uint64_t n_struct, const uint64_t *atom_counts,
                          const double *coords, const uint16_t *elements, const double *cells, double *energies,
                          double *forces, double *stresses);

/* Detailed status of the handle's last call, and the last error message of this thread. */
uint32_t d3x_last_status(d3_handle_t *handle);
const char *d3x_last_error(void);

/* Parameter tables (process-wide; handles copy what they need at init).
 * d3x_table_load: NULL, "" or "synthetic" = built-in synthetic 95 x 5 set (default),
 * "synthetic-extended" = built-in synthetic 103 x 7 set, anything else = path of a table file
 * (format in lasp-d3_amd/csrc/d3_tables.hpp).  The synthetic sets have the reference's layout
 * and invariants but NON-physical values. */
int d3x_table_load(const char *path);
const char *d3x_table_name(void);
int d3x_table_shape(int *nel, int *nref, int *zoff, int *zmax);
int d3x_table_export(float *c6ab, float *r0ab, float *rcov, float *r2r4);

#ifdef __cplusplus
}
#endif

#endif /* D3X_H */
