/* D3 dispersion C API — drop-in for the reference's src/d3.h.
 *
 * Every entry point below replaces the reference function of the same name and signature
 * (src/d3.h:36-128, implemented in src/d3.cu:14-252) and keeps its units and conventions:
 *   - coordinates and cell in Angstrom; the cell is row-major with rows = lattice vectors;
 *   - cutoffs in bohr (not converted);
 *   - energy in eV (total, negative for attraction), forces in eV/Angstrom in the caller's atom
 *     order, stress 3x3 row-major in eV/Angstrom^3 = (1/V) dE/d(strain) with V the signed cell
 *     determinant; outputs are overwritten, not accumulated.
 * The handle owns device buffers on the current HIP device.  Not thread-safe: call one handle
 * from one thread.  Errors are printed to stderr; init functions return NULL, the status
 * function returns non-zero (1), as the reference does.  Unlike the reference, set_atoms with
 * length > max_length reports an error and leaves the handle unchanged instead of exit()ing.
 */
#ifndef D3_H
#define D3_H

#include <stdint.h>

#include "d3_types.h"

typedef void d3_handle_t; /* src/d3.h:9 */

#ifdef __cplusplus
extern "C" {
#endif

/* One-shot evaluation; replaces src/d3.h:36-38 (src/d3.cu:236-252). */
void compute_dispersion_energy(real_t atoms[][3], uint16_t *elements, uint64_t num_atoms, real_t cell[3][3],
                               real_t cutoff_radius, real_t CN_cutoff_radius, damping_type_t damping_type,
                               functional_t functional_type, real_t *energy, real_t *force, real_t *stress);

/* Handle for repeated evaluation; `elements` lists every element that may appear.
 * Replaces src/d3.h:59-61 (src/d3.cu:14-26). */
d3_handle_t *init_d3_handle(uint16_t *elements, uint64_t num_elements, uint64_t max_length, real_t cutoff_radius,
                            real_t coordination_number_cutoff, damping_type_t damping_type,
                            functional_t functional_type);

/* As init_d3_handle, with the Axilrod-Teller-Muto three-body term enabled within atm_cutoff
 * (bohr).  Replaces src/d3.h:63-65 (src/d3.cu:28-40). */
d3_handle_t *init_d3_handle_with_atm(uint16_t *elements, uint64_t length_elements, uint64_t max_length,
                                     real_t cutoff_radius, real_t coordination_number_cutoff, real_t atm_cutoff,
                                     damping_type_t damping_type, functional_t functional_type);

/* coords: length x 3 Angstrom; elements: atomic numbers.  Replaces src/d3.h:81 (src/d3.cu:42-70). */
void set_atoms(d3_handle_t *handle, real_t *coords, uint16_t *elements, uint64_t length);

/* cell: rows = lattice vectors, Angstrom.  Replaces src/d3.h:90 (src/d3.cu:72-93). */
void set_cell(d3_handle_t *handle, real_t cell[3][3]);

/* Replaces src/d3.h:97 (src/d3.cu:110-128). */
void free_d3_handle(d3_handle_t *handle);

/* Clears intermediate results.  Replaces src/d3.h:106 (src/d3.cu:95-108). */
void clear_d3_handle(d3_handle_t *handle);

/* Replaces src/d3.h:115 (src/d3.cu:228-234). */
void compute_dispersion_energy_from_handle(d3_handle_t *handle, real_t *energy, real_t *force, real_t *stress);

/* Returns 0 on success, 1 on failure (NULL handle, no atoms/cell set, neighbour-list
 * capacity exceeded, device error); d3x_last_status() gives the detailed code.
 * Replaces src/d3.h:127-128 (src/d3.cu:130-226). */
uint16_t compute_dispersion_energy_from_handle_status(d3_handle_t *handle, real_t *energy, real_t *force,
                                                      real_t *stress);

#ifdef __cplusplus
}
#endif

#endif /* D3_H */
