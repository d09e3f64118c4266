/* Public types of the D3 dispersion C API.
 *
 * Same names and enum values as the reference's src/d3_types.h:7-34 so that callers compiled
 * against the reference's d3.h keep working: real_t is float, damping and functional selectors
 * are C enums (int-sized).  The reference's device-side structs (atom_t, device_data_t) are
 * internal there and have no counterpart here.
 */
#ifndef D3_TYPES_H
#define D3_TYPES_H

#include <stdint.h>

typedef float real_t; /* src/d3_types.h:7 */

typedef enum {
    ZERO_DAMPING = 0, /* original DFT-D3 damping */
    BJ_DAMPING = 1    /* Becke-Johnson damping */
} damping_type_t;     /* src/d3_types.h:16-19 */

typedef enum {
    PBE = 0,
    PBE0 = 1,
    B3LYP = 2,
    BLYP = 3,
    BP86 = 4,
    REVPBE = 5,
    CUSTOM = 99 /* parameters set with d3x_set_functional_params (all 1.0 until set) */
} functional_t;  /* src/d3_types.h:21-29 */

#endif /* D3_TYPES_H */
