// D3 parameter tables: layout, the built-in synthetic set, and the file loader.
//
// Layout follows the reference's compile-time tables (src/constants_include.h:9-32):
//   c6ab[NEL][NEL][NREF][NREF][3]  = {C6ref, CNref of the first element's reference,
//                                     CNref of the second element's reference}, -1 = missing
//   r0ab[NEL][NEL], rcov[NEL] (already scaled by k2 = 4/3), r2r4[NEL]
// "original": NEL = 95, NREF = 5, slot = Z        (src/d3_buffer.cu:17 IDX_ELEM(x) = x)
// "extended": NEL = 103, NREF = 7, slot = Z - 1   (src/d3_buffer.cu:13 IDX_ELEM(x) = x - 1)
//
// The reference's table blobs are not available, so the default table is a deterministic
// synthetic set with the same layout and invariants (symmetric C6, -1 sentinels, CN
// references that depend only on (element, reference), plausible magnitudes).  Its values are
// NOT physical.  A real table can be loaded from a binary file (see d3_table_load).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace d3 {

struct Table {
    int nel = 0;       // number of element slots
    int nref = 0;      // references per element
    int zoff = 0;      // slot = Z - zoff
    int zmax = 0;      // largest supported atomic number
    std::string name;  // "synthetic-original", "synthetic-extended" or the file path
    std::vector<float> c6ab, r0ab, rcov, r2r4;

    int slot(int Z) const { return Z - zoff; }
    bool supports(int Z) const { return Z >= 1 && Z <= zmax && slot(Z) >= 0 && slot(Z) < nel; }
    size_t c6_index(int sa, int sb, int a, int b) const {
        return ((((size_t)sa * nel + sb) * nref + a) * nref + b) * 3;
    }
};

// Builds the synthetic table. extended = false gives the 95 x 5 layout, true the 103 x 7 one.
Table make_synthetic_table(bool extended);

// Loads a table file written by tables/pack_table.py: magic "D3TAB001", int32 nel, nref, zoff,
// zmax, then float32 c6ab, r0ab, rcov, r2r4 in the layout above.  Throws on any mismatch.
Table load_table_file(const std::string &path);

// The process-wide active table (synthetic "original" unless replaced).
const Table &active_table();
void set_active_table(Table t);

}  // namespace d3
