// Definitions shared by the host engine (d3_host.cpp) and the HIP kernels (d3_*.hip).
//
// HBM layout (one handle = one batch of periodic cells):
//   inputs (caller order): pos_in[3N] double bohr, type_in[N] u16, sys_in[N] i32
//   bins: every cell is cut into nb[0] x nb[1] x nb[2] parallelepiped bins; the bins of all
//         cells are numbered consecutively (SysDev::bin_base); bin_start[nbins+1] i32
//   sorted atoms (bin order, ascending caller index inside a bin, cells contiguous):
//         lx/ly/lz[N] T   position relative to the centre of the atom's bin (bohr)
//         type_s, orig_s (caller index), bin_s, sys_s
//   per sorted atom: cn[N], w/dw[NREF][N] T (normalised Gaussian CN weights and d/dCN),
//         e[N], g[N], f[3N], vir[6N] double (energy E+, dE+/dCN, F = grad E+, virial / V)
//   stencils: for each cell and each cutoff (CN, two-body, ATM) the list of bin offsets d whose
//         bins can hold a pair within the cutoff, with D = -(d/nb) . H, so that for atoms i, j in
//         bins b and b+d (wrapped) the image vector is r_i - r_j - T = l_i - l_j + D.
#pragma once
#include <hip/hip_runtime_api.h>

#include <cstdint>

namespace d3 {

constexpr int MAX_NREF = 7;
constexpr double BOHR_A = 0.52917726;            // src/d3.cu:50
constexpr double HARTREE_EV = 27.211396641308;   // src/d3.cu:204
constexpr double K1 = 16.0;                      // src/d3_kernel.cuh:17
constexpr double K3 = 4.0;                       // src/d3_kernel.cuh:19

enum Cut { CUT_CN = 0, CUT_2B = 1, CUT_ATM = 2, CUT_SUR = 3 };  // SUR: the three-body stand-in's features
constexpr int NCUT = 4;

struct StEntry {
    int dx, dy, dz, pad;
};

// One forward three-body neighbour of a centre: sorted atom index and the stencil entry (bin
// offset and image vector) it was found through.
struct FwdEntry {
    int j, s;
};

struct SysDev {
    double H[9];     // rows = lattice vectors (bohr)
    double Hinv[9];  // H^-1
    double inv_vol;  // 1 / det(H) (signed, src/d3_kernel.cu:90-96)
    int nb[3];       // bins per axis
    int bin_base;    // global index of this cell's bin (0,0,0)
    int atom_base;   // first sorted (and caller) atom index of this cell
    int n_atoms;
    int st_off[NCUT];  // stencil ranges per cutoff
    int st_cnt[NCUT];
    // Bin shape, for skipping stencil bins that no atom within a cutoff of the centre atom can
    // occupy: unit normals of the lattice planes (nrm[3 k .. 3 k + 2] for axis k), half the bin
    // height along each, half the longest bin diagonal, and pad = how far an atom may sit outside
    // its bin (half the neighbour skin) plus a rounding margin.  ortho: the cell is orthorhombic
    // (normals orthogonal, the per-axis gaps add in quadrature).
    double nrm[9];
    double hh[3];
    double hdiag;
    double pad;
    int ortho;
};

// Device pointers and scalars of one evaluation; passed by value to every kernel.
template <class T>
struct Params {
    int n;      // atoms in the batch
    int ne;     // element types of the handle
    int nref;   // references per element (<= MAX_NREF)
    int nsys;
    int damping;
    T rc_cn2, rc_22, rc_32;  // squared cutoffs (bohr^2)
    T rc_cn;                 // CN cutoff (bohr), for the smooth cutoff
    int cn_smooth;           // CN terms times tanh^3(rc_cn - r) (the reference's SMOOTH_CUTOFF_CN)
    T s6, s8, p1, p2;        // s6, s8 and (sr6, sr8) or (a1, a2)
    // structure
    const SysDev *sys;
    const StEntry *st_d;  // stencil bin offsets
    const T *st_D;      // stencil image vectors, 4 per entry
    const int *bin_start;
    // inputs
    const double *pos_in;
    const uint16_t *type_in;
    const int *sys_in;
    // binning scratch
    int *key_in, *bin_count, *cursor, *perm, *perm2;
    double *wrapped;  // 3N
    // sorted atoms
    T *lx, *ly, *lz;
    uint16_t *type_s;
    int *orig_s, *bin_s, *sys_s;
    // per-atom results (sorted order)
    T *cn, *w, *dw;
    double *e, *g, *f, *vir;
    float *npair;  // 3 per atom: two-body, CN, ATM pair counts (for the work model)
    // type tables (T): c6[ne][ne][nref][nref], cnref[ne][nref] (-1 invalid), rcov[ne],
    // q4[ne][ne] = r2r4_A r2r4_B, rdamp[ne][ne] (zero: r0ab; BJ: a1 sqrt(3 q4) + a2),
    // r0[ne][ne] (ATM damping radii)
    const T *c6, *cnref, *rcov, *q4, *rdamp, *r0;
    // Tables whose CN references do not factorise per element: c6g[ne][ne][nref][nref][3] =
    // {C6ref, CN_A ref, CN_B ref} of each pair entry, and the pair kernels interpolate per pair
    // (c6_pairwise) instead of through the per-atom weights w, dw.
    const T *c6g;
    int c6_pairwise;
    // ATM: per-block scratch for one centre's forward neighbours (capacity atm_cap):
    //   atm_i32: j, type (2 x cap); atm_T: record fields (ATM_NF x cap: offset, CN log-derivatives, sqrt C6)
    //   and per tile of 64 slots its bounding box (8 values: min, max per axis, 2 unused);
    //   atm_acc: per-neighbour accumulators (6 x cap, double)
    // and per-atom results (sorted order): fa[3N], ga[N], eas[N] filled with atomics (every
    // triangle's gradient, dE/dCN and e/3 for each vertex), eat[N] = sum of e over the
    // triangles anchored at each centre (deterministic total).  atm_next: the next centre to
    // take (zeroed before every launch).
    int atm_cap;
    int atm_var;  // three-body kernel variant: ATM_KL | ATM_DR bits (d3_atm.hip)
    // C6 rows of the j tile per type B of k: atm_roff[B] = sum of the reference counts of the
    // types before B (a type's count runs to its last valid reference), atm_roff[ne] = all of
    // them, atm_roff[ne + 1] = atm_yrows = that plus padding for the last type's NR-row window
    const int *atm_roff;
    int atm_yrows;
    int *atm_i32;
    T *atm_T;
    double *atm_acc;
    double *fa, *ga, *eas, *eat;
    int *atm_next;
    // Forward-neighbour lists (null: phase 1 walks the three-body stencil).  atm_fl[i * atm_fl_cap
    // + p], p < atm_fl_n[i]: centre i's forward neighbours within sqrt(rc_fl2) (the three-body
    // cutoff plus the skin) in stencil-walk order, built by launch_atm_lists; phase 1 then tests
    // only these.  A centre with more than atm_fl_cap sets status bit 2 like a scratch overflow.
    FwdEntry *atm_fl;
    int *atm_fl_n;
    int atm_fl_cap;
    T rc_fl2;
    // Lattice shifts of the wrap per caller atom (3 doubles), written by the bin sort when the
    // handle keeps its neighbour state (null otherwise) and used to gather moved positions into
    // the kept bin order.
    double *shift;
    // Work items per centre (1..8): with atm_chunks > 1 item t is centre t / atm_chunks and a
    // contiguous range of its k tiles; each item writes its share of the centre's energy to
    // atm_eatc[chunk * n + centre] (summed in chunk order after the launch) and adds its virial
    // and triangle count atomically.  More items than centres fill the GPU on small cells.
    int atm_chunks;
    double *atm_eatc;
    // Two-body waves per centre (1..8): with tb_parts > 1 wave w handles centre w / tb_parts over
    // a contiguous share of its stencil bins and writes its sums to tb_part[(part n + centre) *
    // TB_PART_W] (energy, dE/dCN, force, virial, pair count), added in part order afterwards.
    int tb_parts;
    double *tb_part;
    // three-body stand-in (d3_sur.hip): class of each element type, features [N][64], the
    // network's dE/dx [N][64] and energies (eV), class per sorted atom, fp32 parameters and
    // packed bf16 weights
    const int *type_cls;
    float *feat, *sur_G, *sur_E;
    int *sur_cls;
    // per atom (sorted order): q = C6_ii^(1/4) at its CN and dq/dCN (sur_q[i], sur_q[n + i]),
    // the neighbour-weight factors of the features; the tangent pass's CN push-forward sur_dcn[i]
    float *sur_q, *sur_dcn;
    const float *sur_P;
    const void *sur_PK;
    unsigned *status;
};

constexpr int ATM_NF = 6;  // record fields per ATM neighbour (offset, CN log-derivatives, sqrt C6)
constexpr int TB_PART_W = 12;  // doubles per two-body part sum: E, dE/dCN, F[3], virial[6], pairs
// doubles of ATM accumulator scratch per block: 6 per neighbour slot
__host__ __device__ constexpr size_t atm_acc_stride(int cap) { return 6 * (size_t)cap; }
// values of ATM record scratch per block (cap a multiple of 64): ATM_NF per neighbour slot, then
// 8 per tile of 64 slots
__host__ __device__ constexpr size_t atm_T_stride(int cap) { return ATM_NF * (size_t)cap + 8 * (size_t)(cap / 64); }

// Host-side launchers (d3_kernels.hip).
template <class T>
struct Launch {
    static void bin_sort(const Params<T> &P, int nbins, int *bin_start_mut, hipStream_t s);
    // Moved positions into the bin order of the last sort (wrap shifts and bins kept).
    static void regather(const Params<T> &P, hipStream_t s);
    static void cn(const Params<T> &P, hipStream_t s);
    static void twobody(const Params<T> &P, hipStream_t s);
    static void chain(const Params<T> &P, hipStream_t s);
};

// Three-body term (d3_atm.hip): zeroes the ATM per-atom arrays, runs the triangle kernel on
// nblocks persistent one-wave blocks (when P.rc_32 > 0) and merges forces and dE/dCN into
// P.f / P.g.  atm_resident_blocks: how many of those blocks the current device holds at once.
// The kernel variant (P.atm_var) holds two bits: ATM_KL keeps the k side's tile sums in LDS,
// ATM_DR stages the per-type damping term with the C6 rows; atm_choose_variant picks them for a
// handle's element count, reference count and precision.
enum { ATM_KL = 1, ATM_DR = 2 };
template <class T>
void launch_atm(const Params<T> &P, int nblocks, hipStream_t s);
// Forward-neighbour lists of every centre (P.atm_fl, P.atm_fl_n), one wave per centre.
template <class T>
void launch_atm_lists(const Params<T> &P, hipStream_t s);
size_t atm_lds_bytes(int ne, int nref, int yrows, size_t tsize, int var);
int atm_choose_variant(int ne, int nref, int yrows, size_t tsize);
int atm_resident_blocks(int ne, int nref, int yrows, size_t tsize, int var);
// P.f += fa, P.g += ga (what launch_atm ends with).
template <class T>
void launch_atm_merge(const Params<T> &P, hipStream_t s);

// Three-body stand-in (d3_sur.hip): features only (training sets), or features, network and
// backward writing fa / ga / eas / eat and the virial in place of the ATM kernel, then the merge.
template <class T>
void launch_sur_features(const Params<T> &P, hipStream_t s);
template <class T>
void launch_surrogate(const Params<T> &P, hipStream_t s);
// u = J dir through the features (dir: per atom in sorted order, 3 floats; U: 64 floats per atom)
template <class T>
void launch_sur_tangent(const Params<T> &P, const float *dir, float *U, hipStream_t s);

// Reduction of per-atom energy / virial / pair counts into per-cell totals (deterministic),
// and scatter of per-atom results to caller order with unit conversion.
void launch_reduce(const double *e, const double *eat, const double *vir, const float *npair, const int *chunk_begin,
                   const int *chunk_end, const int *chunk_sys, int nchunks, const int *sys_chunk_off, int nsys,
                   double *partial, double *out_sys, hipStream_t s);
template <class T>
void launch_finalize(const Params<T> &P, double *out_f, double *out_e, double *out_cn, double *out_g,
                     hipStream_t s);

constexpr int RED_W = 10;  // per-cell reduced values: E, vir[6], counts[3]

}  // namespace d3
