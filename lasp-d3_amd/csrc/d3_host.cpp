// Host side of a D3 handle (declarations in d3_host.hpp).
#include "d3_host.hpp"

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>

#include "../../include/d3x.h"
#include "../../include/d3x_sur.h"
#include "d3_sur.hpp"
#include "d3_tables.hpp"

namespace d3 {

namespace {
thread_local std::string g_err;

#define HIPCHK(x)                                                                                         \
    do {                                                                                                  \
        hipError_t e_ = (x);                                                                              \
        if (e_ != hipSuccess)                                                                             \
            throw Error{D3X_ERR_GENERIC | D3X_ERR_DEVICE, std::string(#x) + ": " + hipGetErrorString(e_)}; \
    } while (0)

void *dalloc(size_t bytes) {
    void *p = nullptr;
    HIPCHK(hipMalloc(&p, bytes ? bytes : 16));
    return p;
}
void dfree(void *&p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}
void inv3(const double m[9], double o[9]) {
    const double det = m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
                       m[2] * (m[3] * m[7] - m[4] * m[6]);
    const double id = 1.0 / det;
    o[0] = (m[4] * m[8] - m[5] * m[7]) * id;
    o[1] = (m[2] * m[7] - m[1] * m[8]) * id;
    o[2] = (m[1] * m[5] - m[2] * m[4]) * id;
    o[3] = (m[5] * m[6] - m[3] * m[8]) * id;
    o[4] = (m[0] * m[8] - m[2] * m[6]) * id;
    o[5] = (m[2] * m[3] - m[0] * m[5]) * id;
    o[6] = (m[3] * m[7] - m[4] * m[6]) * id;
    o[7] = (m[1] * m[6] - m[0] * m[7]) * id;
    o[8] = (m[0] * m[4] - m[1] * m[3]) * id;
}
double det3(const double m[9]) {
    return m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) + m[2] * (m[3] * m[7] - m[4] * m[6]);
}
[[noreturn]] void fail(uint32_t code, const std::string &msg) { throw Error{code | D3X_ERR_GENERIC, msg}; }

constexpr int CHUNK = 2048;  // atoms per first-level reduction chunk

// Bin offsets d of one cell for cutoff rc: every d for which some pair of points in bins b and
// b + d can be within rc.  Lower bound of the distance: the larger of the slab bound
// max_k (|d_k| - 1) h_k / n_k (h_k = perpendicular height of the cell) and |D| - diag, with
// D = -(d / n) . H the bin-centre offset and diag the longest bin diagonal.
void make_stencil(const double H[9], const int nb[3], double rc, std::vector<StEntry> &out_d,
                  std::vector<double> &out_D) {
    double Hi[9];
    inv3(H, Hi);
    double hb[3], e[3][3];
    for (int k = 0; k < 3; ++k) {
        const double nk = std::sqrt(Hi[0 * 3 + k] * Hi[0 * 3 + k] + Hi[1 * 3 + k] * Hi[1 * 3 + k] + Hi[2 * 3 + k] * Hi[2 * 3 + k]);
        hb[k] = (1.0 / nk) / nb[k];
        for (int c = 0; c < 3; ++c) e[k][c] = H[k * 3 + c] / nb[k];
    }
    double diag = 0;
    for (int s1 = -1; s1 <= 1; s1 += 2)
        for (int s2 = -1; s2 <= 1; s2 += 2) {
            double v[3];
            for (int c = 0; c < 3; ++c) v[c] = e[0][c] + s1 * e[1][c] + s2 * e[2][c];
            diag = std::max(diag, std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]));
        }
    int R[3];
    for (int k = 0; k < 3; ++k) R[k] = (int)std::ceil(rc / hb[k]) + 1;
    const double lim = rc * (1.0 + 1e-10) + 1e-8;
    for (int dz = -R[2]; dz <= R[2]; ++dz)
        for (int dy = -R[1]; dy <= R[1]; ++dy)
            for (int dx = -R[0]; dx <= R[0]; ++dx) {
                const int d[3] = {dx, dy, dz};
                double D[3];
                for (int c = 0; c < 3; ++c) D[c] = -(dx * e[0][c] + dy * e[1][c] + dz * e[2][c]);
                double slab = 0;
                for (int k = 0; k < 3; ++k) slab = std::max(slab, std::max(0, std::abs(d[k]) - 1) * hb[k]);
                const double sph = std::sqrt(D[0] * D[0] + D[1] * D[1] + D[2] * D[2]) - diag;
                if (std::max(slab, sph) <= lim) {
                    out_d.push_back(StEntry{dx, dy, dz, 0});
                    for (int c = 0; c < 3; ++c) out_D.push_back(D[c]);
                    out_D.push_back(0.0);
                }
            }
}

}  // namespace

void set_last_error(const std::string &msg) { g_err = msg; }
const char *last_error() { return g_err.c_str(); }

bool functional_params(int functional, double out[7]) {
    // {s6, s8_zero, s8_bj, sr6, sr8, a1, a2}: the reference's FUNCTIONAL_PARAMS
    // (src/d3_types.h:46-54), kept as the same float values.
    static const float tab[7][7] = {
        {1.0f, 0.722f, 0.7875f, 1.217f, 1.0f, 0.4289f, 4.4407f},  // PBE
        {1.0f, 0.928f, 1.2177f, 1.287f, 1.0f, 0.4145f, 4.8593f},  // PBE0
        {1.0f, 1.703f, 1.9889f, 1.261f, 1.0f, 0.3981f, 4.4211f},  // B3LYP
        {1.0f, 1.682f, 2.6996f, 1.094f, 1.0f, 0.4298f, 4.2359f},  // BLYP
        {1.0f, 1.683f, 3.2822f, 1.139f, 1.0f, 0.3946f, 4.8516f},  // BP86
        {1.0f, 1.010f, 2.3550f, 0.923f, 1.0f, 0.5238f, 3.5016f},  // REVPBE
        {1.0f, 1.0f, 1.0f, 1.0f, 1.0f, 1.0f, 1.0f},               // CUSTOM (until set)
    };
    int row;
    if (functional >= 0 && functional <= 5)
        row = functional;
    else if (functional == 99)
        row = 6;  // the reference indexes FUNCTIONAL_PARAMS[99] out of bounds (src/d3_buffer.cu:283)
    else
        return false;
    for (int k = 0; k < 7; ++k) out[k] = (double)tab[row][k];
    return true;
}

Engine::Engine(const uint16_t *elements, uint64_t num_elements, uint64_t max_length, double cutoff,
               double cn_cutoff, double atm_cutoff, int damping, int functional, int precision)
    : rc2_(cutoff), rccn_(cn_cutoff), rc3_(atm_cutoff > 0 ? atm_cutoff : 0.0), damping_(damping),
      functional_(functional), precision_(precision) {
    if (!elements && num_elements) fail(D3X_ERR_INPUT, "elements is NULL");
    if (max_length == 0 || max_length > (uint64_t)INT32_MAX / 8) fail(D3X_ERR_INPUT, "max_length out of range");
    if (!(cutoff > 0) || !(cn_cutoff > 0) || !std::isfinite(cutoff) || !std::isfinite(cn_cutoff) ||
        !std::isfinite(rc3_))
        fail(D3X_ERR_INPUT, "cutoffs must be positive and finite");
    if (damping != 0 && damping != 1) fail(D3X_ERR_INPUT, "damping_type must be 0 (zero) or 1 (BJ)");
    if (!functional_params(functional, fp_)) fail(D3X_ERR_INPUT, "unknown functional_type");
    if (precision != D3X_FP64 && precision != D3X_FP32) fail(D3X_ERR_INPUT, "unknown precision");
    const Table &tab = active_table();
    table_name_ = tab.name;
    if (tab.nref != 5 && tab.nref != 7)  // the reference's two table layouts (src/constants_include.h:14-27)
        fail(D3X_ERR_INPUT, "table must have 5 or 7 references per element");
    for (uint64_t k = 0; k < num_elements; ++k) {
        if (!tab.supports(elements[k]))
            fail(D3X_ERR_INPUT, "element " + std::to_string(elements[k]) + " is not in the parameter table");
        elems_.push_back(elements[k]);
    }
    std::sort(elems_.begin(), elems_.end());
    elems_.erase(std::unique(elems_.begin(), elems_.end()), elems_.end());
    ne_ = (int)elems_.size();
    if (ne_ == 0) fail(D3X_ERR_INPUT, "no elements given");
    if (ne_ > 64) fail(D3X_ERR_INPUT, "at most 64 distinct elements per handle");
    // per-handle tables: C6ref rows and factorised CN references (src/d3_buffer.cu:192-277 gathers
    // the same entries per unique element)
    const int ne = ne_, tnr = tab.nref;
    std::vector<double> tcnref((size_t)ne * tnr, -1.0);
    for (int A = 0; A < ne; ++A)
        for (int a = 0; a < tnr; ++a)
            for (int B = 0; B < ne && tcnref[A * tnr + a] <= -1.0; ++B)
                for (int b = 0; b < tnr; ++b) {
                    const size_t ix = tab.c6_index(tab.slot(elems_[A]), tab.slot(elems_[B]), a, b);
                    if (tab.c6ab[ix + 1] > -1.0f && tab.c6ab[ix + 2] > -1.0f) {
                        tcnref[A * tnr + a] = tab.c6ab[ix + 1];
                        break;
                    }
                }
    // References no element of the handle has are dropped: their Gaussian weights are exactly
    // zero, so every sum over references is unchanged, and the kernels loop over fewer terms.
    nref_ = 1;
    for (int A = 0; A < ne; ++A)
        for (int a = 0; a < tnr; ++a)
            if (tcnref[A * tnr + a] > -1.0) nref_ = std::max(nref_, a + 1);
    const int nr = nref_;
    // three-body j-tile rows: each type keeps its references up to its last valid one, plus
    // padding so that the last type's window of nr rows stays inside the staged rows
    atm_roff_.assign(ne + 2, 0);
    for (int A = 0; A < ne; ++A) {
        int nt = 1;
        for (int a = 0; a < nr; ++a)
            if (tcnref[A * tnr + a] > -1.0) nt = a + 1;
        atm_roff_[A + 1] = atm_roff_[A] + nt;
    }
    atm_roff_[ne + 1] = atm_roff_[ne] + nr - (atm_roff_[ne] - atm_roff_[ne - 1]);
    yrows_ = atm_roff_[ne + 1];
    std::vector<double> c6((size_t)ne * ne * nr * nr, 0.0), cnref((size_t)ne * nr, -1.0);
    for (int A = 0; A < ne; ++A)
        for (int a = 0; a < nr; ++a) cnref[A * nr + a] = tcnref[A * tnr + a];
    // Factorisation: every entry (A, B, a, b) valid iff references a of A and b of B are, with CN
    // references equal to the per-element ones.  Otherwise C6 is interpolated per pair over the
    // pair's own entries, as src/d3_kernel.cu:709-800 does for every pair.
    const char *pw_env = std::getenv("D3_C6_PAIRWISE");
    c6_pairwise_ = pw_env && std::strcmp(pw_env, "1") == 0;
    int nr_pw = 1;  // references the pairwise table needs: the last valid index of any entry
    for (int A = 0; A < ne; ++A)
        for (int B = 0; B < ne; ++B)
            for (int a = 0; a < tnr; ++a)
                for (int b = 0; b < tnr; ++b) {
                    const size_t ix = tab.c6_index(tab.slot(elems_[A]), tab.slot(elems_[B]), a, b);
                    const bool valid = tab.c6ab[ix + 1] > -1.0f && tab.c6ab[ix + 2] > -1.0f;
                    const bool fvalid = tcnref[A * tnr + a] > -1.0 && tcnref[B * tnr + b] > -1.0;
                    if (valid) nr_pw = std::max(nr_pw, std::max(a, b) + 1);
                    if (valid != fvalid || (valid && ((double)tab.c6ab[ix + 1] != tcnref[A * tnr + a] ||
                                                      (double)tab.c6ab[ix + 2] != tcnref[B * tnr + b])))
                        c6_pairwise_ = true;
                    if (valid && fvalid && a < nr && b < nr) c6[(((size_t)A * ne + B) * nr + a) * nr + b] = tab.c6ab[ix + 0];
                }
    h_c6g_.clear();
    if (c6_pairwise_) {
        // the per-pair table over nr_pw references; the per-element rows stay for the CN weights
        // (computed, unused by the pairwise kernels); the three-body tile stages one row (CN_j)
        nref_ = std::max(nref_, nr_pw);
        const int np = nref_;
        h_c6g_.assign((size_t)3 * ne * ne * np * np, 0.0);
        for (int A = 0; A < ne; ++A)
            for (int B = 0; B < ne; ++B)
                for (int a = 0; a < np; ++a)
                    for (int b = 0; b < np; ++b) {
                        double *o = &h_c6g_[3 * ((((size_t)A * ne + B) * np + a) * np + b)];
                        o[1] = o[2] = -1.0;
                        if (a >= tnr || b >= tnr) continue;
                        const size_t ix = tab.c6_index(tab.slot(elems_[A]), tab.slot(elems_[B]), a, b);
                        if (tab.c6ab[ix + 1] > -1.0f && tab.c6ab[ix + 2] > -1.0f)
                            for (int q = 0; q < 3; ++q) o[q] = tab.c6ab[ix + q];
                    }
        if (np != nr) {  // per-element rows at the wider reference count
            std::vector<double> c6w((size_t)ne * ne * np * np, 0.0), cnw((size_t)ne * np, -1.0);
            for (int A = 0; A < ne; ++A)
                for (int a = 0; a < nr; ++a) cnw[A * np + a] = cnref[A * nr + a];
            for (int A = 0; A < ne; ++A)
                for (int B = 0; B < ne; ++B)
                    for (int a = 0; a < nr; ++a)
                        for (int b = 0; b < nr; ++b)
                            c6w[(((size_t)A * ne + B) * np + a) * np + b] = c6[(((size_t)A * ne + B) * nr + a) * nr + b];
            c6.swap(c6w);
            cnref.swap(cnw);
        }
        atm_roff_.assign(ne + 2, 0);
        atm_roff_[ne + 1] = 1;
        yrows_ = 1;
    }
    if (rc3_ > 0 && atm_lds_bytes(ne_, knr(), yrows_, precision == D3X_FP64 ? 8 : 4, 0) > 65536)
        fail(D3X_ERR_INPUT, "too many distinct elements for the three-body kernel's LDS tile");
    // layout of h_tab_: c6 | cnref | rcov | r2r4 | r0ab  (q4 / rdamp derived at upload)
    h_tab_.clear();
    h_tab_.insert(h_tab_.end(), c6.begin(), c6.end());
    h_tab_.insert(h_tab_.end(), cnref.begin(), cnref.end());
    for (int A = 0; A < ne; ++A) h_tab_.push_back(tab.rcov[tab.slot(elems_[A])]);
    for (int A = 0; A < ne; ++A) h_tab_.push_back(tab.r2r4[tab.slot(elems_[A])]);
    for (int A = 0; A < ne; ++A)
        for (int B = 0; B < ne; ++B) h_tab_.push_back(tab.r0ab[(size_t)tab.slot(elems_[A]) * tab.nel + tab.slot(elems_[B])]);

    max_len_ = max_length;
    // Bins about half the CN cutoff wide: the CN / chain-rule stencils then skip most corner
    // bins, and the two-body stencil's bins stay small against its cutoff.  A two-body cutoff
    // below the CN cutoff does not shrink them: the CN and chain-rule passes would then walk
    // thousands of near-empty bins per centre while the two-body pass saves next to nothing.
    bin_width_ = std::max(4.0, 0.5 * rccn_);
    if (const char *tc = std::getenv("D3_TUNE_CACHE")) tune_cache_ = tc;
    sys_.resize(1);
    HIPCHK(hipStreamCreate(&own_stream_));
    stream_ = own_stream_;
    alloc_atoms();
    d_status = dalloc(2 * sizeof(unsigned));  // status word | three-body centre counter
    HIPCHK(hipMemset(d_status, 0, 2 * sizeof(unsigned)));
    std::vector<int> cls(ne_);
    for (int A = 0; A < ne_; ++A) cls[A] = sur::element_class(elems_[A]);
    d_type_cls = dalloc(ne_ * sizeof(int));
    HIPCHK(hipMemcpy(d_type_cls, cls.data(), ne_ * sizeof(int), hipMemcpyHostToDevice));
    d_atm_roff = dalloc(atm_roff_.size() * sizeof(int));
    HIPCHK(hipMemcpy(d_atm_roff, atm_roff_.data(), atm_roff_.size() * sizeof(int), hipMemcpyHostToDevice));
}

Engine::~Engine() {
    if (own_stream_) (void)hipStreamSynchronize(own_stream_);
    if (gexec_) (void)hipGraphExecDestroy(gexec_);
    void **all[] = {&d_pos_in, &d_type_in, &d_sys_in, &d_key, &d_perm, &d_perm2, &d_wrapped, &d_lx, &d_ly, &d_lz,
                    &d_type_s, &d_orig_s, &d_bin_s, &d_sys_s, &d_cn, &d_w, &d_dw, &d_e, &d_g, &d_f, &d_vir,
                    &d_npair, &d_out_f, &d_out_e, &d_out_cn, &d_out_g, &d_bin_count, &d_bin_start, &d_cursor,
                    &d_sys, &d_st_d, &d_st_D, &d_chunk_b, &d_chunk_e, &d_chunk_s, &d_sys_chunk, &d_partial,
                    &d_out_sys, &d_tab, &d_status, &d_atm_i32, &d_atm_T, &d_atm_acc, &d_atm_pa, &d_type_cls,
                    &d_feat, &d_sur_G, &d_sur_E, &d_sur_cls, &d_sur_P, &d_sur_PK, &d_tan_dir, &d_tan_U, &d_atm_roff, &d_atm_eatc, &d_tb_part,
                    &d_fl, &d_fl_n, &d_shift, &d_sur_q, &d_sur_dcn};
    for (void **p : all) dfree(*p);
    for (auto &e : ev_)
        if (e) (void)hipEventDestroy(e);
    if (own_stream_) (void)hipStreamDestroy(own_stream_);
}

void Engine::alloc_atoms() {
    const size_t N = max_len_;
    const size_t ts = precision_ == D3X_FP64 ? sizeof(double) : sizeof(float);
    d_pos_in = dalloc(3 * N * sizeof(double));
    d_type_in = dalloc(N * sizeof(uint16_t));
    d_sys_in = dalloc(N * sizeof(int));
    d_key = dalloc(N * sizeof(int));
    d_perm = dalloc(N * sizeof(int));
    d_perm2 = dalloc(N * sizeof(int));
    d_wrapped = dalloc(3 * N * sizeof(double));
    d_lx = dalloc(N * ts);
    d_ly = dalloc(N * ts);
    d_lz = dalloc(N * ts);
    d_type_s = dalloc(N * sizeof(uint16_t));
    d_orig_s = dalloc(N * sizeof(int));
    d_bin_s = dalloc(N * sizeof(int));
    d_sys_s = dalloc(N * sizeof(int));
    d_cn = dalloc(N * ts);
    d_w = dalloc((size_t)nref_ * N * ts);
    d_dw = dalloc((size_t)nref_ * N * ts);
    d_e = dalloc(N * sizeof(double));
    d_g = dalloc(N * sizeof(double));
    d_f = dalloc(3 * N * sizeof(double));
    d_vir = dalloc(6 * N * sizeof(double));
    d_npair = dalloc(3 * N * sizeof(float));
    d_atm_pa = dalloc(6 * N * sizeof(double));
    d_out_f = dalloc(3 * N * sizeof(double));
    d_out_e = dalloc(N * sizeof(double));
    d_out_cn = dalloc(N * sizeof(double));
    d_out_g = dalloc(N * sizeof(double));
}

int Engine::type_of(uint16_t Z) const {
    auto it = std::lower_bound(elems_.begin(), elems_.end(), Z);
    if (it == elems_.end() || *it != Z) return -1;
    return (int)(it - elems_.begin());
}

void Engine::set_params(const double p[7]) {
    for (int k = 0; k < 7; ++k)
        if (!std::isfinite(p[k])) fail(D3X_ERR_INPUT, "functional parameters must be finite");
    std::memcpy(fp_, p, sizeof(fp_));
    tables_dirty_ = true;
}

void Engine::set_atoms(const double *coords_A, const uint16_t *Z, uint64_t n) {
    if (n > max_len_)
        fail(D3X_ERR_INPUT, "length " + std::to_string(n) + " exceeds max_length " + std::to_string(max_len_));
    if (n && (!coords_A || !Z)) fail(D3X_ERR_INPUT, "coords or elements is NULL");
    std::vector<uint16_t> types(n);
    std::vector<double> pos(3 * n);
    for (uint64_t i = 0; i < n; ++i) {
        const int t = type_of(Z[i]);
        if (t < 0) fail(D3X_ERR_INPUT, "element " + std::to_string(Z[i]) + " was not given at init");
        types[i] = (uint16_t)t;
        for (int c = 0; c < 3; ++c) {
            const double v = coords_A[3 * i + c];
            if (!std::isfinite(v)) fail(D3X_ERR_INPUT, "non-finite coordinate");
            pos[3 * i + c] = v / BOHR_A;  // Angstrom -> bohr, src/d3.cu:50-56
        }
    }
    const bool was_batch = sys_.size() != 1;
    if (was_batch) sys_.assign(1, Cell{});
    std::vector<int> sid(n, 0);
    if (n) {
        HIPCHK(hipMemcpyAsync(d_pos_in, pos.data(), pos.size() * sizeof(double), hipMemcpyHostToDevice, stream_));
        HIPCHK(hipMemcpyAsync(d_type_in, types.data(), n * sizeof(uint16_t), hipMemcpyHostToDevice, stream_));
        HIPCHK(hipMemcpyAsync(d_sys_in, sid.data(), n * sizeof(int), hipMemcpyHostToDevice, stream_));
        HIPCHK(hipStreamSynchronize(stream_));
    }
    // the bins and stencils depend on the cell and the atom count only
    if (was_batch || n != n_ || !atoms_set_) geom_dirty_ = true;
    // the kept neighbour state serves while the same atoms have moved less than half the skin
    nbr_moved_far_ = true;
    if (skin_ > 0 && nbr_valid_ && !was_batch && types == nbr_typ_) {
        double m2 = 0;
        for (uint64_t i = 0; i < n; ++i) {
            double d2 = 0;
            for (int c = 0; c < 3; ++c) {
                const double d = pos[3 * i + c] - nbr_pos_[3 * i + c];
                d2 += d * d;
            }
            m2 = std::max(m2, d2);
        }
        nbr_moved_far_ = !(m2 <= 0.25 * skin_ * skin_);
    }
    h_pos_ = std::move(pos);
    h_typ_ = std::move(types);
    sys_[0].n = n;
    n_ = n;
    atoms_set_ = true;
    feat_valid_ = false;
}

void Engine::set_cell(const double cell_A[9]) {
    double H[9];
    for (int k = 0; k < 9; ++k) {
        if (!std::isfinite(cell_A[k])) fail(D3X_ERR_INPUT, "non-finite cell");
        H[k] = cell_A[k] / BOHR_A;
    }
    if (!(std::fabs(det3(H)) > 1e-12)) fail(D3X_ERR_INPUT, "singular cell");
    const bool same = sys_.size() == 1 && sys_[0].set && std::memcmp(sys_[0].H, H, sizeof(H)) == 0;
    if (sys_.size() != 1) {
        // after a batch the device holds the batch's atoms and their cell ids: one cell over
        // them would be wrong, so the atoms must be set again
        sys_.assign(1, Cell{});
        n_ = 0;
        atoms_set_ = false;
    }
    std::memcpy(sys_[0].H, H, sizeof(H));
    sys_[0].set = true;
    sys_[0].n = n_;
    if (!same) {  // the same cell again (an MD step at constant volume) keeps bins and stencils
        geom_dirty_ = true;
        nbr_valid_ = false;
    }
    feat_valid_ = false;
}

void Engine::set_batch(uint64_t nsys, const uint64_t *counts, const double *coords_A, const uint16_t *Z,
                       const double *cells_A) {
    if (nsys == 0 || !counts || !cells_A) fail(D3X_ERR_INPUT, "empty batch");
    uint64_t tot = 0;
    for (uint64_t s = 0; s < nsys; ++s) tot += counts[s];
    if (tot > max_len_) fail(D3X_ERR_INPUT, "batch atom count exceeds max_length");
    if (tot && (!coords_A || !Z)) fail(D3X_ERR_INPUT, "coords or elements is NULL");
    std::vector<Cell> cells(nsys);
    for (uint64_t s = 0; s < nsys; ++s) {
        for (int k = 0; k < 9; ++k) {
            if (!std::isfinite(cells_A[9 * s + k])) fail(D3X_ERR_INPUT, "non-finite cell");
            cells[s].H[k] = cells_A[9 * s + k] / BOHR_A;
        }
        if (!(std::fabs(det3(cells[s].H)) > 1e-12)) fail(D3X_ERR_INPUT, "singular cell");
        cells[s].set = true;
        cells[s].n = counts[s];
    }
    std::vector<uint16_t> types(tot);
    std::vector<double> pos(3 * tot);
    std::vector<int> sid(tot);
    uint64_t i = 0;
    for (uint64_t s = 0; s < nsys; ++s)
        for (uint64_t k = 0; k < counts[s]; ++k, ++i) {
            const int t = type_of(Z[i]);
            if (t < 0) fail(D3X_ERR_INPUT, "element " + std::to_string(Z[i]) + " was not given at init");
            types[i] = (uint16_t)t;
            sid[i] = (int)s;
            for (int c = 0; c < 3; ++c) {
                if (!std::isfinite(coords_A[3 * i + c])) fail(D3X_ERR_INPUT, "non-finite coordinate");
                pos[3 * i + c] = coords_A[3 * i + c] / BOHR_A;
            }
        }
    if (tot) {
        HIPCHK(hipMemcpyAsync(d_pos_in, pos.data(), pos.size() * sizeof(double), hipMemcpyHostToDevice, stream_));
        HIPCHK(hipMemcpyAsync(d_type_in, types.data(), tot * sizeof(uint16_t), hipMemcpyHostToDevice, stream_));
        HIPCHK(hipMemcpyAsync(d_sys_in, sid.data(), tot * sizeof(int), hipMemcpyHostToDevice, stream_));
        HIPCHK(hipStreamSynchronize(stream_));
    }
    sys_ = std::move(cells);
    n_ = tot;
    atoms_set_ = true;
    geom_dirty_ = true;
    nbr_valid_ = false;
    nbr_moved_far_ = true;
    feat_valid_ = false;
}

void Engine::set_skin(double skin) {
    if (!(skin >= 0) || !std::isfinite(skin)) fail(D3X_ERR_INPUT, "skin must be finite and >= 0");
    if (skin != skin_) {
        skin_ = skin;
        geom_dirty_ = true;  // stencils cover cutoff + skin
        nbr_valid_ = false;
    }
    if (skin_ > 0 && !d_shift) d_shift = dalloc(3 * max_len_ * sizeof(double));
}

void Engine::set_neighbour_lists(int mode) {
    if (mode < -1 || mode > 1) fail(D3X_ERR_INPUT, "neighbour list mode must be -1, 0 or 1");
    fl_mode_ = mode;
    nbr_valid_ = false;
}

void Engine::clear() {
    // Every buffer that a compute reads is rewritten by the compute before use (the reference's
    // clear zeroes accumulators it adds into, src/d3_buffer.cu:479-494); only the status resets.
    HIPCHK(hipMemsetAsync(d_status, 0, sizeof(unsigned), stream_));
    HIPCHK(hipStreamSynchronize(stream_));
}

int Engine::stencil_count(int cut) const {
    int c = 0;
    for (const auto &s : sysdev_) c += s.st_cnt[cut];
    return c;
}

void Engine::build_geometry() {
    sysdev_.clear();
    st_d_.clear();
    st_D_.clear();
    chunk_b_.clear();
    chunk_e_.clear();
    chunk_s_.clear();
    sys_chunk_off_.assign(1, 0);
    int bin_base = 0, atom_base = 0;
    for (size_t s = 0; s < sys_.size(); ++s) {
        const Cell &C = sys_[s];
        if (!C.set) fail(D3X_ERR_STATE, "cell not set");
        SysDev d{};
        std::memcpy(d.H, C.H, sizeof(d.H));
        inv3(C.H, d.Hinv);
        d.inv_vol = 1.0 / det3(C.H);
        int64_t nbt = 1;
        for (int k = 0; k < 3; ++k) {
            const double nk = std::sqrt(d.Hinv[0 * 3 + k] * d.Hinv[0 * 3 + k] + d.Hinv[1 * 3 + k] * d.Hinv[1 * 3 + k] +
                                        d.Hinv[2 * 3 + k] * d.Hinv[2 * 3 + k]);
            const double h = 1.0 / nk;
            d.nb[k] = std::max(1, (int)std::floor(h / bin_width_));
            nbt *= d.nb[k];
        }
        if (nbt > (1 << 22)) fail(D3X_ERR_INPUT, "cell too large for the bin grid");
        {
            // bin shape for the per-atom bin skip (Centre::bins)
            bool ortho = true;
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c)
                    if (r != c && C.H[3 * r + c] != 0.0) ortho = false;
            for (int k = 0; k < 3; ++k) {
                const double nk = std::sqrt(d.Hinv[0 * 3 + k] * d.Hinv[0 * 3 + k] + d.Hinv[1 * 3 + k] * d.Hinv[1 * 3 + k] +
                                            d.Hinv[2 * 3 + k] * d.Hinv[2 * 3 + k]);
                for (int c = 0; c < 3; ++c) d.nrm[3 * k + c] = d.Hinv[c * 3 + k] / nk;
                d.hh[k] = 0.5 / (nk * d.nb[k]);
            }
            double diag = 0;
            for (int s1 = -1; s1 <= 1; s1 += 2)
                for (int s2 = -1; s2 <= 1; s2 += 2) {
                    double v[3];
                    for (int c = 0; c < 3; ++c)
                        v[c] = C.H[c] / d.nb[0] + s1 * C.H[3 + c] / d.nb[1] + s2 * C.H[6 + c] / d.nb[2];
                    diag = std::max(diag, std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]));
                }
            d.hdiag = 0.5 * diag;
            d.pad = 0.5 * skin_ + 1e-3;
            d.ortho = ortho ? 1 : 0;
        }
        d.bin_base = bin_base;
        d.atom_base = atom_base;
        d.n_atoms = (int)C.n;
        // the stand-in's feature stencil is built only when a mode needs it
        // with a skin every stencil covers cutoff + skin: atoms binned at the last rebuild may
        // have moved by up to half of it each
        const double rcs[NCUT] = {rccn_, rc2_, rc3_, atm_mode_ != 0 ? sur::RCF : 0.0};
        for (int cut = 0; cut < NCUT; ++cut) {
            d.st_off[cut] = (int)st_d_.size();
            if (rcs[cut] > 0) make_stencil(C.H, d.nb, rcs[cut] + skin_, st_d_, st_D_);
            d.st_cnt[cut] = (int)st_d_.size() - d.st_off[cut];
            if (cut == CUT_ATM && d.st_cnt[cut] > 0) {
                // the three-body kernel takes the entries after the middle one, d = 0, as forward
                const StEntry &m = st_d_[d.st_off[cut] + (d.st_cnt[cut] - 1) / 2];
                if (d.st_cnt[cut] % 2 == 0 || m.dx || m.dy || m.dz) fail(0, "three-body stencil not centred on d = 0");
            }
        }
        for (int a = 0; a < (int)C.n; a += CHUNK) {
            chunk_b_.push_back(atom_base + a);
            chunk_e_.push_back(atom_base + std::min<int>((int)C.n, a + CHUNK));
            chunk_s_.push_back((int)s);
        }
        sys_chunk_off_.push_back((int)chunk_b_.size());
        bin_base += (int)nbt;
        atom_base += (int)C.n;
        sysdev_.push_back(d);
    }
    nbins_ = bin_base;
}

void Engine::upload_geometry() {
    const size_t ts = precision_ == D3X_FP64 ? sizeof(double) : sizeof(float);
    if ((size_t)nbins_ + 1 > bins_cap_) {
        dfree(d_bin_count);
        dfree(d_bin_start);
        dfree(d_cursor);
        bins_cap_ = (size_t)nbins_ + 1;
        d_bin_count = dalloc(bins_cap_ * sizeof(int));
        d_bin_start = dalloc(bins_cap_ * sizeof(int));
        d_cursor = dalloc(bins_cap_ * sizeof(int));
    }
    if (sysdev_.size() > sys_cap_) {
        dfree(d_sys);
        dfree(d_out_sys);
        dfree(d_sys_chunk);
        sys_cap_ = sysdev_.size();
        d_sys = dalloc(sys_cap_ * sizeof(SysDev));
        d_out_sys = dalloc(sys_cap_ * RED_W * sizeof(double));
        d_sys_chunk = dalloc((sys_cap_ + 1) * sizeof(int));
    }
    if (st_d_.size() > st_cap_) {
        dfree(d_st_d);
        dfree(d_st_D);
        st_cap_ = st_d_.size();
        d_st_d = dalloc(st_cap_ * sizeof(StEntry));
        d_st_D = dalloc(st_cap_ * 4 * sizeof(double));
    }
    if (chunk_b_.size() > chunk_cap_ || chunk_cap_ == 0) {
        dfree(d_chunk_b);
        dfree(d_chunk_e);
        dfree(d_chunk_s);
        dfree(d_partial);
        chunk_cap_ = std::max<size_t>(1, chunk_b_.size());
        d_chunk_b = dalloc(chunk_cap_ * sizeof(int));
        d_chunk_e = dalloc(chunk_cap_ * sizeof(int));
        d_chunk_s = dalloc(chunk_cap_ * sizeof(int));
        d_partial = dalloc(chunk_cap_ * RED_W * sizeof(double));
    }
    HIPCHK(hipMemcpyAsync(d_sys, sysdev_.data(), sysdev_.size() * sizeof(SysDev), hipMemcpyHostToDevice, stream_));
    if (!st_d_.empty()) {
        HIPCHK(hipMemcpyAsync(d_st_d, st_d_.data(), st_d_.size() * sizeof(StEntry), hipMemcpyHostToDevice, stream_));
        std::vector<unsigned char> buf(st_D_.size() * ts);
        if (ts == sizeof(double))
            std::memcpy(buf.data(), st_D_.data(), buf.size());
        else
            for (size_t k = 0; k < st_D_.size(); ++k) reinterpret_cast<float *>(buf.data())[k] = (float)st_D_[k];
        HIPCHK(hipMemcpyAsync(d_st_D, buf.data(), buf.size(), hipMemcpyHostToDevice, stream_));
        HIPCHK(hipStreamSynchronize(stream_));
    }
    if (!chunk_b_.empty()) {
        HIPCHK(hipMemcpyAsync(d_chunk_b, chunk_b_.data(), chunk_b_.size() * sizeof(int), hipMemcpyHostToDevice, stream_));
        HIPCHK(hipMemcpyAsync(d_chunk_e, chunk_e_.data(), chunk_e_.size() * sizeof(int), hipMemcpyHostToDevice, stream_));
        HIPCHK(hipMemcpyAsync(d_chunk_s, chunk_s_.data(), chunk_s_.size() * sizeof(int), hipMemcpyHostToDevice, stream_));
    }
    HIPCHK(hipMemcpyAsync(d_sys_chunk, sys_chunk_off_.data(), sys_chunk_off_.size() * sizeof(int),
                          hipMemcpyHostToDevice, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
}

void Engine::ensure_tb_scratch() {
    // Two-body waves per centre: about 4096 waves in all for small cells (512 centres alone leave
    // most SIMDs idle); parts x n stays below 8192 whenever parts > 1.
    const uint64_t ncen = std::max<uint64_t>(n_, 1);
    tb_parts_ = (int)std::min<uint64_t>(8, std::max<uint64_t>(1, (4096 + ncen - 1) / ncen));
    const size_t need = (size_t)tb_parts_ * n_ * TB_PART_W;
    if (tb_parts_ > 1 && need > tb_part_alloc_) {
        dfree(d_tb_part);
        tb_part_alloc_ = std::max<size_t>(need, (size_t)8192 * TB_PART_W);
        d_tb_part = dalloc(tb_part_alloc_ * sizeof(double));
    }
}

void Engine::ensure_atm_scratch() {
    if (rc3_ <= 0) return;
    // forward-neighbour capacity: 1.5 x the mean number of forward neighbours (half of all
    // neighbours within rc3) plus a margin; a centre above it raises D3X_ERR_OVERFLOW and the
    // caller retries with a larger capacity
    double dens = 0;
    for (size_t s = 0; s < sys_.size(); ++s) dens = std::max(dens, (double)sys_[s].n / std::fabs(det3(sys_[s].H)));
    const double mean = 0.5 * dens * 4.18879020478639 * rc3_ * rc3_ * rc3_;
    int cap = (int)std::min(262144.0, 1.5 * mean + 1024.0);
    cap = (std::max(cap, atm_cap_) + 63) / 64 * 64;
    int nb = atm_blocks_req_ > 0 ? atm_blocks_req_ : atm_resident();
    // Work items per centre: small cells would leave most resident waves without a centre (512
    // centres on 1024 SIMDs), and a few centres per wave leave a long tail; split each centre's k
    // tiles into up to 8 items until there are about 8 items per wave.
    const uint64_t ncen = std::max<uint64_t>(n_, 1);
    atm_chunks_ = (int)std::min<uint64_t>(8, std::max<uint64_t>(1, (8 * (uint64_t)nb + ncen - 1) / ncen));
    nb = (int)std::max<uint64_t>(1, std::min<uint64_t>((uint64_t)nb, ncen * atm_chunks_));
    if (atm_chunks_ > 1 && (size_t)atm_chunks_ * n_ > atm_eatc_alloc_) {
        dfree(d_atm_eatc);
        atm_eatc_alloc_ = (size_t)8 * max_len_;
        d_atm_eatc = dalloc(atm_eatc_alloc_ * sizeof(double));
    }
    atm_cap_ = cap;
    atm_blocks_ = nb;
    // forward lists: with a skin (kept across evaluations), or when centres are split into work
    // items that keep every resident block busy several times over (each item would walk the
    // stencil again: water box 62.1 -> 60.9 ms), or always when asked; within 4 GiB of entries.
    // With fewer items the walks run side by side and a list pass before them only adds its own
    // latency (512-atom Si: 1.95 -> 2.01 ms).
    const bool many = atm_chunks_ > 1 && (uint64_t)n_ * atm_chunks_ >= 4 * (uint64_t)nb;
    const bool want = fl_mode_ == 1 || (fl_mode_ < 0 && (skin_ > 0 || many));
    const size_t fl_need = (size_t)std::max<uint64_t>(n_, 1) * (size_t)cap;
    fl_use_ = want && fl_need * sizeof(FwdEntry) <= ((size_t)4 << 30);
    if (fl_use_ && (fl_need > fl_alloc_ || !d_fl_n)) {
        dfree(d_fl);
        dfree(d_fl_n);
        fl_alloc_ = fl_need;
        d_fl = dalloc(fl_alloc_ * sizeof(FwdEntry));
        d_fl_n = dalloc(max_len_ * sizeof(int));
    }
    // the kernel indexes scratch as blockIdx.x * atm_cap_: reallocate whenever either grows
    if ((size_t)cap * nb > atm_alloc_entries_ || nb > atm_alloc_blocks_ || !d_atm_i32) {
        dfree(d_atm_i32);
        dfree(d_atm_T);
        dfree(d_atm_acc);
        const size_t ts = precision_ == D3X_FP64 ? sizeof(double) : sizeof(float);
        atm_alloc_entries_ = (size_t)cap * nb;
        atm_alloc_blocks_ = nb;
        d_atm_i32 = dalloc(atm_alloc_entries_ * 2 * sizeof(int));
        d_atm_T = dalloc((size_t)nb * atm_T_stride(cap) * ts);
        d_atm_acc = dalloc((size_t)nb * atm_acc_stride(cap) * sizeof(double));
    }
}

template <class T>
Params<T> Engine::params() {
    Params<T> P;
    std::memset(&P, 0, sizeof(P));  // padding too: the graph signature compares the raw bytes
    P.n = (int)n_;
    P.ne = ne_;
    P.nref = nref_;
    P.nsys = (int)sysdev_.size();
    P.damping = damping_;
    P.rc_cn2 = (T)(rccn_ * rccn_);
    P.rc_22 = (T)(rc2_ * rc2_);
    P.rc_32 = (T)(rc3_ * rc3_);
    P.rc_cn = (T)rccn_;
    P.cn_smooth = cn_smooth_ ? 1 : 0;
    P.s6 = (T)fp_[0];
    if (damping_ == 0) {
        P.s8 = (T)fp_[1];
        P.p1 = (T)fp_[3];
        P.p2 = (T)fp_[4];
    } else {
        P.s8 = (T)fp_[2];
        P.p1 = (T)fp_[5];
        P.p2 = (T)fp_[6];
    }
    P.sys = (const SysDev *)d_sys;
    P.st_d = (const StEntry *)d_st_d;
    P.st_D = (const T *)d_st_D;
    P.bin_start = (const int *)d_bin_start;
    P.pos_in = (const double *)d_pos_in;
    P.type_in = (const uint16_t *)d_type_in;
    P.sys_in = (const int *)d_sys_in;
    P.key_in = (int *)d_key;
    P.bin_count = (int *)d_bin_count;
    P.cursor = (int *)d_cursor;
    P.perm = (int *)d_perm;
    P.perm2 = (int *)d_perm2;
    P.wrapped = (double *)d_wrapped;
    P.lx = (T *)d_lx;
    P.ly = (T *)d_ly;
    P.lz = (T *)d_lz;
    P.type_s = (uint16_t *)d_type_s;
    P.orig_s = (int *)d_orig_s;
    P.bin_s = (int *)d_bin_s;
    P.sys_s = (int *)d_sys_s;
    P.cn = (T *)d_cn;
    P.w = (T *)d_w;
    P.dw = (T *)d_dw;
    P.e = (double *)d_e;
    P.g = (double *)d_g;
    P.f = (double *)d_f;
    P.vir = (double *)d_vir;
    P.npair = (float *)d_npair;
    const T *tab = (const T *)d_tab;
    P.c6 = tab + tab_off_[0];
    P.cnref = tab + tab_off_[1];
    P.rcov = tab + tab_off_[2];
    P.q4 = tab + tab_off_[3];
    P.rdamp = tab + tab_off_[4];
    P.r0 = tab + tab_off_[5];
    P.c6g = c6_pairwise_ ? tab + tab_off_[6] : nullptr;
    P.c6_pairwise = c6_pairwise_ ? 1 : 0;
    P.atm_cap = atm_cap_;
    P.atm_var = rc3_ > 0 ? atm_var() : 0;
    P.atm_chunks = atm_chunks_;
    P.atm_eatc = (double *)d_atm_eatc;
    P.tb_parts = tb_parts_;
    P.tb_part = (double *)d_tb_part;
    P.atm_roff = (const int *)d_atm_roff;
    P.atm_yrows = yrows_;
    P.atm_fl = fl_use_ ? (FwdEntry *)d_fl : nullptr;
    P.atm_fl_n = (int *)d_fl_n;
    P.atm_fl_cap = atm_cap_;
    P.rc_fl2 = (T)((rc3_ + skin_) * (rc3_ + skin_));
    P.shift = skin_ > 0 ? (double *)d_shift : nullptr;
    P.atm_i32 = (int *)d_atm_i32;
    P.atm_T = (T *)d_atm_T;
    P.atm_acc = (double *)d_atm_acc;
    // packed for the current atom count so that one fill of 6 n doubles clears all four
    P.fa = (double *)d_atm_pa;
    P.ga = P.fa + 3 * n_;
    P.eas = P.ga + n_;
    P.eat = P.eas + n_;
    P.status = (unsigned *)d_status;
    P.atm_next = (int *)d_status + 1;
    P.type_cls = (const int *)d_type_cls;
    P.feat = (float *)d_feat;
    P.sur_G = (float *)d_sur_G;
    P.sur_E = (float *)d_sur_E;
    P.sur_cls = (int *)d_sur_cls;
    P.sur_q = (float *)d_sur_q;
    P.sur_dcn = (float *)d_sur_dcn;
    P.sur_P = (const float *)d_sur_P;
    P.sur_PK = d_sur_PK;
    return P;
}

template <class T>
void Engine::run() {
    if (tables_dirty_) {
        // c6 | cnref | rcov | q4 | rdamp | r0 in T
        const int ne = ne_, nr = nref_;
        const size_t nc6 = (size_t)ne * ne * nr * nr, ncr = (size_t)ne * nr;
        const double *c6 = h_tab_.data(), *cnref = c6 + nc6, *rcov = cnref + ncr, *r2r4 = rcov + ne,
                     *r0ab = r2r4 + ne;
        std::vector<T> t;
        tab_off_[0] = t.size();
        for (size_t k = 0; k < nc6; ++k) t.push_back((T)c6[k]);
        tab_off_[1] = t.size();
        for (size_t k = 0; k < ncr; ++k) t.push_back((T)cnref[k]);
        tab_off_[2] = t.size();
        for (int A = 0; A < ne; ++A) t.push_back((T)rcov[A]);
        tab_off_[3] = t.size();
        for (int A = 0; A < ne; ++A)
            for (int B = 0; B < ne; ++B) t.push_back((T)(r2r4[A] * r2r4[B]));
        tab_off_[4] = t.size();
        for (int A = 0; A < ne; ++A)
            for (int B = 0; B < ne; ++B) {
                // damping radius: zero -> r0ab; BJ -> a1 sqrt(3 r2r4_A r2r4_B) + a2 (src/d3_kernel.cu:1083,388)
                const double rd = damping_ == 0 ? r0ab[A * ne + B] : fp_[5] * std::sqrt(3.0 * r2r4[A] * r2r4[B]) + fp_[6];
                t.push_back((T)rd);
            }
        tab_off_[5] = t.size();
        for (int k = 0; k < ne * ne; ++k) t.push_back((T)r0ab[k]);
        tab_off_[6] = t.size();
        for (double v : h_c6g_) t.push_back((T)v);
        dfree(d_tab);
        d_tab = dalloc(t.size() * sizeof(T));
        HIPCHK(hipMemcpyAsync(d_tab, t.data(), t.size() * sizeof(T), hipMemcpyHostToDevice, stream_));
        HIPCHK(hipStreamSynchronize(stream_));
        tables_dirty_ = false;
    }
    ensure_atm_scratch();
    ensure_tb_scratch();
    ensure_surrogate();
    if (skin_ > 0 && !d_shift) d_shift = dalloc(3 * max_len_ * sizeof(double));
    Params<T> P = params<T>();
    feat_valid_ = mode_now() != D3X_ATM_EXACT;
    // keep the neighbour state of the last rebuild when the atoms stayed within half the skin
    // and the lists this evaluation needs were built then, with this capacity
    const bool need_lists = fl_use_ && rc3_ > 0 && mode_now() != D3X_ATM_SURROGATE;
    const bool reuse = reuse_now() && (!need_lists || nbr_fl_) && atm_cap_ == nbr_cap_;
    if (reuse) {
        ++nbr_reuses_;
    } else {
        ++nbr_rebuilds_;
        if (skin_ > 0 && sys_.size() == 1) {
            nbr_valid_ = true;
            nbr_moved_far_ = false;
            nbr_pos_ = h_pos_;
            nbr_typ_ = h_typ_;
            nbr_fl_ = need_lists;
            nbr_cap_ = atm_cap_;
        }
    }
    if (graph_mode_ && !timing_) {
        // The graph holds every kernel argument and launch shape; re-capture when any changes.
        std::vector<unsigned char> sig(sizeof(P) + 8 * sizeof(int64_t));
        std::memcpy(sig.data(), &P, sizeof(P));
        const int64_t extra[8] = {nbins_, atm_blocks_, (int64_t)chunk_b_.size(), (int64_t)sysdev_.size(),
                                  (int64_t)(intptr_t)stream_, (int64_t)(intptr_t)d_tab, mode_now(), reuse};
        std::memcpy(sig.data() + sizeof(P), extra, sizeof(extra));
        if (!gexec_ || sig != gsig_) {
            if (gexec_) (void)hipGraphExecDestroy(gexec_);
            gexec_ = nullptr;
            hipGraph_t g = nullptr;
            HIPCHK(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
            enqueue<T>(P, reuse);
            HIPCHK(hipStreamEndCapture(stream_, &g));
            const hipError_t ie = hipGraphInstantiate(&gexec_, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            HIPCHK(ie);
            gsig_ = std::move(sig);
            ++graph_captures_;
        }
        HIPCHK(hipGraphLaunch(gexec_, stream_));
        return;
    }
    if (timing_) harvest();
    enqueue<T>(P, reuse);
    if (timing_) timing_pending_ = true;
}

template <class T>
void Engine::enqueue(const Params<T> &P, bool reuse) {
    launch_fill_u32((unsigned *)d_status, 1, 0u, stream_);
    mark(0);
    if (reuse)
        Launch<T>::regather(P, stream_);  // bins and order of the last rebuild
    else
        Launch<T>::bin_sort(P, nbins_, (int *)d_bin_start, stream_);
    mark(1);
    Launch<T>::cn(P, stream_);
    mark(2);
    Launch<T>::twobody(P, stream_);
    mark(3);
    if (mode_now() == D3X_ATM_SURROGATE) {
        launch_surrogate<T>(P, stream_);
    } else {
        if (P.atm_fl && !reuse) launch_atm_lists<T>(P, stream_);
        launch_atm<T>(P, atm_blocks_, stream_);  // zeroes the ATM arrays; kernels only when rc3 > 0
        if (mode_now() == D3X_ATM_EXACT_FEATURES) launch_sur_features<T>(P, stream_);
    }
    mark(4);
    Launch<T>::chain(P, stream_);
    mark(5);
    launch_reduce((const double *)d_e, P.eat, (const double *)d_vir, (const float *)d_npair, (const int *)d_chunk_b,
                  (const int *)d_chunk_e, (const int *)d_chunk_s, (int)chunk_b_.size(), (const int *)d_sys_chunk,
                  (int)sysdev_.size(), (double *)d_partial, (double *)d_out_sys, stream_);
    launch_finalize(P, (double *)d_out_f, (double *)d_out_e, (double *)d_out_cn, (double *)d_out_g, stream_);
    mark(6);
    HIPCHK(hipGetLastError());
}

// ------------------------------------------------------------------------------------ tuning

std::string Engine::tune_key() const {
    // precision | cutoffs (0.01 bohr) | density bucket (quarter octaves of atoms/bohr^3) | elements
    double dens = 0;
    for (const Cell &c : sys_)
        if (c.set) dens = std::max(dens, (double)c.n / std::fabs(det3(c.H)));
    const int db = dens > 0 ? (int)std::lround(4.0 * std::log2(dens)) : -999;
    char buf[160];
    std::snprintf(buf, sizeof(buf), "%s|rc%.2f|cn%.2f|atm%.2f|d%d|ne%d|cells%s%s",
                  precision_ == D3X_FP64 ? "fp64" : "fp32", rc2_, rccn_, rc3_, db, ne_, sys_.size() > 1 ? "N" : "1",
                  c6_pairwise_ ? "|c6pair" : "");
    return buf;
}

namespace {
// Cache file: one JSON object, one entry per line:
//   "key": {"bin_width": w, "atm_blocks": b, "atm_variant": v, "ms": t}   (v = -1: not tuned)
bool cache_lookup(const std::string &path, const std::string &key, double &w, int &b, int &v) {
    FILE *f = std::fopen(path.c_str(), "r");
    if (!f) return false;
    char line[512];
    bool found = false;
    const std::string pat = "\"" + key + "\"";
    while (std::fgets(line, sizeof(line), f)) {
        std::string s(line);
        if (s.find(pat) == std::string::npos) continue;
        const size_t pw = s.find("\"bin_width\":"), pb = s.find("\"atm_blocks\":");
        if (pw == std::string::npos || pb == std::string::npos) continue;
        w = std::atof(s.c_str() + pw + 12);
        b = std::atoi(s.c_str() + pb + 13);
        const size_t pv = s.find("\"atm_variant\":");
        v = pv == std::string::npos ? -1 : std::atoi(s.c_str() + pv + 14);
        found = w > 0 && b > 0;
    }
    std::fclose(f);
    return found;
}
void cache_store(const std::string &path, const std::string &key, double w, int b, int v, double ms) {
    std::vector<std::string> keep;
    if (FILE *f = std::fopen(path.c_str(), "r")) {
        char line[512];
        while (std::fgets(line, sizeof(line), f)) {
            std::string s(line);
            if (s.find("\"bin_width\"") == std::string::npos) continue;
            if (s.find("\"" + key + "\"") != std::string::npos) continue;
            while (!s.empty() && (s.back() == '\n' || s.back() == ',' || s.back() == ' ')) s.pop_back();
            keep.push_back(s);
        }
        std::fclose(f);
    }
    char buf[512];
    std::snprintf(buf, sizeof(buf), "  \"%s\": {\"bin_width\": %.6f, \"atm_blocks\": %d, \"atm_variant\": %d, \"ms\": %.4f}",
                  key.c_str(), w, b, v, ms);
    keep.push_back(buf);
    const std::string tmp = path + ".tmp";
    FILE *f = std::fopen(tmp.c_str(), "w");
    if (!f) return;
    std::fprintf(f, "{\n");
    for (size_t k = 0; k < keep.size(); ++k) std::fprintf(f, "%s%s\n", keep[k].c_str(), k + 1 < keep.size() ? "," : "");
    std::fprintf(f, "}\n");
    std::fclose(f);
    std::rename(tmp.c_str(), path.c_str());
}
}  // namespace

// Bins half the CN cutoff wide on cells at least 6 CN cutoffs across (fcc Cu 63^3: 20 bohr is the
// fastest of 10-30); 3/8 of it on smaller cells, where the forward-neighbour tiles of the
// three-body term span many periodic images (water box 94 bohr: 63.9 ms at 15 bohr vs 66.7 ms at
// 20; the 256-cell batch: 1571 vs 1606 ms).
double Engine::default_bin_width() const {
    double hmin = 1e300;
    for (const Cell &c : sys_) {
        if (!c.set) continue;
        double Hi[9];
        inv3(c.H, Hi);
        for (int k = 0; k < 3; ++k)
            hmin = std::min(hmin, 1.0 / std::sqrt(Hi[k] * Hi[k] + Hi[3 + k] * Hi[3 + k] + Hi[6 + k] * Hi[6 + k]));
    }
    return std::max(4.0, (hmin >= 6.0 * rccn_ ? 0.5 : 0.375) * rccn_);
}

void Engine::set_tuning(double bin_width, int atm_blocks) {
    if (bin_width > 0) {
        if (!std::isfinite(bin_width)) fail(D3X_ERR_INPUT, "bin width must be finite");
        bin_width_ = bin_width;
        geom_dirty_ = true;
    }
    if (atm_blocks > 0) atm_blocks_req_ = atm_blocks;
    tuning_explicit_ = true;
}

void Engine::apply_tune_cache() {
    if (tune_cache_.empty()) return;
    double w;
    int b, v;
    if (cache_lookup(tune_cache_, tune_key(), w, b, v)) {
        bin_width_ = w;
        atm_blocks_req_ = b;
        if (atm_variant_ok(v)) set_atm_variant(v);
    }
}

double Engine::timed_evaluation() {
    nbr_valid_ = false;  // every candidate timed with its neighbour search
    compute_device();    // builds geometry, sizes scratch
    set_timing(true);
    nbr_valid_ = false;
    compute_device();
    double ms[NPHASE];
    phase_times(ms);
    set_timing(false);
    unsigned status = 0;
    HIPCHK(hipMemcpy(&status, d_status, sizeof(unsigned), hipMemcpyDeviceToHost));
    if (status & 2u) {  // three-body capacity too small for this input: grow and time again
        grow_atm_cap();
        return timed_evaluation();
    }
    double t = 0;
    for (double m : ms) t += m;
    return t;
}

double Engine::autotune(const std::string &cache_path) {
    if (!atoms_set_) fail(D3X_ERR_STATE, "set_atoms has not been called");
    // D3_TUNE_LOG=1: one line per timed candidate on stderr (a sweep over a large batch in FP64
    // runs for minutes)
    const char *lg = std::getenv("D3_TUNE_LOG");
    const bool verbose = lg && *lg && std::strcmp(lg, "0") != 0;
    auto report = [&](double w, int b, double t, int var) {
        if (!verbose) return;
        if (var >= 0)
            std::fprintf(stderr, "d3 autotune: bin %.2f bohr, three-body variant %d, blocks %d: %.2f ms\n", w, var, b, t);
        else
            std::fprintf(stderr, "d3 autotune: bin %.2f bohr, three-body blocks %d: %.2f ms\n", w, b, t);
    };
    const bool graph = graph_mode_;
    set_graph(false);
    tuning_explicit_ = true;  // the candidates below, not the default or a cached entry
    const double base = default_bin_width();
    std::vector<double> widths;
    for (double fct : {1.0, 0.75, 0.6, 1.5, 0.5})
        if (base * fct >= 4.0) widths.push_back(base * fct);
    double best_w = bin_width_, best = 1e300;
    for (double w : widths) {
        bin_width_ = w;
        geom_dirty_ = true;
        const double t = timed_evaluation();
        report(w, atm_blocks_effective(), t, -1);
        if (t < best) {
            best = t;
            best_w = w;
        }
    }
    bin_width_ = best_w;
    geom_dirty_ = true;
    if (rc3_ > 0) {
        // three-body kernel variants (KL, DR) that fit, each at its own resident block count
        const int req = atm_blocks_req_;
        int best_v = atm_var();
        best = 1e300;
        for (int v = 0; v < 4; ++v) {
            if (!atm_variant_ok(v)) continue;
            set_atm_variant(v);
            atm_blocks_req_ = 0;
            const double t = timed_evaluation();
            report(bin_width_, atm_resident(), t, v);
            if (t < best) {
                best = t;
                best_v = v;
            }
        }
        set_atm_variant(best_v);
        atm_blocks_req_ = req;
    }
    const int res = atm_resident();
    int best_b = atm_blocks_req_ > 0 ? atm_blocks_req_ : res;
    if (rc3_ > 0) {
        best = 1e300;
        // blocks take centres from a shared counter: beyond the resident count more blocks only
        // add scratch, fewer leave compute units idle
        for (int b : {res / 2, (3 * res) / 4, res, 2 * res}) {
            atm_blocks_req_ = b;
            const double t = timed_evaluation();
            report(bin_width_, b, t, -1);
            if (t < best) {
                best = t;
                best_b = b;
            }
        }
    } else {
        best = timed_evaluation();
    }
    atm_blocks_req_ = best_b;
    tuning_explicit_ = true;
    const std::string path = cache_path.empty() ? tune_cache_ : cache_path;
    if (!path.empty()) cache_store(path, tune_key(), best_w, best_b, rc3_ > 0 ? atm_var() : -1, best);
    set_graph(graph);
    return best;
}

// ------------------------------------------------------------------------- three-body stand-in

void Engine::sur_load(const float *params) {
    if (!params) fail(D3X_ERR_INPUT, "NULL network parameters");
    for (int k = 0; k < sur::NPARAMS; ++k)
        if (!std::isfinite(params[k])) fail(D3X_ERR_INPUT, "network parameters contain non-finite values");
    h_sur_P_.assign(params, params + sur::NPARAMS);
    sur_dirty_ = true;
}

void Engine::set_atm_mode(int mode, int judge_every, double tol_ev_per_atom) {
    if (mode != D3X_ATM_EXACT && mode != D3X_ATM_SURROGATE && mode != D3X_ATM_EXACT_FEATURES)
        fail(D3X_ERR_INPUT, "unknown three-body mode");
    if (mode != D3X_ATM_EXACT && rc3_ <= 0) fail(D3X_ERR_INPUT, "the handle has no three-body term (atm_cutoff <= 0)");
    if (mode != D3X_ATM_EXACT && c6_pairwise_)  // its features weight neighbours by per-element C6 rows
        fail(D3X_ERR_INPUT, "the three-body network needs CN references that factorise per element");
    if (mode == D3X_ATM_SURROGATE && h_sur_P_.empty()) fail(D3X_ERR_STATE, "no network parameters loaded");
    if (judge_every < 0) fail(D3X_ERR_INPUT, "judge_every must be >= 0");
    if (!(tol_ev_per_atom > 0)) fail(D3X_ERR_INPUT, "tolerance must be positive");
    if ((atm_mode_ != D3X_ATM_EXACT) != (mode != D3X_ATM_EXACT)) geom_dirty_ = true;  // feature stencil
    atm_mode_ = mode;
    judge_every_ = judge_every;
    judge_tol_ = tol_ev_per_atom;
    fallen_back_ = false;
    sur_evals_ = 0;
}

void Engine::ensure_surrogate() {
    if (atm_mode_ == D3X_ATM_EXACT) return;
    if (!d_feat) {
        d_feat = dalloc(max_len_ * sur::NFEAT * sizeof(float));
        d_sur_G = dalloc(max_len_ * sur::NFEAT * sizeof(float));
        d_sur_E = dalloc(max_len_ * sizeof(float));
        d_sur_cls = dalloc(max_len_ * sizeof(int));
        d_sur_q = dalloc(2 * max_len_ * sizeof(float));
    }
    if (sur_dirty_ && !h_sur_P_.empty()) {
        if (!d_sur_P) {
            d_sur_P = dalloc(sur::NPARAMS * sizeof(float));
            d_sur_PK = dalloc(sur::PACKED_BYTES);
        }
        HIPCHK(hipMemcpyAsync(d_sur_P, h_sur_P_.data(), sur::NPARAMS * sizeof(float), hipMemcpyHostToDevice, stream_));
        sur::launch_pack((const float *)d_sur_P, d_sur_PK, stream_);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(stream_));
        sur_dirty_ = false;
    }
}

bool Engine::judge_due() const {
    return mode_now() == D3X_ATM_SURROGATE && (sur_evals_ == 0 || (judge_every_ > 0 && sur_evals_ % judge_every_ == 0));
}

bool Engine::record_judgement(const double *e_net, const double *e_exact, const double *f_net,
                              const double *f_exact) {
    double dev = 0, fdev = 0;
    size_t a = 0;
    for (size_t s = 0; s < sys_.size(); ++s) {
        const size_t n = sys_[s].n;
        if (n) dev = std::max(dev, std::fabs(e_net[s] - e_exact[s]) / (double)n);
        if (n && f_net && f_exact) {
            double ss = 0;
            for (size_t k = 3 * a; k < 3 * (a + n); ++k) ss += (f_net[k] - f_exact[k]) * (f_net[k] - f_exact[k]);
            fdev = std::max(fdev, std::sqrt(ss / (3.0 * (double)n)));
        }
        a += n;
    }
    last_dev_ = dev;
    last_fdev_ = (f_net && f_exact) ? fdev : -1.0;
    if (dev > judge_tol_ || (judge_ftol_ > 0 && f_net && f_exact && fdev > judge_ftol_)) {
        fallen_back_ = true;
        ++fallbacks_;
        return false;
    }
    return true;
}

void Engine::fetch_features(float *X) {
    if (!feat_valid_ || !d_feat) fail(D3X_ERR_STATE, "no features: evaluate in a mode that computes them first");
    if (n_ == 0) return;
    std::vector<float> f(n_ * sur::NFEAT);
    std::vector<int> orig(n_);
    HIPCHK(hipMemcpyAsync(f.data(), d_feat, f.size() * sizeof(float), hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipMemcpyAsync(orig.data(), d_orig_s, n_ * sizeof(int), hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
    for (uint64_t i = 0; i < n_; ++i)
        std::memcpy(X + (size_t)orig[i] * sur::NFEAT, f.data() + i * sur::NFEAT, sur::NFEAT * sizeof(float));
}

void Engine::sur_tangent(const double *dir, float *U) {
    if (!feat_valid_ || !d_feat) fail(D3X_ERR_STATE, "no features: evaluate in a mode that computes them first");
    if (n_ == 0) return;
    std::vector<int> orig(n_);
    HIPCHK(hipMemcpyAsync(orig.data(), d_orig_s, n_ * sizeof(int), hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
    std::vector<float> ds(3 * n_);
    for (uint64_t i = 0; i < n_; ++i)
        for (int k = 0; k < 3; ++k) {
            const double v = dir[3 * (size_t)orig[i] + k];
            if (!std::isfinite(v)) fail(D3X_ERR_INPUT, "non-finite direction");
            ds[3 * i + k] = (float)v;
        }
    if (!d_tan_dir) {
        d_tan_dir = dalloc(max_len_ * 3 * sizeof(float));
        d_tan_U = dalloc(max_len_ * sur::NFEAT * sizeof(float));
        d_sur_dcn = dalloc(max_len_ * sizeof(float));
    }
    HIPCHK(hipMemcpyAsync(d_tan_dir, ds.data(), ds.size() * sizeof(float), hipMemcpyHostToDevice, stream_));
    if (precision_ == D3X_FP64)
        launch_sur_tangent<double>(params<double>(), (const float *)d_tan_dir, (float *)d_tan_U, stream_);
    else
        launch_sur_tangent<float>(params<float>(), (const float *)d_tan_dir, (float *)d_tan_U, stream_);
    HIPCHK(hipGetLastError());
    std::vector<float> u(n_ * sur::NFEAT);
    HIPCHK(hipMemcpyAsync(u.data(), d_tan_U, u.size() * sizeof(float), hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
    for (uint64_t i = 0; i < n_; ++i)
        std::memcpy(U + (size_t)orig[i] * sur::NFEAT, u.data() + i * sur::NFEAT, sur::NFEAT * sizeof(float));
}

void Engine::set_graph(bool on) {
    graph_mode_ = on;
    if (!on && gexec_) {
        (void)hipStreamSynchronize(stream_);
        (void)hipGraphExecDestroy(gexec_);
        gexec_ = nullptr;
        gsig_.clear();
    }
}

void Engine::mark(int k) {
    if (timing_) HIPCHK(hipEventRecord(ev_[k], stream_));
}

void Engine::harvest() {
    if (!timing_pending_) return;
    HIPCHK(hipEventSynchronize(ev_[NPHASE]));
    for (int k = 0; k < NPHASE; ++k) {
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, ev_[k], ev_[k + 1]));
        phase_ms_[k] += ms;
    }
    ++timed_runs_;
    timing_pending_ = false;
}

void Engine::set_timing(bool on) {
    if (on && !ev_[0])
        for (auto &e : ev_) HIPCHK(hipEventCreate(&e));
    timing_ = on;
    timing_pending_ = false;
    timed_runs_ = 0;
    for (double &m : phase_ms_) m = 0;
}

int Engine::phase_times(double ms[NPHASE]) {
    if (timing_) harvest();
    for (int k = 0; k < NPHASE; ++k) ms[k] = phase_ms_[k];
    return timed_runs_;
}

void Engine::compute_device() {
    if (!atoms_set_) fail(D3X_ERR_STATE, "set_atoms has not been called");
    for (const Cell &c : sys_)
        if (!c.set) fail(D3X_ERR_STATE, "set_cell has not been called");
    if (geom_dirty_) {
        nbr_valid_ = false;  // new bins or stencils
        if (!tuning_explicit_) {
            bin_width_ = default_bin_width();
            apply_tune_cache();
        }
        build_geometry();
        upload_geometry();
        geom_dirty_ = false;
    }
    if (precision_ == D3X_FP64)
        run<double>();
    else
        run<float>();
}

void Engine::fetch(double *energies, double *forces, double *stresses) {
    const size_t ns = sysdev_.size();
    std::vector<double> red(ns * RED_W);
    unsigned status = 0;
    HIPCHK(hipMemcpyAsync(red.data(), d_out_sys, red.size() * sizeof(double), hipMemcpyDeviceToHost, stream_));
    if (forces && n_)
        HIPCHK(hipMemcpyAsync(forces, d_out_f, 3 * n_ * sizeof(double), hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipMemcpyAsync(&status, d_status, sizeof(unsigned), hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
    if (status & 2u) {
        HIPCHK(hipMemsetAsync(d_status, 0, sizeof(unsigned), stream_));
        throw Error{D3X_ERR_GENERIC | D3X_ERR_OVERFLOW, "three-body neighbour list capacity exceeded"};
    }
    const double sfac = HARTREE_EV / (BOHR_A * BOHR_A * BOHR_A);  // src/d3.cu:213-219
    for (size_t s = 0; s < ns; ++s) {
        const double *r = &red[s * RED_W];
        if (energies) energies[s] = -HARTREE_EV * r[0];
        if (stresses) {
            const double v[6] = {r[1], r[2], r[3], r[4], r[5], r[6]};  // xx yy zz xy xz yz
            double *o = stresses + 9 * s;
            o[0] = v[0] * sfac;
            o[1] = v[3] * sfac;
            o[2] = v[4] * sfac;
            o[3] = v[3] * sfac;
            o[4] = v[1] * sfac;
            o[5] = v[5] * sfac;
            o[6] = v[4] * sfac;
            o[7] = v[5] * sfac;
            o[8] = v[2] * sfac;
        }
    }
}

void Engine::fetch_atom_data(double *cn, double *dedcn, double *atom_energy) {
    if (n_ == 0) return;
    if (cn) HIPCHK(hipMemcpyAsync(cn, d_out_cn, n_ * sizeof(double), hipMemcpyDeviceToHost, stream_));
    if (dedcn) HIPCHK(hipMemcpyAsync(dedcn, d_out_g, n_ * sizeof(double), hipMemcpyDeviceToHost, stream_));
    if (atom_energy) HIPCHK(hipMemcpyAsync(atom_energy, d_out_e, n_ * sizeof(double), hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
}

void Engine::fetch_counts(double *counts) {
    const size_t ns = sysdev_.size();
    std::vector<double> red(ns * RED_W);
    HIPCHK(hipMemcpyAsync(red.data(), d_out_sys, red.size() * sizeof(double), hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
    for (size_t s = 0; s < ns; ++s)
        for (int k = 0; k < 3; ++k) counts[3 * s + k] = red[s * RED_W + 7 + k];
}

template Params<double> Engine::params<double>();
template Params<float> Engine::params<float>();

}  // namespace d3
