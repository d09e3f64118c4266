// Host side of one D3 handle: parameters, device buffers, cell geometry (bins and stencils)
// and the launch sequence.  Replaces the reference's Device_Buffer (src/d3_buffer.cu:151-634)
// and the orchestration of src/d3.cu:130-226.
#pragma once
#include <algorithm>
#include <cstdint>
#include <string>
#include <vector>

#include "d3_engine.hpp"

namespace d3 {

struct Error {
    uint32_t code;
    std::string msg;
};

class Engine {
   public:
    Engine(const uint16_t *elements, uint64_t num_elements, uint64_t max_length, double cutoff, double cn_cutoff,
           double atm_cutoff, int damping, int functional, int precision);
    ~Engine();
    Engine(const Engine &) = delete;
    Engine &operator=(const Engine &) = delete;

    void set_params(const double p[7]);
    // One cell: coordinates in Angstrom (caller order), atomic numbers.
    void set_atoms(const double *coords_A, const uint16_t *Z, uint64_t n);
    void set_cell(const double cell_A[9]);
    // Several cells at once (replaces the single-cell state).
    void set_batch(uint64_t nsys, const uint64_t *counts, const double *coords_A, const uint16_t *Z,
                   const double *cells_A);
    void clear();
    // Enqueues the full pipeline on the stream (no synchronisation).
    void compute_device();
    // Synchronises and copies results: per-cell energy (eV) and stress (eV/A^3, 9 per cell),
    // forces (eV/A, caller order).  Throws Error on device status.
    void fetch(double *energies, double *forces, double *stresses);
    void fetch_atom_data(double *cn, double *dedcn, double *atom_energy);
    // Work counts of the last evaluation per cell: two-body ordered pairs, CN ordered pairs,
    // ATM triangles (each counted once).
    void fetch_counts(double *counts);

    uint64_t num_atoms() const { return n_; }
    uint64_t max_length() const { return max_len_; }
    int num_sys() const { return (int)sys_.size(); }
    hipStream_t stream() const { return stream_; }
    void set_stream(hipStream_t s) { stream_ = s ? s : own_stream_; }
    int precision() const { return precision_; }
    bool use_atm() const { return rc3_ > 0; }
    // Bin width (bohr) used for new cells; tunable.
    void set_bin_width(double w) { bin_width_ = w; geom_dirty_ = true; }
    double bin_width() const { return bin_width_; }
    void set_atm_blocks(int nb) { atm_blocks_req_ = nb; }
    // Quadruples the three-body neighbour capacity (after a D3X_ERR_OVERFLOW); kept neighbour
    // lists of the old capacity are rebuilt.
    void grow_atm_cap() {
        atm_cap_ = std::max(atm_cap_, 1024) * 4;
        nbr_valid_ = false;
    }

    // Neighbour state across evaluations (MD steps).  skin > 0 (bohr): stencils and three-body
    // forward lists cover every cutoff plus the skin, and an evaluation after set_atoms with the
    // same atoms and cell reuses the last bin order and lists while no atom has moved more than
    // skin / 2 from where it was at the last rebuild.  lists: -1 = forward lists when the skin is
    // set or centres are split into more work items than four per resident block, 0 = never,
    // 1 = always (within a 4 GiB budget).  Counters: evaluations that rebuilt / reused the neighbour state.
    // Smooth CN cutoff (the reference's SMOOTH_CUTOFF_CN build option): CN terms times
    // tanh^3(rc_CN - r).
    void set_cn_smooth(bool on) { cn_smooth_ = on; }
    bool cn_smooth() const { return cn_smooth_; }
    void set_skin(double skin);
    double skin() const { return skin_; }
    void set_neighbour_lists(int mode);
    int neighbour_lists() const { return fl_mode_; }
    void neighbour_stats(int &rebuilds, int &reuses) const {
        rebuilds = nbr_rebuilds_;
        reuses = nbr_reuses_;
    }
    int atm_cap() const { return atm_cap_; }
    const std::vector<SysDev> &sysdev() const { return sysdev_; }
    int stencil_count(int cut) const;

    // Per-phase timing with HIP events on the handle's stream (bin sort, CN, two-body, ATM,
    // chain rule, reduction).  enable resets the totals; phase_times harvests pending events and
    // returns the summed milliseconds per phase and the number of timed evaluations.
    static constexpr int NPHASE = 6;
    void set_timing(bool on);
    int phase_times(double ms[NPHASE]);

    // hipGraph mode: the first evaluation captures the whole launch sequence into a graph, later
    // ones replay it; a change of any kernel argument or launch shape re-captures.
    void set_graph(bool on);
    int graph_captures() const { return graph_captures_; }

    // Tuning parameters: bin width (bohr) and the three-body kernel's persistent grid size.
    // autotune() times full evaluations of the current input over a small candidate set, keeps
    // the fastest, and stores it in a JSON cache keyed by tune_key(); the cache is consulted
    // when geometry is built for an input whose key has an entry (unless set explicitly).
    void set_tuning(double bin_width, int atm_blocks);
    double autotune(const std::string &cache_path);
    std::string tune_key() const;
    int atm_blocks_request() const { return atm_blocks_req_; }
    // Three-body grid of the next evaluation: the request, else the device's resident count.
    int atm_blocks_effective() { return atm_blocks_req_ > 0 ? atm_blocks_req_ : atm_resident(); }
    int atm_var() {
        if (atm_var_ < 0)
            atm_var_ = atm_choose_variant(ne_, knr(), yrows_, precision_ == 0 ? sizeof(double) : sizeof(float));
        return atm_var_;
    }
    // Reference count the pair kernels are instantiated for: 0 = pairwise C6 interpolation
    // (tables whose CN references do not factorise per element).
    int knr() const { return c6_pairwise_ ? 0 : nref_; }
    bool c6_pairwise() const { return c6_pairwise_; }
    // A three-body variant this handle can run: its LDS fits, and no DR rows with pairwise C6.
    bool atm_variant_ok(int v) const {
        return v >= 0 && v <= 3 && !(c6_pairwise_ && (v & ATM_DR)) &&
               atm_lds_bytes(ne_, knr(), yrows_, precision_ == 0 ? sizeof(double) : sizeof(float), v) <= 65536;
    }

    // Three-body stand-in (modes of include/d3x_sur.h: 0 exact, 1 network with judge, 2 exact +
    // features).  The judge itself runs in the C API's evaluation (it needs host results):
    // judge_due() says whether this evaluation is judged, set_force_exact() runs the exact
    // kernel once, record_judgement() compares per-cell energies and falls back on failure.
    void sur_load(const float *params);
    void set_atm_mode(int mode, int judge_every, double tol_ev_per_atom);
    int mode_now() const { return (atm_mode_ == 1 && (fallen_back_ || force_exact_)) ? 0 : atm_mode_; }
    bool judge_due() const;
    void set_force_exact(bool on) { force_exact_ = on; }
    // per cell: |E_net - E_exact| / atoms against the energy tolerance and, when forces are given
    // (concatenated in cell order), the RMS over the cell's force components of F_net - F_exact
    // against the force tolerance (<= 0: not judged)
    bool record_judgement(const double *e_net, const double *e_exact, const double *f_net, const double *f_exact);
    void set_judge_force_tol(double tol) { judge_ftol_ = tol; }
    double last_force_deviation() const { return last_fdev_; }
    void count_surrogate_eval() { ++sur_evals_; }
    int fallbacks() const { return fallbacks_; }
    double last_deviation() const { return last_dev_; }
    void fetch_features(float *X);
    // u = J dir of the last evaluation's geometry (caller order; dir per A, atoms x 3; U atoms x 64).
    void sur_tangent(const double *dir, float *U);

   private:
    struct Cell {
        double H[9] = {0};  // bohr
        bool set = false;
        uint64_t n = 0;
    };
    template <class T>
    Params<T> params();
    template <class T>
    void run();
    template <class T>
    void enqueue(const Params<T> &P, bool reuse);
    void build_geometry();
    void upload_geometry();
    void alloc_atoms();
    void ensure_atm_scratch();
    void ensure_tb_scratch();
    double default_bin_width() const;
    void ensure_surrogate();
    int type_of(uint16_t Z) const;

    bool cn_smooth_ = false;
    // neighbour state (set_skin)
    double skin_ = 0.0;
    int fl_mode_ = -1;
    bool fl_use_ = false;           // forward lists in use for the next evaluation
    bool nbr_valid_ = false;        // bin order and lists of the last rebuild are usable
    bool nbr_moved_far_ = true;     // an atom moved more than skin / 2 since the last rebuild
    std::vector<double> h_pos_;     // bohr, caller order, of the last set_atoms
    std::vector<uint16_t> h_typ_;   // types of the last set_atoms
    std::vector<double> nbr_pos_;   // positions at the last rebuild
    std::vector<uint16_t> nbr_typ_;
    int nbr_rebuilds_ = 0, nbr_reuses_ = 0;
    bool nbr_fl_ = false;  // forward lists were built at the last rebuild
    int nbr_cap_ = 0;      // with this capacity
    void *d_fl = nullptr, *d_fl_n = nullptr, *d_shift = nullptr;
    size_t fl_alloc_ = 0;
    bool reuse_now() const { return skin_ > 0 && nbr_valid_ && !nbr_moved_far_ && sys_.size() == 1; }

    // three-body stand-in state
    int atm_mode_ = 0, judge_every_ = 0, fallbacks_ = 0;
    double judge_tol_ = 1e-3, last_dev_ = -1.0, judge_ftol_ = 5e-3, last_fdev_ = -1.0;
    bool fallen_back_ = false, force_exact_ = false, sur_dirty_ = false, feat_valid_ = false;
    uint64_t sur_evals_ = 0;
    std::vector<float> h_sur_P_;
    void *d_type_cls = nullptr, *d_feat = nullptr, *d_sur_G = nullptr, *d_sur_E = nullptr, *d_sur_cls = nullptr,
         *d_sur_P = nullptr, *d_sur_PK = nullptr, *d_tan_dir = nullptr, *d_tan_U = nullptr, *d_atm_roff = nullptr,
         *d_atm_eatc = nullptr, *d_tb_part = nullptr, *d_sur_q = nullptr, *d_sur_dcn = nullptr;

    // configuration
    std::vector<uint16_t> elems_;  // sorted unique atomic numbers of the handle
    int ne_ = 0, nref_ = 0;
    std::vector<int> atm_roff_;  // three-body j-tile row offsets per type (Params::atm_roff)
    int yrows_ = 0;
    // CN references that do not factorise per element (or D3_C6_PAIRWISE=1): C6 interpolated per
    // pair from h_c6g_ = {C6ref, CN_A ref, CN_B ref}[ne][ne][nref][nref] (Params::c6g)
    bool c6_pairwise_ = false;
    std::vector<double> h_c6g_;
    uint64_t max_len_ = 0;
    double rc2_, rccn_, rc3_;
    int damping_, functional_, precision_;
    double fp_[7];
    double bin_width_;
    int atm_blocks_req_ = 0;
    std::string table_name_;

    // state
    uint64_t n_ = 0;
    std::vector<Cell> sys_;
    bool atoms_set_ = false, geom_dirty_ = true, tables_dirty_ = true;
    std::vector<SysDev> sysdev_;
    std::vector<StEntry> st_d_;
    std::vector<double> st_D_;
    std::vector<int> chunk_b_, chunk_e_, chunk_s_, sys_chunk_off_;
    int nbins_ = 0;
    std::vector<uint16_t> h_types_;
    std::vector<double> h_tab_;  // type tables in double (converted to T at upload)

    // device
    hipStream_t own_stream_ = nullptr, stream_ = nullptr;
    void *d_pos_in = nullptr, *d_type_in = nullptr, *d_sys_in = nullptr;
    void *d_key = nullptr, *d_perm = nullptr, *d_perm2 = nullptr, *d_wrapped = nullptr;
    void *d_lx = nullptr, *d_ly = nullptr, *d_lz = nullptr, *d_type_s = nullptr, *d_orig_s = nullptr,
         *d_bin_s = nullptr, *d_sys_s = nullptr;
    void *d_cn = nullptr, *d_w = nullptr, *d_dw = nullptr, *d_e = nullptr, *d_g = nullptr, *d_f = nullptr,
         *d_vir = nullptr, *d_npair = nullptr;
    void *d_out_f = nullptr, *d_out_e = nullptr, *d_out_cn = nullptr, *d_out_g = nullptr;
    void *d_bin_count = nullptr, *d_bin_start = nullptr, *d_cursor = nullptr;
    size_t bins_cap_ = 0;
    void *d_sys = nullptr, *d_st_d = nullptr, *d_st_D = nullptr;
    size_t sys_cap_ = 0, st_cap_ = 0;
    void *d_chunk_b = nullptr, *d_chunk_e = nullptr, *d_chunk_s = nullptr, *d_sys_chunk = nullptr,
         *d_partial = nullptr, *d_out_sys = nullptr;
    size_t chunk_cap_ = 0, sysred_cap_ = 0;
    void *d_tab = nullptr;
    size_t tab_off_[7] = {0};  // c6 | cnref | rcov | q4 | rdamp | r0 | c6g (pairwise only)
    void *d_status = nullptr;
    void *d_atm_i32 = nullptr, *d_atm_T = nullptr, *d_atm_acc = nullptr;
    void *d_atm_pa = nullptr;  // per-atom ATM results: fa[3N] | ga[N] | eas[N] | eat[N] (double)
    int atm_blocks_ = 0, atm_cap_ = 0;
    int atm_resident_ = 0;  // three-body blocks resident on the device at once (0: not queried)
    int atm_var_ = -1;      // three-body kernel variant, ATM_KL | ATM_DR (-1: not chosen yet)
    void set_atm_variant(int v) {  // a fixed variant (autotuner, tuning cache); resident count follows
        atm_var_ = v;
        atm_resident_ = 0;
    }
    int atm_resident() {
        if (atm_resident_ <= 0)
            atm_resident_ =
                atm_resident_blocks(ne_, knr(), yrows_, precision_ == 0 ? sizeof(double) : sizeof(float), atm_var());
        return atm_resident_;
    }
    size_t atm_alloc_entries_ = 0;  // allocated scratch entries (>= atm_blocks_ * atm_cap_)
    int atm_alloc_blocks_ = 0;      // blocks the scratch was allocated for (>= atm_blocks_)
    int atm_chunks_ = 1;            // three-body work items per centre (Params::atm_chunks)
    size_t atm_eatc_alloc_ = 0;     // doubles in d_atm_eatc (per-item centre energies)
    int tb_parts_ = 1;              // two-body waves per centre (Params::tb_parts)
    size_t tb_part_alloc_ = 0;      // doubles in d_tb_part

    bool tuning_explicit_ = false;  // set_tuning / autotune fixed the parameters
    std::string tune_cache_;        // cache consulted at geometry build (empty: none)
    void apply_tune_cache();
    double timed_evaluation();

    bool graph_mode_ = false;
    hipGraphExec_t gexec_ = nullptr;
    std::vector<unsigned char> gsig_;
    int graph_captures_ = 0;

    bool timing_ = false, timing_pending_ = false;
    hipEvent_t ev_[NPHASE + 1] = {};
    double phase_ms_[NPHASE] = {0};
    int timed_runs_ = 0;
    void mark(int k);
    void harvest();
};

void launch_fill_u32(unsigned *p, int n, unsigned v, hipStream_t s);

// Thread-local last error message of the C API.
void set_last_error(const std::string &msg);
const char *last_error();

// Functional parameter table {s6, s8_zero, s8_bj, sr6, sr8, a1, a2}; functional_t order.
bool functional_params(int functional, double out[7]);

}  // namespace d3
