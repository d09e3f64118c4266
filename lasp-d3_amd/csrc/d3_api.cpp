// C ABI of the D3 library: the reference's d3.h entry points (src/d3.cu:14-252) and the d3x.h
// extensions, on top of d3::Engine.  Errors never escape as C++ exceptions: they are printed to
// stderr (as the reference does), recorded for d3x_last_error / d3x_last_status, and turned into
// NULL / non-zero returns.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "../../include/d3x.h"
#include "../../include/d3x_sur.h"
#include "d3_host.hpp"
#include "d3_tables.hpp"

using d3::Engine;
using d3::Error;

namespace {

struct Handle {
    std::unique_ptr<Engine> eng;
    uint32_t status = 0;
};

Handle *H(d3_handle_t *h) { return static_cast<Handle *>(h); }

// Default arithmetic of handles made through the reference-compatible init functions: FP32
// (pair and triplet math in float with compensated sums, as the reference computes in float,
// src/d3_types.h:7) unless D3_PRECISION=fp64 is set in the environment.
int default_precision() {
    const char *p = std::getenv("D3_PRECISION");
    return (p && (std::strcmp(p, "fp64") == 0 || std::strcmp(p, "FP64") == 0)) ? D3X_FP64 : D3X_FP32;
}

void report(const char *where, const std::string &msg) {
    d3::set_last_error(msg);
    std::fprintf(stderr, "Error: %s: %s\n", where, msg.c_str());
}

// Runs fn, converting exceptions into a status code stored on the handle.
template <class F>
uint32_t guarded(Handle *h, const char *where, F &&fn) {
    uint32_t st = 0;
    try {
        fn();
    } catch (const Error &e) {
        st = e.code ? e.code : D3X_ERR_GENERIC;
        report(where, e.msg);
    } catch (const std::bad_alloc &) {
        st = D3X_ERR_GENERIC;
        report(where, "out of host memory");
    } catch (const std::exception &e) {
        st = D3X_ERR_GENERIC;
        report(where, e.what());
    }
    if (h) h->status = st;
    return st;
}

void evaluate_once(Engine &e, double *energies, double *forces, double *stresses);

// Full evaluation with results copied to the host.  With the three-body stand-in active, a
// judged evaluation runs the exact kernel and then the network: on agreement (per cell, |dE| /
// atoms and the RMS force deviation within their tolerances) the network's results are returned,
// otherwise the exact kernel runs once more (so the device holds its per-atom data), its results
// are returned, and the handle stays on the exact kernel.
void evaluate(Engine &e, double *energies, double *forces, double *stresses) {
    if (e.mode_now() != D3X_ATM_SURROGATE) {
        evaluate_once(e, energies, forces, stresses);
        return;
    }
    const bool judge = e.judge_due();
    e.count_surrogate_eval();
    if (!judge) {
        evaluate_once(e, energies, forces, stresses);
        return;
    }
    const size_t ns = (size_t)e.num_sys(), n3 = 3 * (size_t)e.num_atoms();
    std::vector<double> e_net(ns), e_exact(ns), f_net(n3), f_exact(n3), s_net(9 * ns);
    e.set_force_exact(true);
    try {
        evaluate_once(e, e_exact.data(), f_exact.data(), nullptr);
    } catch (...) {
        e.set_force_exact(false);
        throw;
    }
    e.set_force_exact(false);
    evaluate_once(e, e_net.data(), f_net.data(), s_net.data());
    if (e.record_judgement(e_net.data(), e_exact.data(), f_net.data(), f_exact.data())) {
        if (energies) std::copy(e_net.begin(), e_net.end(), energies);
        if (forces) std::copy(f_net.begin(), f_net.end(), forces);
        if (stresses) std::copy(s_net.begin(), s_net.end(), stresses);
    } else {
        evaluate_once(e, energies, forces, stresses);  // fallen back: the exact kernel's results
    }
}

// One evaluation; a three-body neighbour-list overflow is retried with a larger capacity
// (capacity grows x4, at most 4 times).
void evaluate_once(Engine &e, double *energies, double *forces, double *stresses) {
    for (int attempt = 0;; ++attempt) {
        e.compute_device();
        try {
            e.fetch(energies, forces, stresses);
            return;
        } catch (const Error &err) {
            if ((err.code & D3X_ERR_OVERFLOW) && attempt < 4) {
                e.grow_atm_cap();
                continue;
            }
            throw;
        }
    }
}

d3_handle_t *make_handle(const uint16_t *elements, uint64_t ne, uint64_t max_length, double rc, double rccn,
                         double rc3, int damping, int functional, int precision) {
    Handle *h = nullptr;
    uint32_t st = guarded(nullptr, "init_d3_handle", [&] {
        std::unique_ptr<Handle> hh(new Handle);
        hh->eng.reset(new Engine(elements, ne, max_length, rc, rccn, rc3, damping, functional, precision));
        h = hh.release();
    });
    return st ? nullptr : h;
}

// Options of reference-compatible handles from the environment: D3_GRAPH=1 evaluates through a
// captured hipGraph, D3_SMOOTH_CN=1 the smooth CN cutoff, D3_SKIN=<bohr> keeps the neighbour
// state across MD steps.
d3_handle_t *with_env_graph(d3_handle_t *h) {
    if (!h) return h;
    const char *g = std::getenv("D3_GRAPH");
    if (g && std::strcmp(g, "1") == 0) H(h)->eng->set_graph(true);
    const char *sm = std::getenv("D3_SMOOTH_CN");
    if (sm && std::strcmp(sm, "1") == 0) H(h)->eng->set_cn_smooth(true);
    const char *sk = std::getenv("D3_SKIN");
    if (sk && *sk) {
        const double v = std::atof(sk);
        if (v > 0 && guarded(H(h), "D3_SKIN", [&] { H(h)->eng->set_skin(v); })) {
            delete H(h);
            return nullptr;
        }
    }
    return h;
}

}  // namespace

extern "C" {

d3_handle_t *init_d3_handle(uint16_t *elements, uint64_t num_elements, uint64_t max_length, real_t cutoff_radius,
                            real_t coordination_number_cutoff, damping_type_t damping_type,
                            functional_t functional_type) {
    return with_env_graph(make_handle(elements, num_elements, max_length, cutoff_radius, coordination_number_cutoff,
                                      0.0, (int)damping_type, (int)functional_type, default_precision()));
}

d3_handle_t *init_d3_handle_with_atm(uint16_t *elements, uint64_t length_elements, uint64_t max_length,
                                     real_t cutoff_radius, real_t coordination_number_cutoff, real_t atm_cutoff,
                                     damping_type_t damping_type, functional_t functional_type) {
    if (!(atm_cutoff > 0)) {
        report("init_d3_handle_with_atm", "atm_cutoff must be positive");
        return nullptr;
    }
    return with_env_graph(make_handle(elements, length_elements, max_length, cutoff_radius,
                                      coordination_number_cutoff, atm_cutoff, (int)damping_type, (int)functional_type,
                                      default_precision()));
}

void set_atoms(d3_handle_t *handle, real_t *coords, uint16_t *elements, uint64_t length) {
    if (!handle) {
        report("set_atoms", "handle is NULL");
        return;
    }
    guarded(H(handle), "set_atoms", [&] {
        std::vector<double> c(3 * length);
        for (uint64_t k = 0; k < 3 * length; ++k) c[k] = coords[k];
        H(handle)->eng->set_atoms(c.data(), elements, length);
    });
}

void set_cell(d3_handle_t *handle, real_t cell[3][3]) {
    if (!handle) {
        report("set_cell", "handle is NULL");
        return;
    }
    guarded(H(handle), "set_cell", [&] {
        double c[9];
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) c[a * 3 + b] = cell[a][b];
        H(handle)->eng->set_cell(c);
    });
}

void free_d3_handle(d3_handle_t *handle) {
    if (!handle) {
        report("free_d3_handle", "handle is NULL");
        return;
    }
    delete H(handle);
}

void clear_d3_handle(d3_handle_t *handle) {
    if (!handle) {
        report("clear_d3_handle", "handle is NULL");
        return;
    }
    guarded(H(handle), "clear_d3_handle", [&] { H(handle)->eng->clear(); });
}

uint16_t compute_dispersion_energy_from_handle_status(d3_handle_t *handle, real_t *energy, real_t *force,
                                                      real_t *stress) {
    if (!handle) {
        report("compute_dispersion_energy_from_handle", "handle is NULL");
        return 1;
    }
    Handle *h = H(handle);
    uint32_t st = guarded(h, "compute_dispersion_energy_from_handle", [&] {
        Engine &e = *h->eng;
        if (e.num_sys() != 1) throw Error{D3X_ERR_GENERIC | D3X_ERR_STATE, "handle holds a batch; call set_atoms first"};
        const uint64_t n = e.num_atoms();
        std::vector<double> f(3 * n);
        double E, S[9];
        evaluate(e, &E, f.data(), S);
        if (energy) *energy = (real_t)E;
        if (force)
            for (uint64_t k = 0; k < 3 * n; ++k) force[k] = (real_t)f[k];
        if (stress)
            for (int k = 0; k < 9; ++k) stress[k] = (real_t)S[k];
    });
    return st ? 1 : 0;
}

void compute_dispersion_energy_from_handle(d3_handle_t *handle, real_t *energy, real_t *force, real_t *stress) {
    uint16_t st = compute_dispersion_energy_from_handle_status(handle, energy, force, stress);
    if (st != 0) std::fprintf(stderr, "Error: compute_dispersion_energy_from_handle failed with status %d\n", st);
}

void compute_dispersion_energy(real_t atoms[][3], uint16_t *elements, uint64_t num_atoms, real_t cell[3][3],
                               real_t cutoff_radius, real_t CN_cutoff_radius, damping_type_t damping_type,
                               functional_t functional_type, real_t *energy, real_t *force, real_t *stress) {
    // as src/d3.cu:236-252: a handle over the structure's own elements, used once
    d3_handle_t *h = init_d3_handle(elements, num_atoms, num_atoms ? num_atoms : 1, cutoff_radius, CN_cutoff_radius,
                                    damping_type, functional_type);
    if (!h) return;
    set_atoms(h, (real_t *)atoms, elements, num_atoms);
    set_cell(h, cell);
    compute_dispersion_energy_from_handle(h, energy, force, stress);
    free_d3_handle(h);
}

// ------------------------------------------------------------------------------------------ d3x

d3_handle_t *d3x_init(const uint16_t *elements, uint64_t num_elements, uint64_t max_length, double cutoff,
                      double cn_cutoff, double atm_cutoff, int damping, int functional, int precision) {
    return make_handle(elements, num_elements, max_length, cutoff, cn_cutoff, atm_cutoff, damping, functional,
                       precision);
}

int d3x_set_functional_params(d3_handle_t *handle, const double params[7]) {
    if (!handle || !params) return (int)(D3X_ERR_GENERIC | D3X_ERR_INPUT);
    return (int)guarded(H(handle), "d3x_set_functional_params", [&] { H(handle)->eng->set_params(params); });
}

int d3x_set_atoms_f64(d3_handle_t *handle, const double *coords, const uint16_t *elements, uint64_t length) {
    if (!handle) return (int)(D3X_ERR_GENERIC | D3X_ERR_INPUT);
    return (int)guarded(H(handle), "d3x_set_atoms_f64", [&] { H(handle)->eng->set_atoms(coords, elements, length); });
}

int d3x_set_cell_f64(d3_handle_t *handle, const double cell[9]) {
    if (!handle || !cell) return (int)(D3X_ERR_GENERIC | D3X_ERR_INPUT);
    return (int)guarded(H(handle), "d3x_set_cell_f64", [&] { H(handle)->eng->set_cell(cell); });
}

int d3x_compute_f64(d3_handle_t *handle, double *energy, double *forces, double *stress) {
    if (!handle) return (int)(D3X_ERR_GENERIC | D3X_ERR_INPUT);
    return (int)guarded(H(handle), "d3x_compute_f64", [&] {
        Engine &e = *H(handle)->eng;
        if (e.num_sys() != 1) throw Error{D3X_ERR_GENERIC | D3X_ERR_STATE, "handle holds a batch; use d3x_compute_batch_f64"};
        evaluate(e, energy, forces, stress);
    });
}

int d3x_get_atom_data(d3_handle_t *handle, double *cn, double *dedcn, double *atom_energy) {
    if (!handle) return (int)(D3X_ERR_GENERIC | D3X_ERR_INPUT);
    return (int)guarded(H(handle), "d3x_get_atom_data",
                        [&] { H(handle)->eng->fetch_atom_data(cn, dedcn, atom_energy); });
}

int d3x_compute_device(d3_handle_t *handle) {
    if (!handle) return (int)(D3X_ERR_GENERIC | D3X_ERR_INPUT);
    return (int)guarded(H(handle), "d3x_compute_device", [&] { H(handle)->eng->compute_device(); });
}

int d3x_fetch_results(d3_handle_t *handle, double *energy, double *forces, double *stress) {
    if (!handle) return (int)(D3X_ERR_GENERIC | D3X_ERR_INPUT);
    return (int)guarded(H(handle), "d3x_fetch_results", [&] { H(handle)->eng->fetch(energy, forces, stress); });
}

int d3x_get_counts(d3_handle_t *handle, double *counts) {
    if (!handle || !counts) return (int)(D3X_ERR_GENERIC | D3X_ERR_INPUT);
    return (int)guarded(H(handle), "d3x_get_counts", [&] { H(handle)->eng->fetch_counts(counts); });
}

int d3x_set_timing(d3_handle_t *handle, int enable) {
    if (!handle) return (int)(D3X_ERR_GENERIC | D3X_ERR_INPUT);
    return (int)guarded(H(handle), "d3x_set_timing", [&] { H(handle)->eng->set_timing(enable != 0); });
}

int d3x_phase_times(d3_handle_t *handle, double ms[6]) {
    if (!handle || !ms) return -1;
    int n = -1;
    uint32_t st = guarded(H(handle), "d3x_phase_times", [&] { n = H(handle)->eng->phase_times(ms); });
    return st ? -1 : n;
}

int d3x_set_tuning(d3_handle_t *handle, double bin_width_bohr, int atm_blocks) {
    if (!handle) return (int)(D3X_ERR_GENERIC | D3X_ERR_INPUT);
    return (int)guarded(H(handle), "d3x_set_tuning",
                        [&] { H(handle)->eng->set_tuning(bin_width_bohr, atm_blocks); });
}

int d3x_get_tuning(d3_handle_t *handle, double *bin_width_bohr, int *atm_blocks) {
    if (!handle) return (int)(D3X_ERR_GENERIC | D3X_ERR_INPUT);
    if (bin_width_bohr) *bin_width_bohr = H(handle)->eng->bin_width();
    if (atm_blocks) *atm_blocks = H(handle)->eng->atm_blocks_effective();
    return 0;
}

int d3x_set_cn_smooth(d3_handle_t *handle, int on) {
    if (!handle) return (int)(D3X_ERR_GENERIC | D3X_ERR_INPUT);
    H(handle)->eng->set_cn_smooth(on != 0);
    return 0;
}

int d3x_set_skin(d3_handle_t *handle, double skin_bohr) {
    if (!handle) return (int)(D3X_ERR_GENERIC | D3X_ERR_INPUT);
    return (int)guarded(H(handle), "d3x_set_skin", [&] { H(handle)->eng->set_skin(skin_bohr); });
}

int d3x_set_neighbour_lists(d3_handle_t *handle, int mode) {
    if (!handle) return (int)(D3X_ERR_GENERIC | D3X_ERR_INPUT);
    return (int)guarded(H(handle), "d3x_set_neighbour_lists", [&] { H(handle)->eng->set_neighbour_lists(mode); });
}

int d3x_neighbour_stats(d3_handle_t *handle, int *rebuilds, int *reuses) {
    if (!handle) return (int)(D3X_ERR_GENERIC | D3X_ERR_INPUT);
    int a = 0, b = 0;
    H(handle)->eng->neighbour_stats(a, b);
    if (rebuilds) *rebuilds = a;
    if (reuses) *reuses = b;
    return 0;
}

int d3x_get_atm_capacity(d3_handle_t *handle) {
    if (!handle) return -1;
    return H(handle)->eng->atm_cap();
}

int d3x_get_atm_variant(d3_handle_t *handle) {
    if (!handle) return -1;
    int v = -1;
    uint32_t st = guarded(H(handle), "d3x_get_atm_variant", [&] { v = H(handle)->eng->atm_var(); });
    return st ? -1 : v;
}

int d3x_get_c6_pairwise(d3_handle_t *handle) {
    if (!handle) return -1;
    return H(handle)->eng->c6_pairwise() ? 1 : 0;
}

double d3x_autotune(d3_handle_t *handle, const char *cache_path) {
    if (!handle) return -1.0;
    double best = -1.0;
    uint32_t st = guarded(H(handle), "d3x_autotune",
                          [&] { best = H(handle)->eng->autotune(cache_path ? cache_path : ""); });
    return st ? -1.0 : best;
}

int d3x_sur_load(d3_handle_t *handle, const float *params) {
    if (!handle) return (int)(D3X_ERR_GENERIC | D3X_ERR_INPUT);
    return (int)guarded(H(handle), "d3x_sur_load", [&] { H(handle)->eng->sur_load(params); });
}

int d3x_sur_load_file(d3_handle_t *handle, const char *path) {
    if (!handle) return (int)(D3X_ERR_GENERIC | D3X_ERR_INPUT);
    return (int)guarded(H(handle), "d3x_sur_load_file", [&] {
        if (!path) throw Error{D3X_ERR_GENERIC | D3X_ERR_INPUT, "NULL path"};
        FILE *f = std::fopen(path, "rb");
        if (!f) throw Error{D3X_ERR_GENERIC | D3X_ERR_INPUT, std::string("cannot open ") + path};
        char magic[8] = {0};
        uint32_t count = 0;
        std::vector<float> p(D3X_SUR_NPARAMS);
        const bool ok = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, "D3SUR001", 8) == 0 &&
                        std::fread(&count, 4, 1, f) == 1 && count == (uint32_t)D3X_SUR_NPARAMS &&
                        std::fread(p.data(), sizeof(float), p.size(), f) == p.size();
        std::fclose(f);
        if (!ok) throw Error{D3X_ERR_GENERIC | D3X_ERR_INPUT, std::string("not a network parameter file: ") + path};
        H(handle)->eng->sur_load(p.data());
    });
}

int d3x_set_atm_mode(d3_handle_t *handle, int mode, int judge_every, double tol_ev_per_atom) {
    if (!handle) return (int)(D3X_ERR_GENERIC | D3X_ERR_INPUT);
    return (int)guarded(H(handle), "d3x_set_atm_mode",
                        [&] { H(handle)->eng->set_atm_mode(mode, judge_every, tol_ev_per_atom); });
}

int d3x_sur_state(d3_handle_t *handle, int *mode, int *fallbacks, double *last_dev_ev_per_atom) {
    if (!handle) return (int)(D3X_ERR_GENERIC | D3X_ERR_INPUT);
    const Engine &e = *H(handle)->eng;
    if (mode) *mode = e.mode_now();
    if (fallbacks) *fallbacks = e.fallbacks();
    if (last_dev_ev_per_atom) *last_dev_ev_per_atom = e.last_deviation();
    return 0;
}

int d3x_sur_set_force_tol(d3_handle_t *handle, double tol_rms_ev_per_A) {
    if (!handle) return (int)(D3X_ERR_GENERIC | D3X_ERR_INPUT);
    return (int)guarded(H(handle), "d3x_sur_set_force_tol", [&] {
        if (!std::isfinite(tol_rms_ev_per_A)) throw Error{D3X_ERR_GENERIC | D3X_ERR_INPUT, "tolerance must be finite"};
        H(handle)->eng->set_judge_force_tol(tol_rms_ev_per_A);
    });
}

double d3x_sur_last_force_dev(d3_handle_t *handle) { return handle ? H(handle)->eng->last_force_deviation() : -1.0; }

int d3x_sur_features(d3_handle_t *handle, float *features) {
    if (!handle || !features) return (int)(D3X_ERR_GENERIC | D3X_ERR_INPUT);
    return (int)guarded(H(handle), "d3x_sur_features", [&] { H(handle)->eng->fetch_features(features); });
}

int d3x_sur_tangent(d3_handle_t *handle, const double *dir, float *U) {
    if (!handle || !dir || !U) return (int)(D3X_ERR_GENERIC | D3X_ERR_INPUT);
    return (int)guarded(H(handle), "d3x_sur_tangent", [&] { H(handle)->eng->sur_tangent(dir, U); });
}

int d3x_set_graph(d3_handle_t *handle, int enable) {
    if (!handle) return (int)(D3X_ERR_GENERIC | D3X_ERR_INPUT);
    return (int)guarded(H(handle), "d3x_set_graph", [&] { H(handle)->eng->set_graph(enable != 0); });
}

void *d3x_get_stream(d3_handle_t *handle) { return handle ? (void *)H(handle)->eng->stream() : nullptr; }

int d3x_set_stream(d3_handle_t *handle, void *stream) {
    if (!handle) return (int)(D3X_ERR_GENERIC | D3X_ERR_INPUT);
    H(handle)->eng->set_stream((hipStream_t)stream);
    return 0;
}

int d3x_compute_batch_f64(d3_handle_t *handle, uint64_t n_struct, const uint64_t *atom_counts,
                          const double *coords, const uint16_t *elements, const double *cells, double *energies,
                          double *forces, double *stresses) {
    if (!handle) return (int)(D3X_ERR_GENERIC | D3X_ERR_INPUT);
    return (int)guarded(H(handle), "d3x_compute_batch_f64", [&] {
        Engine &e = *H(handle)->eng;
        e.set_batch(n_struct, atom_counts, coords, elements, cells);
        evaluate(e, energies, forces, stresses);
    });
}

uint32_t d3x_last_status(d3_handle_t *handle) { return handle ? H(handle)->status : D3X_ERR_GENERIC; }

const char *d3x_last_error(void) { return d3::last_error(); }

int d3x_table_load(const char *path) {
    try {
        if (!path || !*path || std::strcmp(path, "synthetic") == 0)
            d3::set_active_table(d3::make_synthetic_table(false));
        else if (std::strcmp(path, "synthetic-extended") == 0)
            d3::set_active_table(d3::make_synthetic_table(true));
        else
            d3::set_active_table(d3::load_table_file(path));
        return 0;
    } catch (const std::exception &e) {
        report("d3x_table_load", e.what());
        return (int)(D3X_ERR_GENERIC | D3X_ERR_INPUT);
    }
}

const char *d3x_table_name(void) { return d3::active_table().name.c_str(); }

int d3x_table_shape(int *nel, int *nref, int *zoff, int *zmax) {
    const d3::Table &t = d3::active_table();
    if (nel) *nel = t.nel;
    if (nref) *nref = t.nref;
    if (zoff) *zoff = t.zoff;
    if (zmax) *zmax = t.zmax;
    return 0;
}

int d3x_table_export(float *c6ab, float *r0ab, float *rcov, float *r2r4) {
    const d3::Table &t = d3::active_table();
    if (c6ab) std::memcpy(c6ab, t.c6ab.data(), t.c6ab.size() * sizeof(float));
    if (r0ab) std::memcpy(r0ab, t.r0ab.data(), t.r0ab.size() * sizeof(float));
    if (rcov) std::memcpy(rcov, t.rcov.data(), t.rcov.size() * sizeof(float));
    if (r2r4) std::memcpy(r2r4, t.r2r4.data(), t.r2r4.size() * sizeof(float));
    return 0;
}

}  // extern "C"
