// d3_cpp: command-line driver over the C API, a drop-in for the reference's executable
// (src/main.cpp:10-114, with the .arc reader of src/parser.cpp:73-176 and the cell
// construction of src/modules.h:233-278).  It reads LASP .arc files and extended-XYZ files,
// evaluates every structure through init_d3_handle[_with_atm] / set_atoms / set_cell /
// compute_dispersion_energy_from_handle, and prints the reference's report lines (structure
// count, cell, elapsed time, energy, per-atom forces, force sum, stress).
//
// Usage: d3_cpp [--use-atm] [--bj] [--fp32 | --fp64] [--cutoff R] [--cn-cutoff R] [--atm-cutoff R] FILE...
// Defaults as in the reference's driver: cutoff 60 bohr, CN and ATM cutoffs 40 bohr, zero
// damping, PBE, the library's default FP32 arithmetic unless --fp64 (or D3_PRECISION=fp64).
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/d3.h"

namespace {

const char *const SYMBOLS[] = {
    "H",  "He", "Li", "Be", "B",  "C",  "N",  "O",  "F",  "Ne", "Na", "Mg", "Al", "Si", "P",  "S",  "Cl", "Ar",
    "K",  "Ca", "Sc", "Ti", "V",  "Cr", "Mn", "Fe", "Co", "Ni", "Cu", "Zn", "Ga", "Ge", "As", "Se", "Br", "Kr",
    "Rb", "Sr", "Y",  "Zr", "Nb", "Mo", "Tc", "Ru", "Rh", "Pd", "Ag", "Cd", "In", "Sn", "Sb", "Te", "I",  "Xe",
    "Cs", "Ba", "La", "Ce", "Pr", "Nd", "Pm", "Sm", "Eu", "Gd", "Tb", "Dy", "Ho", "Er", "Tm", "Yb", "Lu", "Hf",
    "Ta", "W",  "Re", "Os", "Ir", "Pt", "Au", "Hg", "Tl", "Pb", "Bi", "Po", "At", "Rn", "Fr", "Ra", "Ac", "Th",
    "Pa", "U",  "Np", "Pu", "Am", "Cm", "Bk", "Cf", "Es", "Fm", "Md", "No", "Lr", "Rf", "Db", "Sg", "Bh", "Hs",
    "Mt", "Ds", "Rg", "Cn", "Nh", "Fl", "Mc", "Lv", "Ts", "Og"};

// Atomic number of an element symbol (0 if unknown).
int atomic_number(const std::string &s) {
    for (int z = 0; z < 118; ++z)
        if (s == SYMBOLS[z]) return z + 1;
    return 0;
}

struct Structure {
    std::vector<uint16_t> Z;
    std::vector<float> pos;  // Angstrom, atoms x 3
    float cell[3][3] = {{0}};
    bool has_cell = false;
};

std::vector<std::string> split(const std::string &line) {
    std::istringstream in(line);
    std::vector<std::string> t;
    std::string w;
    while (in >> w) t.push_back(w);
    return t;
}

// Lattice vectors (rows) from lengths and angles in degrees: a along x, b in the xy plane.
void cell_from_parameters(double a, double b, double c, double al, double be, double ga, float out[3][3]) {
    const double d = std::acos(-1.0) / 180.0;
    const double ca = std::cos(al * d), cb = std::cos(be * d), cg = std::cos(ga * d), sg = std::sin(ga * d);
    const double v[3][3] = {{a, 0.0, 0.0},
                            {b * cg, b * sg, 0.0},
                            {c * cb, c * (ca - cb * cg) / sg,
                             c * std::sqrt(std::max(0.0, 1.0 - ca * ca - cb * cb - cg * cg + 2.0 * ca * cb * cg)) / sg}};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) out[i][j] = (float)v[i][j];
}

// LASP .arc: a structure starts at a line "Energy <n> <x> <energy> [symmetry]", its cell is the
// line "PBC a b c alpha beta gamma", its atoms are lines of exactly 10 fields starting with an
// element symbol followed by x y z (Angstrom).  Other lines are reported and skipped.
bool read_arc(const std::string &path, std::vector<Structure> &out) {
    std::ifstream f(path);
    if (!f) return false;
    std::string line;
    bool open = false;
    while (std::getline(f, line)) {
        const std::vector<std::string> t = split(line);
        try {
            if (t.size() == 10 && atomic_number(t[0]) > 0) {
                if (open) {
                    // parse all three coordinates before storing anything, so that a bad number
                    // leaves the structure's Z and pos the same length
                    const float x = std::stof(t[1]), y = std::stof(t[2]), z = std::stof(t[3]);
                    out.back().Z.push_back((uint16_t)atomic_number(t[0]));
                    out.back().pos.insert(out.back().pos.end(), {x, y, z});
                }
                continue;
            }
            if (t.size() == 7 && t[0] == "PBC") {
                if (open) {
                    cell_from_parameters(std::stod(t[1]), std::stod(t[2]), std::stod(t[3]), std::stod(t[4]),
                                         std::stod(t[5]), std::stod(t[6]), out.back().cell);
                    out.back().has_cell = true;
                }
                continue;
            }
            if ((t.size() == 4 || t.size() == 5) && t[0] == "Energy") {
                (void)std::stod(t[3]);
                out.emplace_back();
                open = true;
                continue;
            }
        } catch (const std::exception &) {
        }
        if (t.size() == 10 && atomic_number(t[0]) == 0) std::cerr << "Element not found: " << t[0] << std::endl;
        std::cerr << "Invalid line: " << line << std::endl;
    }
    return !out.empty();
}

// Extended XYZ: blocks of "<count>", a comment line with Lattice="a1 a2 a3 b1 b2 b3 c1 c2 c3"
// (rows = lattice vectors) and optionally Properties=name:type:cols:..., then <count> atom lines
// (species and pos columns located from Properties, default species x y z).
bool read_extxyz(const std::string &path, std::vector<Structure> &out) {
    std::ifstream f(path);
    if (!f) return false;
    std::string line;
    while (std::getline(f, line)) {
        const std::vector<std::string> head = split(line);
        if (head.empty()) continue;
        long n = 0;
        try {
            n = std::stol(head[0]);
        } catch (const std::exception &) {
            return false;
        }
        std::string comment;
        if (n < 0 || !std::getline(f, comment)) return false;
        Structure s;
        const size_t lp = comment.find("Lattice=\"");
        if (lp != std::string::npos) {
            const size_t e = comment.find('"', lp + 9);
            std::istringstream in(comment.substr(lp + 9, e - lp - 9));
            double v[9];
            for (double &x : v)
                if (!(in >> x)) return false;
            for (int i = 0; i < 9; ++i) s.cell[i / 3][i % 3] = (float)v[i];
            s.has_cell = true;
        }
        int col_species = 0, col_pos = 1;
        const size_t pp = comment.find("Properties=");
        if (pp != std::string::npos) {
            std::string spec = comment.substr(pp + 11);
            spec = spec.substr(0, spec.find(' '));
            std::vector<std::string> parts;
            std::stringstream ss(spec);
            std::string p;
            while (std::getline(ss, p, ':')) parts.push_back(p);
            int col = 0;
            for (size_t k = 0; k + 2 < parts.size(); k += 3) {
                const int w = std::atoi(parts[k + 2].c_str());
                if (parts[k] == "species") col_species = col;
                if (parts[k] == "pos") col_pos = col;
                col += w;
            }
        }
        for (long a = 0; a < n; ++a) {
            if (!std::getline(f, line)) return false;
            const std::vector<std::string> t = split(line);
            if ((int)t.size() < std::max(col_species, col_pos + 2) + 1) return false;
            const int z = atomic_number(t[col_species]);
            if (z == 0) {
                std::cerr << "Element not found: " << t[col_species] << std::endl;
                return false;
            }
            float x[3];
            try {
                for (int k = 0; k < 3; ++k) x[k] = std::stof(t[col_pos + k]);
            } catch (const std::exception &) {
                std::cerr << "Invalid line: " << line << std::endl;
                return false;
            }
            s.Z.push_back((uint16_t)z);
            s.pos.insert(s.pos.end(), x, x + 3);
        }
        out.push_back(std::move(s));
    }
    return !out.empty();
}

bool ends_with(const std::string &s, const char *suf) {
    const size_t n = std::strlen(suf);
    return s.size() >= n && s.compare(s.size() - n, n, suf) == 0;
}

void usage() {
    std::cout << "Usage: d3_cpp [options] FILE...   (.arc or .xyz/.extxyz)" << std::endl;
    std::cout << "Options:" << std::endl;
    std::cout << "  -h, --help        Display this help message" << std::endl;
    std::cout << "  --use-atm         Include the ATM three-body term" << std::endl;
    std::cout << "  --bj              Becke-Johnson damping (default: zero damping)" << std::endl;
    std::cout << "  --fp32            FP32 arithmetic (default)" << std::endl;
    std::cout << "  --fp64            FP64 arithmetic" << std::endl;
    std::cout << "  --cutoff R        two-body cutoff in bohr (default 60)" << std::endl;
    std::cout << "  --cn-cutoff R     CN cutoff in bohr (default 40)" << std::endl;
    std::cout << "  --atm-cutoff R    ATM cutoff in bohr (default 40)" << std::endl;
}

}  // namespace

int main(int argc, char **argv) {
    std::vector<std::string> files;
    bool use_atm = false, bj = false;
    float rc = 60.0f, rccn = 40.0f, rc3 = 40.0f;
    if (argc < 2) {
        std::cerr << "No input file provided. Use -h or --help for usage." << std::endl;
        return EXIT_FAILURE;
    }
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "-h" || a == "--help") {
            usage();
            return 0;
        } else if (a == "--use-atm") {
            use_atm = true;
        } else if (a == "--bj") {
            bj = true;
        } else if (a == "--fp32") {
            setenv("D3_PRECISION", "fp32", 1);
        } else if (a == "--fp64") {
            setenv("D3_PRECISION", "fp64", 1);
        } else if ((a == "--cutoff" || a == "--cn-cutoff" || a == "--atm-cutoff") && i + 1 < argc) {
            const float v = std::strtof(argv[++i], nullptr);
            (a == "--cutoff" ? rc : a == "--cn-cutoff" ? rccn : rc3) = v;
        } else if (std::ifstream(a).good()) {
            files.push_back(a);
        } else {
            std::cerr << "Unknown argument or file does not exist: " << a << std::endl;
            return 1;
        }
    }
    for (const std::string &file : files) {
        std::vector<Structure> structures;
        const bool xyz = ends_with(file, ".xyz") || ends_with(file, ".extxyz");
        if (!(xyz ? read_extxyz(file, structures) : read_arc(file, structures))) {
            std::cerr << "Error parsing file: " << file << std::endl;
            return EXIT_FAILURE;
        }
        std::cout << "Parsed " << structures.size() << " structures from file: " << file << std::endl;
        for (Structure &s : structures) {
            const uint64_t n = s.Z.size();
            if (!s.has_cell) {
                std::cerr << "Structure without a periodic cell in " << file << std::endl;
                return EXIT_FAILURE;
            }
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) std::cout << "Cell[" << i << "][" << j << "]: " << s.cell[i][j] << std::endl;
            std::cout << "Computing dispersion energy for " << n << " atoms..." << std::endl;
            const damping_type_t damp = bj ? BJ_DAMPING : ZERO_DAMPING;
            d3_handle_t *h = use_atm ? init_d3_handle_with_atm(s.Z.data(), n, n, rc, rccn, rc3, damp, PBE)
                                     : init_d3_handle(s.Z.data(), n, n, rc, rccn, damp, PBE);
            if (!h) {
                std::cerr << "Failed to initialize D3 handle." << std::endl;
                return EXIT_FAILURE;
            }
            std::vector<float> force(3 * n), stress(9);
            float energy = 0.0f;
            set_atoms(h, s.pos.data(), s.Z.data(), n);
            set_cell(h, s.cell);
            clear_d3_handle(h);
            const auto t0 = std::chrono::high_resolution_clock::now();
            const uint16_t st = compute_dispersion_energy_from_handle_status(h, &energy, force.data(), stress.data());
            const auto t1 = std::chrono::high_resolution_clock::now();
            free_d3_handle(h);
            if (st != 0) {
                std::cerr << "Dispersion evaluation failed." << std::endl;
                return EXIT_FAILURE;
            }
            std::cout << "Elapsed time: " << std::chrono::duration<double>(t1 - t0).count() << " seconds" << std::endl;
            std::cout << std::fixed << std::setprecision(20) << "Energy: " << energy << " eV" << std::endl;
            float sum[3] = {0.0f, 0.0f, 0.0f};
            for (uint64_t i = 0; i < n; ++i) {
                for (int k = 0; k < 3; ++k) sum[k] += force[3 * i + k];
                std::cout << "Force[" << i << "]: " << force[3 * i] << " " << force[3 * i + 1] << " " << force[3 * i + 2]
                          << std::endl;
            }
            std::cout << "Force sum: " << sum[0] << " " << sum[1] << " " << sum[2] << std::endl;
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) std::cout << "Stress[" << i << "][" << j << "]: " << stress[3 * i + j] << std::endl;
            std::cout.unsetf(std::ios::fixed);
            std::cout << std::setprecision(6);
        }
    }
    return EXIT_SUCCESS;
}
