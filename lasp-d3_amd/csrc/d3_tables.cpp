// Synthetic D3 table generator and table-file loader (layout in d3_tables.hpp).
#include "d3_tables.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <stdexcept>

namespace d3 {
namespace {

// splitmix64 finaliser: a stateless hash, so every table entry depends only on its key.
uint64_t mix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
double u01(uint64_t a, uint64_t b = 0, uint64_t c = 0, uint64_t d = 0) {
    uint64_t h = mix64(a * 0x100000001B3ull ^ mix64(b + 0x51ED27ull) ^ mix64(c * 31 + 7) ^ mix64(d * 131 + 11));
    return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}

enum Key : uint64_t { K_NREF = 1, K_CNREF, K_C6, K_C6PAIR, K_R0, K_R0PAIR, K_RCOV, K_R2R4 };

}  // namespace

Table make_synthetic_table(bool extended) {
    Table t;
    t.nel = extended ? 103 : 95;
    t.nref = extended ? 7 : 5;
    t.zoff = extended ? 1 : 0;
    t.zmax = extended ? 103 : 94;
    t.name = extended ? "synthetic-extended" : "synthetic-original";
    const int nel = t.nel, nref = t.nref;
    t.c6ab.assign((size_t)nel * nel * nref * nref * 3, -1.0f);
    t.r0ab.assign((size_t)nel * nel, 0.0f);
    t.rcov.assign(nel, 0.0f);
    t.r2r4.assign(nel, 0.0f);

    // per-element quantities
    std::vector<int> nr(t.zmax + 1, 0);
    std::vector<std::vector<double>> cnref(t.zmax + 1), c6self(t.zmax + 1);
    std::vector<double> r0el(t.zmax + 1, 0.0);
    for (int Z = 1; Z <= t.zmax; ++Z) {
        const double g = 1.0 - std::exp(-Z / 14.0);  // grows with Z, saturates for heavy elements
        nr[Z] = 1 + (int)(u01(K_NREF, Z) * nref);
        if (nr[Z] > nref) nr[Z] = nref;
        double cn = 0.0;
        for (int a = 0; a < nr[Z]; ++a) {
            if (a > 0) cn = std::min(12.0, cn + 0.6 + 2.6 * u01(K_CNREF, Z, a));
            cnref[Z].push_back((double)(float)cn);
            // C6 of the free atom between ~3 (H) and ~1500 (heavy), decreasing with CN
            double base = std::exp(std::log(3.0) + (std::log(1500.0) - std::log(3.0)) * g * (0.75 + 0.25 * u01(K_C6, Z)));
            c6self[Z].push_back(base * (1.0 - 0.035 * cn) * (0.9 + 0.2 * u01(K_C6, Z, a + 1)));
        }
        t.rcov[t.slot(Z)] = (float)std::clamp(0.8 + 3.6 * g + 0.4 * (u01(K_RCOV, Z) - 0.5), 0.8, 5.0);
        t.r2r4[t.slot(Z)] = (float)std::clamp(1.5 + 7.0 * g + 0.5 * (u01(K_R2R4, Z) - 0.5), 1.5, 9.0);
        r0el[Z] = std::clamp(2.2 + 6.0 * g + 0.6 * (u01(K_R0, Z) - 0.5), 2.0, 9.0);
    }
    for (int A = 1; A <= t.zmax; ++A) {
        for (int B = 1; B <= t.zmax; ++B) {
            const int lo = std::min(A, B), hi = std::max(A, B);
            t.r0ab[(size_t)t.slot(A) * nel + t.slot(B)] =
                (float)(0.5 * (r0el[A] + r0el[B]) * (1.0 + 0.06 * (u01(K_R0PAIR, lo, hi) - 0.5)));
            for (int a = 0; a < nr[A]; ++a) {
                for (int b = 0; b < nr[B]; ++b) {
                    // symmetric key: (A,a) and (B,b) ordered so that c6[A][B][a][b] == c6[B][A][b][a]
                    uint64_t ka = (uint64_t)A * 16 + a, kb = (uint64_t)B * 16 + b;
                    double pert = u01(K_C6PAIR, std::min(ka, kb), std::max(ka, kb));
                    double c6 = std::sqrt(c6self[A][a] * c6self[B][b]) * (0.92 + 0.16 * pert);
                    size_t idx = t.c6_index(t.slot(A), t.slot(B), a, b);
                    t.c6ab[idx + 0] = (float)c6;
                    t.c6ab[idx + 1] = (float)cnref[A][a];
                    t.c6ab[idx + 2] = (float)cnref[B][b];
                }
            }
        }
    }
    return t;
}

Table load_table_file(const std::string &path) {
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) throw std::runtime_error("cannot open table file " + path);
    char magic[8];
    int32_t hdr[4];
    auto rd = [&](void *p, size_t n) {
        if (std::fread(p, 1, n, f) != n) {
            std::fclose(f);
            throw std::runtime_error("truncated table file " + path);
        }
    };
    rd(magic, 8);
    if (std::memcmp(magic, "D3TAB001", 8) != 0) {
        std::fclose(f);
        throw std::runtime_error("bad magic in table file " + path);
    }
    rd(hdr, sizeof(hdr));
    Table t;
    t.nel = hdr[0];
    t.nref = hdr[1];
    t.zoff = hdr[2];
    t.zmax = hdr[3];
    t.name = path;
    if (t.nel <= 0 || t.nel > 256 || t.nref <= 0 || t.nref > 8 || t.zoff < 0 || t.zoff > 1 || t.zmax < 1 ||
        t.zmax - t.zoff >= t.nel) {
        std::fclose(f);
        throw std::runtime_error("bad table header in " + path);
    }
    t.c6ab.resize((size_t)t.nel * t.nel * t.nref * t.nref * 3);
    t.r0ab.resize((size_t)t.nel * t.nel);
    t.rcov.resize(t.nel);
    t.r2r4.resize(t.nel);
    rd(t.c6ab.data(), t.c6ab.size() * 4);
    rd(t.r0ab.data(), t.r0ab.size() * 4);
    rd(t.rcov.data(), t.rcov.size() * 4);
    rd(t.r2r4.data(), t.r2r4.size() * 4);
    std::fclose(f);
    return t;
}

namespace {
std::mutex g_table_mu;
Table *g_table = nullptr;
}  // namespace

const Table &active_table() {
    std::lock_guard<std::mutex> lk(g_table_mu);
    if (!g_table) g_table = new Table(make_synthetic_table(false));
    return *g_table;
}

void set_active_table(Table t) {
    std::lock_guard<std::mutex> lk(g_table_mu);
    Table *old = g_table;
    g_table = new Table(std::move(t));
    // Handles copy what they need at init, so the previous table can go (kept alive is simpler
    // and costs a few MB; freed here because handles never point into it).
    delete old;
}

}  // namespace d3
