// Tuning tables: launch parameters per (kind, field mode, size bucket, batch
// size), written by modlin-rs_amd/autotune.py and read when a context is
// created.  Host code only.
#include <dlfcn.h>
#include <stdio.h>
#include <string.h>

#include <cmath>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "limb.hpp"
#include "tuning.hpp"

namespace mdl {

static std::string mode_of(uint64_t p) {
    if (p == P61) return "m61";
    if (p == P31) return "m31";
    return "g" + std::to_string(limb_count(p));
}

static uint64_t side_bucket(uint64_t n) {
    uint64_t s = 256;
    while (s < n && s < (1ull << 62)) s <<= 1;
    return s;
}

bool parse_tuning(const std::string &path, std::vector<TuneEntry> &out, std::string &err) {
    std::ifstream in(path);
    if (!in) {
        err = "cannot read tuning table " + path;
        return false;
    }
    std::vector<TuneEntry> table;
    std::string line;
    int lineno = 0;
    while (std::getline(in, line)) {
        lineno++;
        const size_t hash = line.find('#');
        if (hash != std::string::npos) line.resize(hash);
        std::istringstream ss(line);
        std::string kind, mode, count;
        uint64_t side = 0;
        if (!(ss >> kind)) continue;  // blank or comment
        TuneEntry e{};
        uint32_t v[9] = {};
        bool ok = (kind == "rank" || kind == "bm") && (ss >> mode >> side >> count) && (count == "1" || count == "many");
        for (int i = 0; ok && i < 7; i++) ok = static_cast<bool>(ss >> v[i]);
        // two optional trailing fields (far_depth far_quarters), both or neither
        std::string extra;
        if (ok && (ss >> extra)) {
            try {
                v[7] = (uint32_t)std::stoul(extra);
            } catch (...) {
                ok = false;
            }
            ok = ok && static_cast<bool>(ss >> v[8]) && !(ss >> extra);
        }
        if (!ok || side == 0 || (side & (side - 1)) || (v[3] & 1) || v[4] > 64 || v[5] > 1024 ||
            !bm_threads_ok(v[6]) || v[7] > 2 || v[8] > 2) {
            err = path + ":" + std::to_string(lineno) + ": malformed tuning entry";
            return false;
        }
        e.bm = kind == "bm";
        e.mode = mode;
        e.side = side;
        e.many = count == "many";
        e.t.far_per_cu = v[0];
        e.t.near_per_cu = v[1];
        e.t.gather_wgs = v[2];
        e.t.check_every = v[3];
        e.t.panel = v[4];
        e.t.bm_coop_wgs = v[5];
        e.t.bm_threads = v[6];
        e.t.far_depth = v[7];
        e.t.far_quarters = v[8];
        table.push_back(e);
    }
    out.swap(table);
    return true;
}

std::string default_tuning_path() {
    Dl_info info;
    if (dladdr((const void *)&default_tuning_path, &info) && info.dli_fname) {
        std::string lib(info.dli_fname);
        const size_t slash = lib.rfind('/');
        return (slash == std::string::npos ? std::string(".") : lib.substr(0, slash)) + "/tuning.txt";
    }
    return "tuning.txt";
}

modlin_tuning_t select_tuning(const std::vector<TuneEntry> &table, const modlin_tuning_t *override_, uint64_t p,
                              uint64_t max_n, uint64_t count, bool bm) {
    modlin_tuning_t t{};
    const std::string mode = mode_of(p);
    const uint64_t side = side_bucket(max_n);
    const bool many = count > 1;
    const TuneEntry *best = nullptr;
    double bestd = 0;
    for (int pass = 0; pass < 2 && !best; pass++) {  // same count first, then any count
        for (const TuneEntry &e : table) {
            if (e.bm != bm || e.mode != mode || (pass == 0 && e.many != many)) continue;
            const double d = std::fabs(std::log2((double)e.side) - std::log2((double)side));
            if (!best || d < bestd) {
                best = &e;
                bestd = d;
            }
        }
    }
    if (best) t = best->t;
    if (override_) {
        const uint32_t *o = &override_->far_per_cu;
        uint32_t *d = &t.far_per_cu;
        for (int i = 0; i < 9; i++)
            if (o[i]) d[i] = o[i];
    }
    return t;
}

}  // namespace mdl

extern "C" int modlin_tuning_select(const char *path, uint64_t p, uint64_t max_n, uint64_t count, int bm,
                                    modlin_tuning_t *out) {
    if (!path || !out || p < 2 || (p >> 63)) return MODLIN_EINVAL;
    std::vector<mdl::TuneEntry> table;
    std::string err;
    if (!mdl::parse_tuning(path, table, err)) return MODLIN_EINVAL;
    *out = mdl::select_tuning(table, nullptr, p, max_n, count, bm != 0);
    return MODLIN_OK;
}
