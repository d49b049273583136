// Tuning tables of the launch parameters (host code; see modlin_ctx_load_tuning).
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/modlin.h"

namespace mdl {

// One line of a tuning table.
struct TuneEntry {
    bool bm;           // a Berlekamp-Massey entry (else rank)
    std::string mode;  // m61, m31 or gL
    uint64_t side;     // power of two
    bool many;         // count > 1
    modlin_tuning_t t;
};

// 0 (default) or a thread count the batched BM kernel is built for.
inline bool bm_threads_ok(uint32_t v) { return v == 0 || v == 256 || v == 512 || v == 1024; }
// Reads a table; false (with err set) if the file is unreadable or a line is malformed.
bool parse_tuning(const std::string &path, std::vector<TuneEntry> &out, std::string &err);
// tuning.txt in the directory of this library.
std::string default_tuning_path();
// The launch parameters of a call: the table entry of its kind, mode and count
// with the nearest size bucket, with the override's nonzero fields on top; zero
// fields mean built-in defaults.
modlin_tuning_t select_tuning(const std::vector<TuneEntry> &table, const modlin_tuning_t *override_, uint64_t p,
                              uint64_t max_n, uint64_t count, bool bm);

}  // namespace mdl
