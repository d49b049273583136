// modlin: command-line drop-in for the reference's binary (src/main.rs,
// src/cli.rs).  Same flags, validation and stdout lines; the rank and
// linear-complexity work of all reps goes through one batched GPU call per
// group of reps (modlin_rank_batch / modlin_lc_batch) instead of one CPU call
// per rep.
//
//   modlin -R SIDE [REPS] -p P [-S SEED]      rank test            (main.rs:66-127)
//   modlin -L LENGTH [REPS] -p P [-S SEED]    linear-complexity    (main.rs:129-176)
//
// The reference selects its generator when it is compiled (Cargo features,
// prng.rs:1-45); here -G/--generator selects it at run time (default: mixmax,
// the N = 240 generator) and --lux adds the keep-5/drop-9 decimation
// (prng.rs:246-279).  Matrix and sequence entries are generated on the GPU
// (modlin_prng_fill_device, exact jump-ahead to each group's first output) and
// never cross PCIe; --host-gen generates them on the host instead.  Progress
// goes to stderr; stdout carries exactly the reference's lines.
#include <ctype.h>
#include <inttypes.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/modlin.h"
#include "prng.hpp"

namespace {

const char *kUsage =
    "Usage: modlin [OPTIONS] --modulus <MODULUS> <--rank <SIDE>|--linear-complexity <LENGTH>> [REPS]";

[[noreturn]] void die(const std::string &msg) {  // cli.rs:183-189: clap-style error, status 2
    fprintf(stderr, "error: %s\n\n%s\n\nFor more information, try '--help'.\n", msg.c_str(), kUsage);
    exit(2);
}

void help() {
    printf(
        "Runs a modular rank or linear-complexity test on the GPU.\n\n%s\n\n"
        "Arguments:\n  [REPS]\n          Number of matrices or sequences to test. [default: 1]\n\n"
        "Options:\n"
        "  -R, --rank <SIDE>\n          Run the modular rank test on SIDE×SIDE matrices; mutually exclusive with -L.\n"
        "  -L, --linear-complexity <LENGTH>\n"
        "          Run the modular linear-complexity test on sequences of the given length; mutually exclusive "
        "with -R.\n"
        "  -p, --modulus <MODULUS>\n          Prime field modulus p: a prime below 2⁶³.\n"
        "  -S, --seed <SEED>\n          PRNG seed (decimal, or 0x/0o/0b prefix; underscores allowed). [default: 0]\n"
        "      --log-interval <LOG_INTERVAL>\n          How often to log progress (accepted for compatibility). "
        "[default: 10s]\n"
        "  -G, --generator <NAME>\n          xoroshiro128pp, mixmax (= mixmax240), mixmax17 or mixmax256. "
        "[default: mixmax]\n"
        "      --lux\n          MIXMAX decimation: keep 5 outputs, drop 9.\n"
        "      --method <auto|plain|mfma>\n          Trailing-update method of the rank test. [default: auto]\n"
        "      --batch-gib <GIB>\n          Memory per batched call. [default: 8]\n"
        "      --host-gen\n          Generate the entries on the host and copy them to the GPU.\n"
        "  -h, --help\n          Print help\n",
        kUsage);
}

// cli.rs:54-72: decimal, or 0x/0o/0b (any case), underscores allowed.
bool parse_u64(const char *s, uint64_t &out) {
    std::string t(s);
    while (!t.empty() && isspace((unsigned char)t.front())) t.erase(t.begin());
    while (!t.empty() && isspace((unsigned char)t.back())) t.pop_back();
    unsigned radix = 10;
    if (t.size() >= 2 && t[0] == '0' && strchr("xXoObB", t[1])) {
        radix = (t[1] == 'x' || t[1] == 'X') ? 16 : (t[1] == 'o' || t[1] == 'O') ? 8 : 2;
        t = t.substr(2);
    }
    std::string d;
    for (char c : t)
        if (c != '_') d.push_back(c);
    if (d.empty()) return false;
    unsigned __int128 v = 0;
    for (char c : d) {
        unsigned dig;
        if (c >= '0' && c <= '9') dig = c - '0';
        else if (c >= 'a' && c <= 'f') dig = c - 'a' + 10;
        else if (c >= 'A' && c <= 'F') dig = c - 'A' + 10;
        else return false;
        if (dig >= radix) return false;
        v = v * radix + dig;
        if (v > UINT64_MAX) return false;
    }
    out = (uint64_t)v;
    return true;
}

// Deterministic Miller-Rabin for 64-bit n with the first twelve prime bases
// (exact below 3.3e24), in place of the reference's primal::is_prime (cli.rs:178).
uint64_t mulmod(uint64_t a, uint64_t b, uint64_t m) { return (uint64_t)((unsigned __int128)a * b % m); }
uint64_t powmod(uint64_t a, uint64_t e, uint64_t m) {
    uint64_t r = 1 % m;
    a %= m;
    while (e) {
        if (e & 1) r = mulmod(r, a, m);
        a = mulmod(a, a, m);
        e >>= 1;
    }
    return r;
}
bool is_prime(uint64_t n) {
    if (n < 2) return false;
    static const uint64_t small[] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
    for (uint64_t q : small)
        if (n % q == 0) return n == q;
    uint64_t d = n - 1;
    int s = 0;
    while ((d & 1) == 0) {
        d >>= 1;
        s++;
    }
    for (uint64_t a : small) {
        uint64_t x = powmod(a, d, n);
        if (x == 1 || x == n - 1) continue;
        bool comp = true;
        for (int i = 1; i < s && comp; i++) {
            x = mulmod(x, x, n);
            if (x == n - 1) comp = false;
        }
        if (comp) return false;
    }
    return true;
}

// main.rs:49-60
std::string format_memory(double bytes) {
    const double GIB = (double)(1ull << 30), MIB = (double)(1ull << 20);
    char buf[64];
    if (bytes >= GIB) snprintf(buf, sizeof buf, "%.3f GiB", bytes / GIB);
    else snprintf(buf, sizeof buf, "%.3f MiB", bytes / MIB);
    return buf;
}

// One generator of prng.rs behind a common interface, with the optional
// decimation of the mixmax-lux feature (keep 5 outputs, then drop 9).
struct Source {
    virtual ~Source() {}
    virtual uint64_t raw() = 0;
    bool lux = false;
    uint32_t kept = 0;
    uint64_t next() {
        if (lux) {
            if (kept == 5) {
                for (int i = 0; i < 9; i++) (void)raw();
                kept = 0;
            }
            kept++;
        }
        return raw();
    }
};
template <class G>
struct SourceOf : Source {
    G g;
    explicit SourceOf(const G &g0) : g(g0) {}
    uint64_t raw() override { return g.next_u64(); }
};

struct Picked {
    Source *src;
    std::string name;
    uint32_t kind;
};

Picked make_source(const std::string &gen, bool lux, uint64_t seed) {
    using namespace modlin::prng;
    Picked pk{nullptr, "", MODLIN_PRNG_XOROSHIRO128PP};
    const char *err = "the zero seed is not allowed for MIXMAX";  // prng.rs:153-155
    if (gen == "xoroshiro128pp") {
        pk.src = new SourceOf<Xoroshiro128pp>(Xoroshiro128pp(seed));
        pk.name = modlin_prng_name(MODLIN_PRNG_XOROSHIRO128PP);
        if (lux) die("--lux applies to the MIXMAX generators only");
        return pk;
    }
    uint32_t kind;
    if (gen == "mixmax" || gen == "mixmax240") kind = MODLIN_PRNG_MIXMAX240;
    else if (gen == "mixmax17") kind = MODLIN_PRNG_MIXMAX17;
    else if (gen == "mixmax256") kind = MODLIN_PRNG_MIXMAX256;
    else die("unknown generator '" + gen + "' (xoroshiro128pp, mixmax, mixmax17, mixmax256)");
    pk.name = modlin_prng_name(kind);
    pk.kind = kind;
    if (lux) pk.name += ", lux: keep 5 drop 9";
    bool ok = false;
    if (kind == MODLIN_PRNG_MIXMAX240) {
        auto *s = new SourceOf<MixMax240>(MixMax240());
        ok = s->g.seed(seed);
        pk.src = s;
    } else if (kind == MODLIN_PRNG_MIXMAX17) {
        auto *s = new SourceOf<MixMax17>(MixMax17());
        ok = s->g.seed(seed);
        pk.src = s;
    } else {
        auto *s = new SourceOf<MixMax256>(MixMax256());
        ok = s->g.seed(seed);
        pk.src = s;
    }
    if (!ok) {
        delete pk.src;
        pk.src = nullptr;
        pk.name += std::string("\n") + err;  // reported after the header lines, as main.rs:29-34 does
    }
    if (pk.src) pk.src->lux = lux;
    return pk;
}

}  // namespace

int main(int argc, char **argv) {
    uint64_t side = 0, length = 0, reps = 1, p = 0, seed = 0;
    bool have_r = false, have_l = false, have_p = false, have_reps = false, lux = false, host_gen = false;
    std::string gen = "mixmax", method = "auto";
    double batch_gib = 8.0;
    auto need = [&](int &i, const char *flag) -> const char * {
        if (i + 1 >= argc) die(std::string("a value is required for '") + flag + "' but none was supplied");
        return argv[++i];
    };
    for (int i = 1; i < argc; i++) {
        std::string a = argv[i];
        std::string val;
        bool has_eq = false;
        if (a.rfind("--", 0) == 0 && a.find('=') != std::string::npos) {
            val = a.substr(a.find('=') + 1);
            a = a.substr(0, a.find('='));
            has_eq = true;
        }
        auto value = [&](const char *flag) -> std::string { return has_eq ? val : std::string(need(i, flag)); };
        auto u64 = [&](const char *flag) -> uint64_t {
            std::string v = value(flag);
            uint64_t x;
            if (!parse_u64(v.c_str(), x)) die("invalid value '" + v + "' for '" + flag + "'");
            return x;
        };
        if (a == "-h" || a == "--help") {
            help();
            return 0;
        } else if (a == "-R" || a == "--rank") {
            side = u64("--rank <SIDE>");
            have_r = true;
        } else if (a == "-L" || a == "--linear-complexity") {
            length = u64("--linear-complexity <LENGTH>");
            have_l = true;
        } else if (a == "-p" || a == "--modulus") {
            p = u64("--modulus <MODULUS>");
            have_p = true;
        } else if (a == "-S" || a == "--seed") {
            seed = u64("--seed <SEED>");
        } else if (a == "--log-interval") {
            (void)value("--log-interval <LOG_INTERVAL>");
        } else if (a == "-G" || a == "--generator") {
            gen = value("--generator <NAME>");
        } else if (a == "--lux") {
            lux = true;
        } else if (a == "--method") {
            method = value("--method <METHOD>");
        } else if (a == "--host-gen") {
            host_gen = true;
        } else if (a == "--batch-gib") {
            batch_gib = atof(value("--batch-gib <GIB>").c_str());
        } else if (!a.empty() && a[0] == '-' && a != "-") {
            die("unexpected argument '" + a + "' found");
        } else {
            if (have_reps) die("unexpected argument '" + a + "' found");
            if (!parse_u64(a.c_str(), reps)) die("invalid value '" + a + "' for '[REPS]'");
            have_reps = true;
        }
    }
    if (have_r && have_l) die("the argument '--rank <SIDE>' cannot be used with '--linear-complexity <LENGTH>'");
    if (!have_r && !have_l) die("the following required arguments were not provided:\n  <--rank <SIDE>|--linear-complexity <LENGTH>>");
    if (!have_p) die("the following required arguments were not provided:\n  --modulus <MODULUS>");
    // cli.rs:159-181
    const uint64_t size = have_r ? side : length;
    if (size < 2) die("SIDE/LENGTH must be at least 2");
    if (have_r && size > UINT32_MAX) die("SIDE is too large: SIDE² overflows a 64-bit index");
    if (reps < 1) die("reps must be at least 1");
    if (p < 2) die("modulus must be at least 2");
    if (p >= (1ull << 63)) die("modulus must be below 2⁶³");
    if (!is_prime(p)) die("modulus must be prime");
    uint32_t m = MODLIN_METHOD_AUTO;
    if (method == "plain") m = MODLIN_METHOD_PLAIN;
    else if (method == "mfma") m = MODLIN_METHOD_MFMA;
    else if (method != "auto") die("invalid value '" + method + "' for '--method <METHOD>'");

    // main.rs:22-34
    Picked pk = make_source(gen, lux, seed);
    const std::string name = pk.name.substr(0, pk.name.find('\n'));
    printf("Generator: %s\n", name.c_str());
    printf("Seed: 0x%016" PRIx64 "\n", seed);
    fflush(stdout);
    if (!pk.src) {
        fprintf(stderr, "Error initializing the generator: %s\n", pk.name.substr(pk.name.find('\n') + 1).c_str());
        return 1;
    }
    modlin_field_t field;
    if (modlin_field_init(p, &field) != MODLIN_OK) die("modulus must be below 2⁶³");

    if (have_r) {  // main.rs:72-127
        const uint64_t n = side;
        printf("Running a modular rank test using %s of RAM: %" PRIu64 " %" PRIu64 "×%" PRIu64 " %s over the field of size %" PRIu64 "\n",
               format_memory((double)n * n * 8).c_str(), reps, n, n, reps == 1 ? "matrix" : "matrices", p);
    } else {  // main.rs:138-148
        const uint64_t n = length;
        printf("Running a modular linear-complexity test using %s of RAM: %" PRIu64 " %s of length %" PRIu64
               " over the field of size %" PRIu64 "\n",
               format_memory(4.0 * n * 8).c_str(), reps, reps == 1 ? "sequence" : "sequences", n, p);
    }
    fflush(stdout);

    modlin_ctx_t *ctx = nullptr;
    int rc = modlin_ctx_create(0, &ctx);
    if (rc != MODLIN_OK) {
        fprintf(stderr, "Error: no usable gfx950 GPU (status %d)\n", rc);
        return 1;
    }
    const uint64_t n = have_r ? side : length;
    const uint64_t item = have_r ? n * n : n;
    uint64_t group = (uint64_t)(batch_gib * (double)(1ull << 30) / (8.0 * (double)item));
    if (group < 1) group = 1;
    if (group > reps) group = reps;
    std::vector<uint64_t> data;
    std::vector<uint64_t> res(group);
    void *dev = nullptr;  // device-generated entries of one group
    if (!host_gen && (rc = modlin_dev_alloc(ctx, group * item * 8, &dev)) != MODLIN_OK) {
        fprintf(stderr, "Error: %s (status %d)\n", modlin_ctx_last_error(ctx), rc);
        modlin_ctx_destroy(ctx);
        return 1;
    }
    for (uint64_t first = 0; first < reps; first += group) {
        const uint64_t cnt = std::min<uint64_t>(group, reps - first);
        uint64_t *base;
        uint32_t mem;
        if (host_gen) {
            data.resize(cnt * item);
            for (uint64_t i = 0; i < cnt * item; i++) data[i] = pk.src->next() % p;  // field.reduce (main.rs:96, 154)
            base = data.data();
            mem = MODLIN_MEM_HOST;
        } else {
            // the same successive outputs, from output first * item on (main.rs:94-105, 153-155)
            rc = modlin_prng_fill_device(ctx, pk.kind, seed, first * item, lux ? 5 : 0, lux ? 9 : 0, &field,
                                         MODLIN_ELEM_U64, dev, cnt * item);
            if (rc != MODLIN_OK) {
                fprintf(stderr, "Error: %s (status %d)\n", modlin_ctx_last_error(ctx), rc);
                modlin_ctx_destroy(ctx);
                return 1;
            }
            base = (uint64_t *)dev;
            mem = MODLIN_MEM_DEVICE;
        }
        fprintf(stderr, "%s %" PRIu64 "-%" PRIu64 "/%" PRIu64 " on the GPU...\n", have_r ? "Matrices" : "Sequences",
                first + 1, first + cnt, reps);
        if (have_r) {
            std::vector<modlin_mat_desc_t> d(cnt);
            for (uint64_t i = 0; i < cnt; i++) d[i] = {base + i * item, n, MODLIN_ELEM_U64, mem};
            modlin_opts_t o{m, 0, 0, 0};
            rc = modlin_rank_batch(ctx, &field, d.data(), cnt, &o, res.data());
        } else {
            std::vector<modlin_seq_desc_t> d(cnt);
            for (uint64_t i = 0; i < cnt; i++) d[i] = {base + i * item, n, MODLIN_ELEM_U64, mem};
            rc = modlin_lc_batch(ctx, &field, d.data(), cnt, nullptr, res.data());
        }
        if (rc != MODLIN_OK) {
            fprintf(stderr, "Error: %s (status %d)\n", modlin_ctx_last_error(ctx), rc);
            modlin_ctx_destroy(ctx);
            return 1;
        }
        char pbuf[96];
        for (uint64_t i = 0; i < cnt; i++) {
            if (have_r) {  // main.rs:119-125
                const uint64_t corank = n - res[i];
                modlin_pretty_p_value(modlin_corank_tail_pvalue(p, n, corank), pbuf, sizeof pbuf);
                printf("Matrix %" PRIu64 "/%" PRIu64 "\tcorank=%" PRIu64 "\tp=%s\n", first + i + 1, reps, corank, pbuf);
            } else {  // main.rs:169-174
                modlin_pretty_p_value(modlin_lc_left_tail_pvalue(p, n, res[i]), pbuf, sizeof pbuf);
                printf("Sequence %" PRIu64 "/%" PRIu64 "\tlinear complexity=%" PRIu64 "\tp=%s\n", first + i + 1, reps,
                       res[i], pbuf);
            }
        }
        fflush(stdout);
    }
    if (dev) modlin_dev_free(ctx, dev);
    modlin_ctx_destroy(ctx);
    delete pk.src;
    return 0;
}
