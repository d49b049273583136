// Host statistics of modlin (src/stats.rs): p-values of a corank and of a
// linear complexity, and their printing.
#pragma once
#include <stdint.h>

#include <string>

namespace modlin {
namespace stats {

constexpr double P_FLOOR = 1e-307;  // stats.rs:16

double corank_prob(uint64_t p, uint64_t n, uint64_t d);
double corank_tail_pvalue(uint64_t p, uint64_t n, uint64_t c);
double lc_left_tail_pvalue(uint64_t p, uint64_t n, uint64_t ell);
std::string debug_f64(double x);
std::string pretty_p_value(double p);

}  // namespace stats
}  // namespace modlin
