// table_cache.hpp -- the on-disk cache of the Laplace operator tables (table_cache.cpp).
#pragma once

#include <cstdint>
#include <vector>

namespace fmmh {

// The directory of the cache for this process (empty or null: no cache), and the loads
// that hit or missed since it was set.
void set_table_cache(const char *directory);
void table_cache_stats(int64_t *hits, int64_t *misses);
// build_tables through the cache: loaded when the directory holds valid files of degree p,
// else built and stored.
void cached_tables(int p, std::vector<double> &m2m, std::vector<double> &l2l, std::vector<double> &m2l);

}  // namespace fmmh
