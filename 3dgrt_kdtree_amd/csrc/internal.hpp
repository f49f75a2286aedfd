// Internal declarations shared by the host core, the device runtime and the
// serving loop.  Not part of the C ABI.
#pragma once

#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "kdgrt.h"

struct kdgrt_tree {
    std::vector<kdgrt_node> nodes;
    std::vector<uint32_t> offsets;
    float aabb[6] = {0, 0, 0, 0, 0, 0}; // xmin,ymin,zmin,xmax,ymax,zmax (.kdt order)
    kdgrt_tree_stats stats{};
};

namespace kdgrt {

// Thread-local last error; set_error returns the status so call sites can
// `return set_error(...)`.
kdgrt_status set_error(kdgrt_status st, const std::string& msg);
void clear_error();

// Tree statistics (depth, leaves, ...) by a walk from the root.  Returns
// false if the walk meets an out-of-range child (the tree is then invalid).
bool compute_tree_stats(const kdgrt_tree& t, kdgrt_tree_stats& st, std::string* err);
kdgrt_status validate_tree(const kdgrt_node* nodes, uint32_t n_nodes, uint64_t n_offsets,
                           const uint32_t* offsets, uint64_t n_gaussians);

// SplitMix64 step (Steele, Lea, Flood 2014): used to derive independent seeds.
inline uint64_t splitmix64(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

inline uint64_t seed_for(uint64_t master, uint64_t index) {
    uint64_t s = master ^ (0xD1B54A32D192ED03ull * (index + 1));
    return splitmix64(s);
}

double wall_seconds();

} // namespace kdgrt
