// Index arithmetic of the per-group gate kernel (gates.hip), shared with its
// CPU check (tests/native/gate_index_check.cpp), which replays every lane of
// every launch the host makes on the host with bounds checks.
//
// A k-qubit block splits the 2^n amplitudes into 2^(n-k) groups of 2^k: group
// g's first amplitude is g with a zero bit inserted at each operand position
// (ascending), and its member l sets operand ops[j] where bit j of l is set
// (bit j of a matrix row/column index is operand j, reference
// src/kernels/dense_kernels.h:32-37).
#pragma once

#include <cstdint>

namespace mh {

constexpr uint32_t kGroupThreads = 256;
constexpr uint64_t kGroupMaxBlocks = uint64_t{1} << 20;

__host__ __device__ inline uint64_t group_base(uint64_t g, const uint32_t* sorted, uint32_t k) {
    uint64_t base = g;
    for (uint32_t j = 0; j < k; ++j) {
        const uint64_t low = (uint64_t{1} << sorted[j]) - 1;
        base = ((base & ~low) << 1) | (base & low);
    }
    return base;
}

__host__ __device__ inline uint64_t group_member(uint64_t base, uint32_t l, const uint32_t* ops, uint32_t k) {
    uint64_t idx = base;
    for (uint32_t j = 0; j < k; ++j)
        if ((l >> j) & 1u) idx |= uint64_t{1} << ops[j];
    return idx;
}

// Workgroups of the launch over `groups` groups: one thread per group, at
// most kGroupMaxBlocks workgroups (threads then stride by the grid).
inline uint32_t group_grid(uint64_t groups) {
    uint64_t blocks = (groups + kGroupThreads - 1) / kGroupThreads;
    if (blocks < 1) blocks = 1;
    if (blocks > kGroupMaxBlocks) blocks = kGroupMaxBlocks;
    return static_cast<uint32_t>(blocks);
}

}  // namespace mh
