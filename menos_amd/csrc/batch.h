// Batched rounds (batch.hip): many planned circuits executed together.
#pragma once

#include <cstdint>
#include <functional>
#include <map>
#include <utility>
#include <vector>

#include "mh_internal.h"
#include "window.h"

namespace mh {

// One planned circuit of a round.
struct RoundItem {
    uint32_t n = 0;
    const std::vector<Block>* blocks = nullptr;
    uint32_t max_pass_cost = 0;  // 0: the default
    double global_phase = 0.0;
    // Expectation terms: results land at terms_out[2 * (term_out + t)].
    uint32_t num_terms = 0;
    const uint64_t* x = nullptr;
    const uint64_t* z = nullptr;
    uint32_t term_out = 0;
    // Sampling (shots > 0): counts keyed by the clbit string.
    uint64_t shots = 0;
    uint64_t seed = 0;
    const std::vector<std::pair<uint32_t, uint32_t>>* targets = nullptr;
    std::map<uint64_t, uint64_t>* counts = nullptr;
    // Host destination of the final amplitudes, or null.
    void* statevector = nullptr;
};

// Wall time of a round's phases (each ends with a synchronisation).
struct RoundStats {
    uint32_t launches = 0;
    double prepare_seconds = 0;      // host: the caller's prepare callbacks
    double compile_seconds = 0;      // host: pass planning and tables
    double pass_seconds = 0;         // initialisation and gate passes, compile excluded
    double expectation_seconds = 0;
    double sample_seconds = 0;
};

// Device bytes one state of a round takes in the round's arena (a multiple of 256).
uint64_t round_state_bytes(uint32_t n, int32_t precision);

// Runs one round: every item's state in one pooled arena (no initialisation
// pass: each state's first window pass generates phase * |0...0>), pass p of
// every item in one launch, all expectation terms in two launches, then each
// sampling item's draw and statevector copy.
// Synchronises. Every item needs window_width(precision) to 30 qubits.
// `arena` (device memory of at least the items' round_state_bytes in total)
// holds the states; null takes an arena from the pool for the call.
// The gate passes go to the device in groups of items: prepare(lo, hi), when
// given, runs on the host just before items [lo, hi) are compiled (to fill in
// their blocks), while the device runs the groups before them.
void execute_round(int32_t precision, std::vector<RoundItem>& items, uint32_t threads, double* terms_out,
                   RoundStats* stats, void* arena = nullptr,
                   const std::function<void(std::size_t, std::size_t)>& prepare = nullptr,
                   const BatchSchedule& schedule = {});

// Runs body(i) for i in [0, count) on up to `threads` host threads; rethrows
// the first exception.
void parallel_for(std::size_t count, uint32_t threads, const std::function<void(std::size_t)>& body);

}  // namespace mh
