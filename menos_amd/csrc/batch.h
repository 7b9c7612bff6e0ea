// Batched rounds (batch.hip): many planned circuits executed together.
#pragma once

#include <cstdint>
#include <functional>
#include <map>
#include <utility>
#include <vector>

#include "batch_program.h"
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

// The term reductions of a batch program, in buckets of states whose passes
// end at the same launch. A bucket's two reduction launches go on the
// reduction stream as soon as the launch that finishes its last state is
// queued (after(), the program's AfterLaunch hook), so the reductions of
// shallow circuits run beside the passes of deep ones. Within each launch
// group, states are taken in order of their pass counts into buckets of about
// half, a third and the last sixth of the group's reduction work (one bucket
// for small groups). Every term lands at out[2 * (out_index + t)] of a device
// array; each state's terms take the same kernels, blocks and order as
// mh_expectation, whatever its bucket.
class EarlyReductions {
public:
    // terms[i]: program item i's terms (terms == 0: none).
    EarlyReductions(int32_t precision, const BatchProgram& prog, const std::vector<ExpectationItem>& terms,
                    double* out, hipStream_t reduce_stream);
    ~EarlyReductions();
    EarlyReductions(const EarlyReductions&) = delete;
    EarlyReductions& operator=(const EarlyReductions&) = delete;
    // Copies every bucket's descriptors on `copy`, not waited for; when
    // `ordered`, the first reduction waits for the copy on the device (leave
    // it false when the copy is synchronised before a graph capture).
    void upload(hipStream_t copy, bool ordered);
    void after(const BatchLaunch& l, hipStream_t stream);
    // Queues the buckets no launch finished (states without passes) behind
    // everything queued so far on `streams`.
    void flush(const std::vector<hipStream_t>& streams);
    uint32_t launches() const { return launches_; }

private:
    struct Bucket {
        uint32_t group = 0, pass = 0;
        std::unique_ptr<ExpectationProgram> program;
        DeviceBuffer workspace;
        bool queued = false;
    };
    void queue(Bucket& b);
    hipEvent_t event();
    std::vector<Bucket> buckets_;
    std::vector<hipEvent_t> events_;
    hipEvent_t uploaded_ = nullptr;
    bool wait_upload_ = false;
    double* out_ = nullptr;
    hipStream_t reduce_ = nullptr;
    uint32_t launches_ = 0;
};

// Runs body(i) for i in [0, count) on up to `threads` host threads; rethrows
// the first exception.
void parallel_for(std::size_t count, uint32_t threads, const std::function<void(std::size_t)>& body);

}  // namespace mh
