// The window-pass planner and launcher (window.hip).
#pragma once

#include <cstdint>
#include <memory>
#include <vector>

#include "mh_internal.h"

namespace mh {

// Per-amplitude FP operations a pass may carry before a new one starts:
// about the point where a pass's arithmetic takes as long as its HBM read
// and write of the state.
constexpr uint32_t kDefaultPassCost = 160;
// The cap for batched rounds: measured on MI355X over 4096 ansatz circuits
// (10..22 qubits), 320 gives about 3% more circuits per second than 160 at
// both precisions, while the single-state random circuit at 33 qubits is
// window-limited and gains nothing from it.
constexpr uint32_t kBatchPassCost = 320;

// Window qubits of one pass: the 6 lane bits of a wavefront plus the
// register bits of a lane (4).
uint32_t window_width(int32_t precision) noexcept;

struct WindowPass {
    uint64_t dense_mask = 0;      // qubits its dense blocks act on
    uint32_t kmax = 0;            // widest dense block
    std::vector<uint32_t> blocks; // consecutive plan blocks
    bool generic = false;         // one block on the per-group generic kernel
};

// Groups consecutive blocks into passes over windows of at most w qubits.
std::vector<WindowPass> plan_window_passes(uint32_t n, int32_t precision, const std::vector<Block>& blocks,
                                           uint32_t w, uint32_t max_cost);

void run_window_passes(void* state, uint32_t n, int32_t precision, const std::vector<Block>& blocks, uint32_t w,
                       uint32_t max_cost, bool time_passes, PassStats* stats);

// One state of a batch: its amplitudes and the blocks to apply to them.
struct BatchItem {
    void* state = nullptr;
    uint32_t n = 0;
    const std::vector<Block>* blocks = nullptr;
    uint32_t max_cost = 0;  // per-pass FP work cap; 0: kBatchPassCost
    // initialise: the state starts as phase * |0...0> (its memory need not be
    // initialised): the first pass generates it instead of loading it, or, when
    // that pass is not a window pass, an initialisation kernel runs first.
    bool initialise = false;
    double phase_re = 1.0, phase_im = 0.0;
};

// How a batch's launches are laid out on the device.
//   group_bytes = 0: pass p of every state of an add() call in one launch, the
//   calls alternating between two kernel streams.
//   group_bytes > 0: the states of an add() call are split, largest first,
//   into cache groups of at most group_bytes of state (a state larger than
//   that is a group of its own); each group runs all its passes back to back
//   and the groups rotate over kernel_streams streams (1..4), so that a
//   group's passes find its states in the 256 MB Infinity Cache.
//   temporal: ordinary loads and stores of the states (they stay in the
//   caches) instead of non-temporal ones.
struct BatchSchedule {
    uint64_t group_bytes = 0;
    uint32_t kernel_streams = 2;
    bool temporal = false;
};

struct BatchStats {
    uint32_t launches = 0;
    uint32_t max_passes = 0;
    double compile_seconds = 0;  // host: pass planning and table building
};

// Applies the blocks of groups of items as they are added, without waiting:
// add() plans and compiles a group's passes on host threads, copies their
// tables on the copy stream and launches them behind the work already queued
// (so the host prepares the next group while the device runs this one).
// Groups alternate between the two kernel streams, both of which start behind
// the work queued on the main stream before the runner was made; the groups
// must touch disjoint states. `hooks` (optional) sees the group's compiled
// program and each of its launches. The reduction stream also starts behind
// the work queued before the runner. finish() synchronises every stream, the
// reduction stream included. The destructor waits for what was launched.
struct BatchHooks;  // batch_program.h

class BatchPassRunner {
public:
    BatchPassRunner(int32_t precision, uint32_t threads, const BatchSchedule& schedule = {});
    ~BatchPassRunner();
    BatchPassRunner(const BatchPassRunner&) = delete;
    BatchPassRunner& operator=(const BatchPassRunner&) = delete;
    void add(const BatchItem* items, std::size_t count, BatchStats* stats, const BatchHooks* hooks = nullptr);
    void finish();

private:
    struct Impl;
    std::unique_ptr<Impl> impl_;
};

// Applies every item's blocks to its state (default window and pass cap):
// pass p of all items in one launch. Plans and compiles on `threads` host
// threads; synchronises. Every item needs at least window_width() qubits.
void run_window_passes_batch(int32_t precision, const std::vector<BatchItem>& items, uint32_t threads,
                             BatchStats* stats);

// A block the window passes do not take, on its own pass (gates.hip): the
// one-thread-per-group kernel, or the matrix cores for complex64 blocks of 5
// or 6 qubits; synchronises.
void launch_generic(void* state, uint32_t n, int32_t precision, const cplx* matrix, const uint32_t* operands,
                    uint32_t k);

}  // namespace mh
