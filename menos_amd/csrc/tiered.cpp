// The host-offloaded tier (mh_tier_estimate, mh_tier_execute): the dense
// algorithm over a state held in pinned host memory, run on the device through
// windows. It is the reference's Tier::HostOffloaded (include/menos/plan.h:
// 88-98, measured there and not built) on the plan of its disk tier: a state of
// n qubits is 2^(n-L) chunks of 2^L amplitudes, qubits below L local and the
// rest global; a pass holds a set of at most w = W - L global qubits, and the
// greedy rule of src/engines/dense_tiered/pass_plan.cpp:46-122 groups the
// plan's blocks into passes (the chunk with the fewest passes, pass_plan.cpp:
// 124-151). For each pass and each group of 2^w chunks its set selects, the
// chunks are copied into a device window of 2^W amplitudes, the pass's blocks
// run on it with their operands remapped (a global qubit becomes window bit
// L + its rank in the set, chunked_executor.cpp:106-115) through the same
// window kernels as the in-memory engine, and the chunks are copied back
// (chunked_executor.cpp:117-226). Each block multiplies the same amplitudes by
// the same matrix in the same order as in memory, so the state is bitwise the
// in-memory engine's.
//
// On the device side it differs from the disk tier where the hardware does:
//  * two windows are in flight: while one group runs, the other's chunks
//    travel, host-to-device and device-to-host on streams of their own (the
//    link is full duplex); the first pass makes its windows on the device
//    (the initial state) instead of reading a written one;
//  * a pass's set is padded with the lowest free global qubits up to w, so
//    every group fills the window (fewer, longer copies; adjacent chunks are
//    copied as one);
//  * the outputs read the final state once per kind of term instead of once
//    per term: counts and the terms whose X mask lies inside a window share
//    one read in index order (the sampler carries its exact running sum from
//    window to window, bit for bit the in-memory sampler's); a term whose X
//    mask reaches above the window is summed over pairs of half-windows
//    (m, m ^ x), all terms with the same high X bits in one read.
#include <algorithm>
#include <bit>
#include <chrono>
#include <cstring>
#include <map>
#include <string>

#include "engine.hpp"
#include "gate_table.h"
#include "window.h"

namespace mh {
namespace {

// Chunks of at least 2^24 amplitudes are tried first (reference pass_plan.h:76-80).
constexpr uint32_t kPreferredMinChunkQubits = 24;
// A chosen chunk leaves at most 2^6 chunks to a window (a 6-qubit block
// still fits): a group's copies stay at least 1/64 of the window, where a DMA
// copy runs at the link's rate. Below it the reference would take chunks down
// to one amplitude in windows under 2^24.
constexpr uint32_t kMaxChosenWindowGlobals = 6;
// The window with no device budget: 2^28 amplitudes (4 GiB at complex128).
constexpr uint32_t kDefaultWindowQubits = 28;
constexpr uint32_t kWindows = 2;
constexpr uint64_t kMaxTrackedChunks = uint64_t{1} << 16;
constexpr uint32_t kMaxQubits = 40;

uint64_t state_bytes_of(uint32_t n, int32_t precision) { return (uint64_t{1} << n) * amp_bytes(precision); }

// The bits of `value`, lowest first, at the set bits of `mask`, lowest first.
uint64_t deposit(uint64_t value, uint64_t mask) {
    uint64_t out = 0;
    for (uint64_t rest = mask; rest != 0; rest &= rest - 1, value >>= 1)
        if (value & 1u) out |= rest & (~rest + 1);
    return out;
}

uint64_t global_mask_of(const std::vector<uint32_t>& operands, uint32_t chunk_qubits) {
    uint64_t mask = 0;
    for (uint32_t q : operands)
        if (q >= chunk_qubits) mask |= uint64_t{1} << (q - chunk_qubits);
    return mask;
}

struct TierPass {
    uint64_t mask = 0;
    uint32_t first = 0, end = 0;
};

struct TierPlan {
    uint32_t n = 0, chunk = 0, window = 0;
    std::vector<TierPass> passes;
    uint64_t global_blocks = 0;
    uint32_t max_global = 0;
    bool too_small = false;
};

// The greedy rule: a pass opens with a block's global operands and takes each
// following block while the union fits w qubits (reference
// pass_plan.cpp:46-122, over fusion blocks).
TierPlan build_tier_plan(const std::vector<Block>& blocks, uint32_t n, uint32_t chunk, uint32_t window) {
    TierPlan p;
    p.n = n;
    p.chunk = chunk;
    p.window = window;
    const int wg = static_cast<int>(window - chunk);
    bool open = false;
    uint64_t group = 0;
    uint32_t first = 0;
    for (uint32_t i = 0; i < blocks.size(); ++i) {
        if (blocks[i].operands.empty()) continue;
        const uint64_t mask = global_mask_of(blocks[i].operands, chunk);
        const int g = std::popcount(mask);
        p.max_global = std::max<uint32_t>(p.max_global, static_cast<uint32_t>(g));
        if (g > wg) p.too_small = true;
        if (g > 0) ++p.global_blocks;
        if (!open) {
            open = true;
            group = mask;
            continue;
        }
        const uint64_t merged = group | mask;
        if (std::popcount(merged) <= wg) {
            group = merged;
        } else {
            p.passes.push_back(TierPass{group, first, i});
            first = i;
            group = mask;
        }
    }
    if (open) p.passes.push_back(TierPass{group, first, static_cast<uint32_t>(blocks.size())});
    return p;
}

// The chunk with the fewest passes in a window of `window` qubits, the larger
// on a tie; chunks under the preferred minimum only when nothing above fits
// (reference pass_plan.cpp:124-151), and none below window - 6. too_small
// when no chunk fits.
TierPlan choose_tier_plan(const std::vector<Block>& blocks, uint32_t n, uint32_t window) {
    const uint32_t lowest = window > kPreferredMinChunkQubits ? kPreferredMinChunkQubits : 0;
    const uint32_t floor = window > kMaxChosenWindowGlobals ? window - kMaxChosenWindowGlobals : 0;
    TierPlan best;
    bool have = false;
    for (uint32_t chunk = window; chunk >= floor; --chunk) {
        if (chunk < lowest && have) break;
        TierPlan plan = build_tier_plan(blocks, n, chunk, window);
        if (!plan.too_small && (!have || plan.passes.size() < best.passes.size())) {
            best = std::move(plan);
            have = true;
        }
        if (chunk == 0 || chunk == floor) break;
    }
    if (!have) {
        best = build_tier_plan(blocks, n, floor, window);
        best.too_small = true;
    }
    return best;
}

struct TierRequest {
    uint32_t chunk_qubits = 0, window_qubits = 0;
};

// The device bytes beside the two windows: the sampler's tables for two
// windows and its shots, the expectation partials, the pass programs.
uint64_t tier_scratch_bytes(uint32_t window, uint64_t shots, std::size_t terms) {
    uint64_t bytes = 64ull << 10;
    if (shots) {
        const uint64_t blocks = window > 13 ? (uint64_t{1} << (window - 13)) : 1;
        bytes += 2 * (blocks * 112 + 4096) + shots * 16 + 4096;
    }
    if (terms) bytes += 2048ull * 16 * 16 + 4096;
    return bytes;
}

uint64_t tier_device_bytes(uint32_t window, int32_t precision, uint64_t shots, std::size_t terms) {
    return kWindows * state_bytes_of(window, precision) + tier_scratch_bytes(window, shots, terms);
}

// The terms, split the way the outputs read the state: X inside the window
// (one read in index order, shared with the sampler), or reaching above it
// (one read of half-window pairs per value of x >> (W - 1)).
struct TermSets {
    std::vector<uint32_t> local;
    std::map<uint64_t, std::vector<uint32_t>> paired;  // x >> (W - 1) -> terms
};

TermSets split_terms(const std::vector<std::pair<uint64_t, uint64_t>>& terms, uint32_t window) {
    TermSets s;
    for (uint32_t t = 0; t < terms.size(); ++t) {
        const uint64_t x = terms[t].first;
        if (window >= 64 || (x >> window) == 0)
            s.local.push_back(t);
        else
            s.paired[x >> (window - 1)].push_back(t);
    }
    return s;
}

// The window bit of qubit q in a pass whose (padded) set is `set`.
uint32_t window_bit(uint32_t q, uint32_t chunk, uint64_t set) {
    if (q < chunk) return q;
    const uint64_t below = (uint64_t{1} << (q - chunk)) - 1;
    return chunk + static_cast<uint32_t>(std::popcount(set & below));
}

// A pass's set padded with the lowest free global qubits up to w of them.
uint64_t padded_set(uint64_t set, uint32_t globals, uint32_t w) {
    for (uint32_t b = 0; b < globals && static_cast<uint32_t>(std::popcount(set)) < w; ++b) set |= uint64_t{1} << b;
    return set;
}

// A pass's blocks with their operands remapped into its windows.
std::vector<Block> pass_blocks(const std::vector<Block>& all_blocks, const TierPlan& tier, const TierPass& pass) {
    const uint32_t globals = tier.n - tier.chunk;
    const uint64_t set = padded_set(pass.mask, globals, tier.window - tier.chunk);
    std::vector<Block> blocks(all_blocks.begin() + pass.first, all_blocks.begin() + pass.end);
    for (Block& b : blocks)
        for (uint32_t& q : b.operands) q = window_bit(q, tier.chunk, set);
    return blocks;
}

struct Layout {
    Plan plan;
    TierPlan tier;
    uint64_t state_bytes = 0, device_bytes = 0;
    uint64_t link_bytes = 0, link_write_bytes = 0;
    uint64_t output_reads = 0;
    double seconds = 0;
};

Layout lay_out(const ExecutionRequest& r, const TierRequest& t) {
    amp_bytes(r.precision);
    if (r.circuit == nullptr) fail(MH_ERR_INVALID, "null circuit");
    const uint32_t n = r.circuit->num_qubits;
    if (n > kMaxQubits) fail(MH_ERR_INVALID, "num_qubits above 40");
    const uint32_t wmin = window_default(r.precision);
    if (n < wmin) fail(MH_ERR_INVALID, "the host-offloaded tier needs at least " + std::to_string(wmin) + " qubits");
    if (r.want_region_sampler) fail(MH_ERR_INVALID, "the host-offloaded tier keeps no state on the device: no region sampler");
    if (r.bitstring_width > 64) fail(MH_ERR_INVALID, "bitstring_width above 64");
    for (const auto& [q, c] : r.measure_targets)
        if (q >= n || c >= std::max<uint32_t>(r.bitstring_width, 1)) fail(MH_ERR_INVALID, "measure target out of range");
    for (const auto& [x, z] : r.terms)
        if (((x | z) >> n) != 0) fail(MH_ERR_INVALID, "a term's masks reach above the state's qubits");
    Layout L;
    L.plan = make_plan(r);
    L.state_bytes = state_bytes_of(n, r.precision);
    uint32_t window = t.window_qubits;
    if (window == 0) {
        if (r.budget_bytes == 0) {
            window = std::min(n, kDefaultWindowQubits);
        } else {
            for (uint32_t w = wmin; w <= n; ++w)
                if (tier_device_bytes(w, r.precision, r.shots, r.terms.size()) <= r.budget_bytes) window = w;
            if (window == 0)
                fail(MH_ERR_RESOURCE, "the host-offloaded tier's smallest windows need " +
                                          std::to_string(tier_device_bytes(wmin, r.precision, r.shots, r.terms.size())) +
                                          " device bytes, budget " + std::to_string(r.budget_bytes));
        }
    }
    if (window < wmin || window > n)
        fail(MH_ERR_INVALID, "window_qubits must lie in [" + std::to_string(wmin) + ", " + std::to_string(n) + "]");
    if (t.chunk_qubits > window) fail(MH_ERR_INVALID, "chunk_qubits above window_qubits");
    L.tier = t.chunk_qubits == 0 ? choose_tier_plan(L.plan.blocks, n, window)
                                 : build_tier_plan(L.plan.blocks, n, t.chunk_qubits, window);
    if (L.tier.too_small)
        fail(MH_ERR_RESOURCE, "a block has " + std::to_string(L.tier.max_global) +
                                  " operands above the chunk, more than the window of 2^" + std::to_string(window) +
                                  " amplitudes brings together");
    L.device_bytes = tier_device_bytes(window, r.precision, r.shots, r.terms.size());
    if (r.budget_bytes != 0 && L.device_bytes > r.budget_bytes)
        fail(MH_ERR_RESOURCE, "the windows need " + std::to_string(L.device_bytes) + " device bytes, budget " +
                                  std::to_string(r.budget_bytes));
    // The link: every pass writes the state back; every pass after the first
    // reads it; then the outputs' reads.
    const TermSets sets = split_terms(r.terms, window);
    const uint64_t passes = std::max<uint64_t>(L.tier.passes.size(), 1);
    L.output_reads = ((r.shots > 0 || !sets.local.empty()) ? 1 : 0) + sets.paired.size();
    L.link_write_bytes = L.state_bytes * passes;
    L.link_bytes = L.link_write_bytes + L.state_bytes * (passes - 1 + L.output_reads);
    // The time model: a pass takes the longer of its copies (the two
    // directions overlap) and its window passes over every group.
    const mh_calibration& c = calibration();
    const double sb = static_cast<double>(L.state_bytes);
    const uint64_t groups = uint64_t{1} << (n - window);
    double seconds = 0;
    for (std::size_t p = 0; p < passes; ++p) {
        double link = sb / c.d2h_bytes_per_second;
        if (p > 0) link = std::max(link, sb / c.h2d_bytes_per_second);
        double compute = 0;
        if (!L.tier.passes.empty()) {
            const std::vector<Block> sub = pass_blocks(L.plan.blocks, L.tier, L.tier.passes[p]);
            const uint32_t wp = count_passes(window, r.precision, sub, r.window_qubits, r.max_pass_cost);
            compute = static_cast<double>(wp) * 2.0 * sb / c.pass_bytes_per_second +
                      static_cast<double>(groups * (wp + 2)) * c.launch_seconds;
        }
        seconds += std::max(link, compute);
    }
    seconds += static_cast<double>(L.output_reads) * sb / c.h2d_bytes_per_second;
    L.seconds = seconds;
    return L;
}

// ---- execution

class PinnedState {
public:
    explicit PinnedState(uint64_t bytes) {
        if (hipHostMalloc(&ptr_, bytes, hipHostMallocDefault) != hipSuccess || ptr_ == nullptr) {
            (void)hipGetLastError();
            ptr_ = nullptr;
            fail(MH_ERR_RESOURCE, "could not pin " + std::to_string(bytes) + " bytes of host memory for the state");
        }
    }
    ~PinnedState() {
        if (ptr_) (void)hipHostFree(ptr_);
    }
    PinnedState(const PinnedState&) = delete;
    PinnedState& operator=(const PinnedState&) = delete;
    char* get() const { return static_cast<char*>(ptr_); }

private:
    void* ptr_ = nullptr;
};

// The copy streams (default priority) and the events that order the windows.
class Pipeline {
public:
    Pipeline() {
        try {
            MH_HIP(hipStreamCreateWithFlags(&in, hipStreamNonBlocking));
            MH_HIP(hipStreamCreateWithFlags(&out, hipStreamNonBlocking));
            for (uint32_t b = 0; b < kWindows; ++b) {
                MH_HIP(hipEventCreateWithFlags(&in_done[b], hipEventDisableTiming));
                MH_HIP(hipEventCreateWithFlags(&used[b], hipEventDisableTiming));
                MH_HIP(hipEventCreateWithFlags(&out_done[b], hipEventDisableTiming));
            }
            MH_HIP(hipEventCreateWithFlags(&drained, hipEventDisableTiming));
        } catch (...) {
            release();
            throw;
        }
    }
    ~Pipeline() { release(); }
    Pipeline(const Pipeline&) = delete;
    Pipeline& operator=(const Pipeline&) = delete;

    // Waits for every copy and kernel the tier queued.
    void drain() {
        MH_HIP(hipStreamSynchronize(in));
        MH_HIP(hipStreamSynchronize(out));
        synchronize();
    }

    // The event after chunk c's last write-back, made on first use.
    hipEvent_t chunk_done(uint64_t c) { return lazy(chunk_events_, c, hipEventDisableTiming); }
    // Timing events at the start (0) and after each pass's write-backs.
    hipEvent_t pass_end(std::size_t p) { return lazy(pass_events_, p, hipEventDefault); }

    hipStream_t in = nullptr, out = nullptr;
    hipEvent_t in_done[kWindows] = {}, used[kWindows] = {}, out_done[kWindows] = {};
    hipEvent_t drained = nullptr;
    bool out_pending[kWindows] = {}, use_pending[kWindows] = {};

private:
    std::vector<hipEvent_t> chunk_events_, pass_events_;

    static hipEvent_t lazy(std::vector<hipEvent_t>& list, uint64_t k, unsigned flags) {
        if (list.size() <= k) list.resize(k + 1, nullptr);
        if (list[k] == nullptr) MH_HIP(hipEventCreateWithFlags(&list[k], flags));
        return list[k];
    }

    void release() noexcept {
        if (in) (void)hipStreamSynchronize(in);
        if (out) (void)hipStreamSynchronize(out);
        (void)hipStreamSynchronize(streams().main);
        for (uint32_t b = 0; b < kWindows; ++b) {
            if (in_done[b]) (void)hipEventDestroy(in_done[b]);
            if (used[b]) (void)hipEventDestroy(used[b]);
            if (out_done[b]) (void)hipEventDestroy(out_done[b]);
            in_done[b] = used[b] = out_done[b] = nullptr;
        }
        if (drained) (void)hipEventDestroy(drained);
        drained = nullptr;
        for (auto* list : {&chunk_events_, &pass_events_}) {
            for (hipEvent_t e : *list)
                if (e) (void)hipEventDestroy(e);
            list->clear();
        }
        if (in) (void)hipStreamDestroy(in);
        if (out) (void)hipStreamDestroy(out);
        in = out = nullptr;
    }
};

// Calls f(first slot, first chunk, run length) for the chunks
// base | deposit(s, set), s = 0 .. 2^popcount(set) - 1, of a group's window
// slots, chunks adjacent in both the window and the host state as one run.
template <typename F>
void for_runs(uint64_t base, uint64_t set, F&& f) {
    const uint64_t slots = uint64_t{1} << std::popcount(set);
    uint64_t s = 0;
    while (s < slots) {
        const uint64_t c0 = base | deposit(s, set);
        uint64_t run = 1;
        while (s + run < slots && (base | deposit(s + run, set)) == c0 + run) ++run;
        f(s, c0, run);
        s += run;
    }
}

struct TierRun {
    uint32_t passes_run = 0;
    std::vector<float> pass_ms;
    double gate_seconds = 0;
};

// Every pass of the plan over the host state (an empty plan: one pass that
// only writes the initial state).
TierRun run_tier_passes(const ExecutionRequest& r, const Layout& L, char* host, void* const* windows, Pipeline& pl) {
    const uint32_t n = L.tier.n, chunk = L.tier.chunk, W = L.tier.window;
    const uint32_t globals = n - chunk, w = W - chunk;
    const uint64_t all = globals >= 64 ? ~uint64_t{0} : (uint64_t{1} << globals) - 1;
    const uint64_t chunk_bytes = state_bytes_of(chunk, r.precision);
    const uint64_t window_bytes = state_bytes_of(W, r.precision);
    std::vector<TierPass> passes = L.tier.passes;
    if (passes.empty()) passes.push_back(TierPass{0, 0, 0});
    const cplx phase = host_phase(r.global_phase);
    const uint64_t groups = uint64_t{1} << (globals - std::min(globals, w));
    // A load of a chunk waits for that chunk's last write-back only (an event
    // per chunk), so a pass's first loads run beside the last write-backs of
    // the pass before it; with more chunks than this, a pass waits for the
    // whole of the last one instead.
    const bool tracked = (uint64_t{1} << globals) <= kMaxTrackedChunks;
    std::vector<uint8_t> written(tracked ? (uint64_t{1} << globals) : 0, 0);
    const bool timing = r.record_pass_times;
    if (timing) MH_HIP(hipEventRecord(pl.pass_end(0), streams().main));
    TierRun run;
    using clock = std::chrono::steady_clock;
    const auto t0 = clock::now();
    for (std::size_t p = 0; p < passes.size(); ++p) {
        const uint64_t set = padded_set(passes[p].mask, globals, w);
        const uint64_t rest = all & ~set;
        const std::vector<Block> blocks = pass_blocks(L.plan.blocks, L.tier, passes[p]);
        const bool first = p == 0;
        if (!first && !tracked) MH_HIP(hipEventRecord(pl.drained, pl.out));
        auto issue_in = [&](uint64_t g) {
            const uint32_t b = static_cast<uint32_t>(g % kWindows);
            char* win = static_cast<char*>(windows[b]);
            if (pl.out_pending[b]) MH_HIP(hipStreamWaitEvent(pl.in, pl.out_done[b], 0));
            if (!tracked) MH_HIP(hipStreamWaitEvent(pl.in, pl.drained, 0));
            for_runs(deposit(g, rest), set, [&](uint64_t s0, uint64_t c0, uint64_t len) {
                if (tracked)
                    for (uint64_t c = c0; c < c0 + len; ++c)
                        if (written[c]) MH_HIP(hipStreamWaitEvent(pl.in, pl.chunk_done(c), 0));
                MH_HIP(hipMemcpyAsync(win + s0 * chunk_bytes, host + c0 * chunk_bytes, len * chunk_bytes,
                                      hipMemcpyHostToDevice, pl.in));
            });
            MH_HIP(hipEventRecord(pl.in_done[b], pl.in));
        };
        if (!first)
            for (uint64_t g = 0; g < std::min<uint64_t>(groups, kWindows); ++g) issue_in(g);
        for (uint64_t g = 0; g < groups; ++g) {
            const uint32_t b = static_cast<uint32_t>(g % kWindows);
            char* win = static_cast<char*>(windows[b]);
            if (first) {
                if (pl.out_pending[b]) MH_HIP(hipStreamWaitEvent(streams().main, pl.out_done[b], 0));
                if (deposit(g, rest) == 0)
                    launch_init(win, W, r.precision, phase.real(), phase.imag());
                else
                    MH_HIP(hipMemsetAsync(win, 0, window_bytes, streams().main));
            } else {
                MH_HIP(hipStreamWaitEvent(streams().main, pl.in_done[b], 0));
            }
            if (!blocks.empty())
                apply_blocks(win, W, r.precision, blocks, r.window_qubits, false, nullptr, r.max_pass_cost);
            MH_HIP(hipEventRecord(pl.used[b], streams().main));
            MH_HIP(hipStreamWaitEvent(pl.out, pl.used[b], 0));
            for_runs(deposit(g, rest), set, [&](uint64_t s0, uint64_t c0, uint64_t len) {
                MH_HIP(hipMemcpyAsync(host + c0 * chunk_bytes, win + s0 * chunk_bytes, len * chunk_bytes,
                                      hipMemcpyDeviceToHost, pl.out));
                if (tracked)
                    for (uint64_t c = c0; c < c0 + len; ++c) {
                        MH_HIP(hipEventRecord(pl.chunk_done(c), pl.out));
                        written[c] = 1;
                    }
            });
            MH_HIP(hipEventRecord(pl.out_done[b], pl.out));
            pl.out_pending[b] = true;
            if (!first && g + kWindows < groups) issue_in(g + kWindows);
        }
        if (timing) MH_HIP(hipEventRecord(pl.pass_end(p + 1), pl.out));
        run.passes_run = static_cast<uint32_t>(p + 1);
        if (r.monitor) {
            // The pass's safe boundary: every chunk holds the state after its
            // last block (reference chunked_executor.cpp:216-221).
            MH_HIP(hipStreamSynchronize(pl.out));
            if (r.monitor(r.monitor_data, p + 1) != 0) {
                pl.drain();
                fail(MH_ERR_ABANDONED, "the execution monitor abandoned the run after " + std::to_string(p + 1) + " of " +
                                           std::to_string(passes.size()) + " passes");
            }
        }
    }
    pl.drain();
    run.gate_seconds = std::chrono::duration<double>(clock::now() - t0).count();
    if (timing) {
        run.pass_ms.resize(passes.size());
        for (std::size_t p = 0; p < passes.size(); ++p)
            MH_HIP(hipEventElapsedTime(&run.pass_ms[p], pl.pass_end(p), pl.pass_end(p + 1)));
    }
    return run;
}

// Reads the final state back through the windows for counts and raw terms.
void tier_outputs(const ExecutionRequest& r, const Layout& L, char* host, void* const* windows, Pipeline& pl,
                  ExecutionResult& res) {
    const uint32_t n = L.tier.n, W = L.tier.window;
    const uint64_t window_bytes = state_bytes_of(W, r.precision);
    const uint64_t ranges = uint64_t{1} << (n - W);
    const TermSets sets = split_terms(r.terms, W);
    std::vector<cplx> terms(r.terms.size(), cplx(0.0, 0.0));
    for (uint32_t b = 0; b < kWindows; ++b) pl.use_pending[b] = false;

    // One streaming read: load(i) copies window i's memory (ready at in_done),
    // use(i, window) queues or runs its work on the main stream.
    auto stream_read = [&](uint64_t count, const auto& load, const auto& use) {
        auto issue = [&](uint64_t i) {
            const uint32_t b = static_cast<uint32_t>(i % kWindows);
            if (pl.use_pending[b]) MH_HIP(hipStreamWaitEvent(pl.in, pl.used[b], 0));
            load(i, static_cast<char*>(windows[b]));
            MH_HIP(hipEventRecord(pl.in_done[b], pl.in));
        };
        for (uint64_t i = 0; i < std::min<uint64_t>(count, kWindows); ++i) issue(i);
        for (uint64_t i = 0; i < count; ++i) {
            const uint32_t b = static_cast<uint32_t>(i % kWindows);
            MH_HIP(hipStreamWaitEvent(streams().main, pl.in_done[b], 0));
            use(i, windows[b]);
            MH_HIP(hipEventRecord(pl.used[b], streams().main));
            pl.use_pending[b] = true;
            if (i + kWindows < count) issue(i + kWindows);
        }
        pl.drain();
    };

    using clock = std::chrono::steady_clock;
    // Counts and the terms with X inside the window: index order.
    if (r.shots > 0 || !sets.local.empty()) {
        const auto s0 = clock::now();
        std::unique_ptr<ChunkedSampler> sampler;
        std::vector<double> u;
        if (r.shots > 0) {
            u = sorted_uniforms(r.seed, r.shots);
            sampler = std::make_unique<ChunkedSampler>(n, W, r.precision, u.data(), r.shots, streams().main);
        }
        std::vector<uint64_t> xs, zs;
        for (uint32_t t : sets.local) {
            xs.push_back(r.terms[t].first & ((uint64_t{1} << W) - 1));
            zs.push_back(r.terms[t].second & ((uint64_t{1} << W) - 1));
        }
        std::vector<cplx> part(sets.local.size());
        stream_read(
            ranges,
            [&](uint64_t i, char* win) {
                MH_HIP(hipMemcpyAsync(win, host + i * window_bytes, window_bytes, hipMemcpyHostToDevice, pl.in));
            },
            [&](uint64_t i, void* win) {
                if (sampler) sampler->add(win);
                if (sets.local.empty()) return;
                expectation_terms(win, W, r.precision, xs.data(), zs.data(), static_cast<uint32_t>(xs.size()),
                                  part.data());
                // The Z bits above the window give every amplitude of window i
                // the sign (-1)^popcount((z >> W) & i).
                for (std::size_t k = 0; k < sets.local.size(); ++k) {
                    const uint32_t t = sets.local[k];
                    const bool neg = W < 64 && (std::popcount((r.terms[t].second >> W) & i) & 1) != 0;
                    terms[t] += neg ? -part[k] : part[k];
                }
            });
        if (sampler) {
            res.counts = fold_counts(sampler->basis(), r.measure_targets);
            res.sample_seconds = std::chrono::duration<double>(clock::now() - s0).count();
        }
    }
    // Terms whose X mask reaches above the window: for h = x >> (W - 1), the
    // pairs of half-windows (c, c ^ h) with bit msb(h) of c clear, slot 0 and
    // slot 1 of a window. In the window the term is x' = (x mod 2^(W-1)) +
    // 2^(W-1) and z' = (z mod 2^(W-1)) + b 2^(W-1), b = popcount(zh & h) mod 2
    // (zh = z >> (W - 1)); the pair adds sign(c) * (-1)^((popcount(zh & h) - b) / 2)
    // times the window's raw term, sign(c) = (-1)^popcount(zh & c): the signs
    // of the high Z bits and the part of i^popcount(x & z) the window's own
    // phase does not carry.
    const uint32_t half = W - 1;
    const uint64_t half_bytes = window_bytes / 2;
    const uint64_t low = (uint64_t{1} << half) - 1;
    for (const auto& [h, members] : sets.paired) {
        const uint32_t top = 63 - static_cast<uint32_t>(std::countl_zero(h));
        const uint64_t pairs_mask = ((uint64_t{1} << (n - half)) - 1) & ~(uint64_t{1} << top);
        std::vector<uint64_t> xs, zs;
        std::vector<double> scale;
        for (uint32_t t : members) {
            const uint64_t x = r.terms[t].first, z = r.terms[t].second;
            const uint64_t zh = z >> half;
            const uint32_t pc = static_cast<uint32_t>(std::popcount(zh & h));
            const uint32_t b = pc & 1u;
            xs.push_back((x & low) | (uint64_t{1} << half));
            zs.push_back((z & low) | (uint64_t{b} << half));
            scale.push_back(((pc - b) / 2) % 2 ? -1.0 : 1.0);
        }
        std::vector<cplx> part(members.size());
        const auto e0 = clock::now();
        stream_read(
            ranges,
            [&](uint64_t i, char* win) {
                const uint64_t c = deposit(i, pairs_mask);
                MH_HIP(hipMemcpyAsync(win, host + c * half_bytes, half_bytes, hipMemcpyHostToDevice, pl.in));
                MH_HIP(hipMemcpyAsync(win + half_bytes, host + (c ^ h) * half_bytes, half_bytes, hipMemcpyHostToDevice,
                                      pl.in));
            },
            [&](uint64_t i, void* win) {
                const uint64_t c = deposit(i, pairs_mask);
                expectation_terms(win, W, r.precision, xs.data(), zs.data(), static_cast<uint32_t>(xs.size()),
                                  part.data());
                for (std::size_t k = 0; k < members.size(); ++k) {
                    const uint32_t t = members[k];
                    const bool neg = (std::popcount((r.terms[t].second >> half) & c) & 1) != 0;
                    terms[t] += (neg ? -scale[k] : scale[k]) * part[k];
                }
            });
        res.expectation_seconds += std::chrono::duration<double>(clock::now() - e0).count();
    }
    if (!r.terms.empty()) res.expectation_terms = std::move(terms);
}

struct TierOutcome {
    ExecutionResult result;
    Layout layout;
};

TierOutcome tier_execute(const ExecutionRequest& r, const TierRequest& t, void* statevector_out) {
    using clock = std::chrono::steady_clock;
    const auto t0 = clock::now();
    TierOutcome o;
    o.layout = lay_out(r, t);
    const Layout& L = o.layout;
    DeviceLedger& ledger = DeviceLedger::current();
    ledger.reset();
    context();
    PinnedState host(L.state_bytes);
    ExecutionResult& res = o.result;
    res.num_blocks = static_cast<uint32_t>(L.plan.blocks.size());
    res.fusion_used = L.plan.fusion_used;
    res.width_used = L.plan.width_used;
    res.fused_blocks = L.plan.fusion.fused_blocks;
    res.gates_absorbed = L.plan.fusion.gates_absorbed;
    {
        const uint64_t window_bytes = state_bytes_of(L.tier.window, r.precision);
        ChargedBuffer w0(window_bytes), w1(window_bytes);
        void* const windows[kWindows] = {w0.get(), w1.get()};
        // Declared after the windows: its destructor waits for the copies and
        // kernels that use them before they are freed.
        Pipeline pl;
        const TierRun run = run_tier_passes(r, L, host.get(), windows, pl);
        res.num_passes = static_cast<uint32_t>(L.tier.passes.size());
        res.pass_ms = run.pass_ms;
        res.gate_seconds = run.gate_seconds;
        tier_outputs(r, L, host.get(), windows, pl, res);
        res.actual_bytes = kWindows * window_bytes;
        res.peak_device_bytes = std::max(ledger.peak(), res.actual_bytes);
    }
    if (r.want_statevector && statevector_out) std::memcpy(statevector_out, host.get(), L.state_bytes);
    res.execute_seconds = std::chrono::duration<double>(clock::now() - t0).count();
    return o;
}

TierRequest to_tier(const mh_tier_request* t) {
    TierRequest q;
    if (t) {
        q.chunk_qubits = t->chunk_qubits;
        q.window_qubits = t->window_qubits;
    }
    return q;
}

void write_cost(const Layout& L, const mh_tier_request* t, mh_tier_cost* out) {
    out->feasible = 1;
    out->passes = L.tier.passes.size();
    out->chunk_qubits = L.tier.chunk;
    out->window_qubits = L.tier.window;
    out->link_bytes = L.link_bytes;
    out->host_state_bytes = L.state_bytes;
    out->link_write_bytes = L.link_write_bytes;
    out->global_blocks = L.tier.global_blocks;
    out->max_global_operands = L.tier.max_global;
    out->blocks = static_cast<uint32_t>(L.plan.blocks.size());
    out->device_bytes = L.device_bytes;
    out->seconds = L.seconds;
    out->fusion_used = L.plan.fusion_used;
    out->fusion_width_used = L.plan.width_used;
    if (t && t->passes) {
        const std::size_t m = std::min<std::size_t>(t->pass_capacity, L.tier.passes.size());
        for (std::size_t i = 0; i < m; ++i)
            t->passes[i] = mh_tier_pass{L.tier.passes[i].mask, L.tier.passes[i].first, L.tier.passes[i].end};
    }
}

}  // namespace
}  // namespace mh

using namespace mh;

extern "C" {

int mh_tier_estimate(const mh_execution_request* request, const mh_tier_request* tier, mh_tier_cost* out) {
    return guarded([&] {
        if (out == nullptr) fail(MH_ERR_INVALID, "null output");
        std::memset(out, 0, sizeof(*out));
        try {
            write_cost(lay_out(to_request(request), to_tier(tier)), tier, out);
        } catch (const std::exception& ex) {
            out->feasible = 0;
            std::strncpy(out->infeasible_reason, ex.what(), sizeof(out->infeasible_reason) - 1);
        }
    });
}

int mh_tier_execute(const mh_execution_request* request, const mh_tier_request* tier, mh_execution_result* out,
                    mh_tier_cost* plan) {
    return guarded([&] {
        if (out == nullptr) fail(MH_ERR_INVALID, "null result");
        out->region_sampler = nullptr;
        const ExecutionRequest r = to_request(request);
        TierOutcome o = tier_execute(r, to_tier(tier), out->statevector);
        o.result.statevector.reset();
        write_result(o.result, out);
        if (plan) {
            std::memset(plan, 0, sizeof(*plan));
            write_cost(o.layout, tier, plan);
        }
    });
}

}  // extern "C"
