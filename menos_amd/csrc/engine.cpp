// HipDenseEngine and its C ABI (menos_hip.h).
//
// execute() follows the reference engines' order (CudaDenseEngine::execute,
// src/engines/dense_cuda/cuda_dense_engine.cpp:325-412, and DenseEngine::execute,
// src/engines/dense/dense_engine.cpp:333-398): plan first, then the state
// (budget-checked before allocation), initialise, sweep the blocks, then
// counts, expectation terms and finally the statevector. The differences are
// where the work runs: the sweep is a few LDS-window passes instead of one
// launch-and-sync per block, and counts come from the device sampler (exact,
// sampler.hip) instead of a copy-back and a host walk.
#include "engine.hpp"
#include "gate_table.h"
#include "window.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <random>

namespace mh {

DeviceLedger& DeviceLedger::current() {
    thread_local DeviceLedger ledger;
    return ledger;
}

void DeviceLedger::charge(uint64_t bytes) {
    live_ += bytes;
    peak_ = std::max(peak_, live_);
}

void DeviceLedger::release(uint64_t bytes) { live_ -= std::min(live_, bytes); }

void DeviceLedger::reset() { peak_ = live_; }

namespace {

constexpr uint32_t kMaxQubits = 40;

uint64_t state_bytes(uint32_t n, int32_t precision) { return (uint64_t{1} << n) * amp_bytes(precision); }

// Device scratch beside the state: the expectation partials and the sampler's
// per-block tables (sampler.hip: 2^(n-13) blocks of about 100 bytes).
uint64_t scratch_bytes(uint32_t n, uint64_t shots, std::size_t terms) {
    uint64_t bytes = 0;
    if (terms) bytes = std::max<uint64_t>(bytes, 2048ull * 16 * 16 + 4096);
    if (shots) {
        const uint64_t blocks = n > 13 ? (uint64_t{1} << (n - 13)) : 1;
        bytes = std::max<uint64_t>(bytes, blocks * 112 + shots * 16 + 4096);
    }
    return bytes + (64ull << 10);  // pass programs
}

// The calibrated time model. Defaults are MI355X datasheet figures derated to
// what a streaming kernel reaches; mh_calibrate() replaces them by measurement.
mh_calibration g_calibration = {
    /*hbm_bytes_per_second=*/6.0e12, /*pass_bytes_per_second=*/4.0e12, /*launch_seconds=*/8e-6,
    /*h2d_bytes_per_second=*/50e9, /*d2h_bytes_per_second=*/50e9, /*fp64_ops_per_second=*/30e12,
    /*mfma_flops_per_second=*/100e12};

double word_to_uniform(uint64_t w) { return static_cast<double>(w >> 11) * (1.0 / 9007199254740992.0); }

uint32_t window_for(const ExecutionRequest& r) {
    return r.window_qubits == 0 ? window_default(r.precision) : r.window_qubits;
}

// The calibrated time model's device seconds for the gate passes of a block
// list (launch costs apart): a window pass moves the state in and out at the measured pass rate; a
// block on a pass of its own (the per-group kernel, or the matrix cores for
// complex64 blocks of 5-6 qubits) takes the larger of its HBM time and its
// arithmetic (8 * 2^k operations per amplitude) at the measured FP64 rate
// (complex128, no FMA), twice that (complex64 FMA) or the matrix-core rate.
double pass_seconds(uint32_t n, int32_t precision, const std::vector<Block>& blocks, uint32_t w, uint32_t max_cost) {
    const mh_calibration& c = g_calibration;
    const double sb = static_cast<double>(state_bytes(n, precision));
    const double amps = static_cast<double>(uint64_t{1} << n);
    double seconds = 0;
    for (const WindowPass& pass : plan_window_passes(n, precision, blocks, w, max_cost == 0 ? kDefaultPassCost : max_cost)) {
        if (!pass.generic) {
            seconds += 2.0 * sb / c.pass_bytes_per_second;
            continue;
        }
        const uint32_t k = static_cast<uint32_t>(blocks[pass.blocks[0]].operands.size());
        const double ops = 8.0 * static_cast<double>(1u << k) * amps;
        const double rate = precision == MH_C128 ? c.fp64_ops_per_second
                            : (precision == MH_C64 && k >= 5 && n >= k + 5) ? c.mfma_flops_per_second
                                                                           : 2.0 * c.fp64_ops_per_second;
        seconds += std::max(2.0 * sb / c.hbm_bytes_per_second, ops / rate);
    }
    return seconds;
}

}  // namespace

const mh_calibration& calibration() { return g_calibration; }

std::vector<double> sorted_uniforms(uint64_t seed, uint64_t shots) {
    std::mt19937_64 rng(seed);
    std::vector<double> u(shots);
    for (auto& x : u) x = word_to_uniform(rng());
    std::sort(u.begin(), u.end());
    return u;
}

std::map<uint64_t, uint64_t> fold_counts(const std::vector<uint64_t>& basis,
                                         const std::vector<std::pair<uint32_t, uint32_t>>& targets) {
    std::map<uint64_t, uint64_t> counts;
    for (uint64_t b : basis) {
        uint64_t key = 0;
        for (const auto& [q, c] : targets)
            if ((b >> q) & 1ULL) key |= uint64_t{1} << c;
        counts[key] += 1;
    }
    return counts;
}

Plan make_plan(const ExecutionRequest& r) {
    validate_circuit(r.circuit);
    const CircuitView cv{r.circuit};
    Plan p;
    if (r.fusion != MH_FUSE_AUTO) {
        p.fusion = plan_fusion(cv, r.fusion, r.fusion_width);
        p.blocks = build_blocks(cv, p.fusion);
        p.fusion_used = r.fusion;
        p.width_used = r.fusion == MH_FUSE_WIDTH ? r.fusion_width : 0;
        return p;
    }
    // The autotuner: containment first, then widths 2..6; strictly faster wins.
    const uint32_t n = r.circuit->num_qubits;
    double best = 0;
    bool have = false;
    for (uint32_t cand = 0; cand <= 5; ++cand) {
        const int32_t fusion = cand == 0 ? MH_FUSE_CONTAINMENT : MH_FUSE_WIDTH;
        const uint32_t width = cand == 0 ? 0 : cand + 1;
        Plan q;
        q.fusion = plan_fusion(cv, fusion, width);
        q.blocks = build_blocks(cv, q.fusion);
        q.fusion_used = fusion;
        q.width_used = width;
        const double t = n < window_default(r.precision)
                             ? static_cast<double>(q.blocks.size())
                             : pass_seconds(n, r.precision, q.blocks, window_for(r), r.max_pass_cost) +
                                   g_calibration.launch_seconds *
                                       count_passes(n, r.precision, q.blocks, window_for(r), r.max_pass_cost);
        if (!have || t < best) {
            best = t;
            p = std::move(q);
            have = true;
        }
    }
    return p;
}

CostEstimate HipDenseEngine::estimate(const ExecutionRequest& r) const noexcept {
    CostEstimate e;
    try {
        amp_bytes(r.precision);
        if (r.circuit == nullptr) fail(MH_ERR_INVALID, "null circuit");
        const uint32_t n = r.circuit->num_qubits;
        if (n > kMaxQubits) fail(MH_ERR_INVALID, "num_qubits above 40");
        const Plan p = make_plan(r);
        e.blocks = static_cast<uint32_t>(p.blocks.size());
        e.passes = count_passes(n, r.precision, p.blocks, window_for(r), r.max_pass_cost);
        e.fusion_used = p.fusion_used;
        e.width_used = p.width_used;
        const uint64_t sb = state_bytes(n, r.precision);
        e.device_bytes = sb + scratch_bytes(n, r.shots, r.terms.size());
        e.host_bytes = r.want_statevector ? sb : 0;
        // Work units as the reference counts them (dense_engine.cpp:309-331).
        uint64_t work = uint64_t{1} << n;
        for (const Block& b : p.blocks) work += uint64_t{1} << (n + b.operands.size());
        if (r.want_statevector) work += uint64_t{1} << n;
        if (r.shots) work += uint64_t{1} << n;
        work += r.terms.size() * (uint64_t{1} << n);
        e.work_units = work;
        e.launches = 1 + e.passes + (r.shots ? 7 : 0) + 2 * r.terms.size();
        e.h2d_bytes = 4096 + r.shots * 8;
        e.d2h_bytes = (r.want_statevector ? sb : 0) + r.shots * 8 + r.terms.size() * 16;
        const mh_calibration& c = g_calibration;
        double seconds = static_cast<double>(e.launches) * c.launch_seconds;
        seconds += n < window_default(r.precision)
                       ? static_cast<double>(e.passes) * 2.0 * static_cast<double>(sb) / c.pass_bytes_per_second
                       : pass_seconds(n, r.precision, p.blocks, window_for(r), r.max_pass_cost);
        seconds += static_cast<double>(sb) / c.hbm_bytes_per_second;  // initialisation
        seconds += (r.shots ? 3.0 : 0.0) * static_cast<double>(sb) / c.hbm_bytes_per_second;
        seconds += static_cast<double>(r.terms.size()) * static_cast<double>(sb) / c.hbm_bytes_per_second;
        seconds += static_cast<double>(e.h2d_bytes) / c.h2d_bytes_per_second;
        seconds += static_cast<double>(e.d2h_bytes) / c.d2h_bytes_per_second;
        e.seconds = seconds;
        if (r.budget_bytes != 0 && e.device_bytes > r.budget_bytes)
            fail(MH_ERR_RESOURCE, "needs " + std::to_string(e.device_bytes) + " device bytes, budget " +
                                      std::to_string(r.budget_bytes));
        e.feasible = true;
    } catch (const std::exception& ex) {
        e.feasible = false;
        e.infeasible_reason = ex.what();
    }
    return e;
}

std::map<uint64_t, uint64_t> sample_counts(const void* state, uint32_t n, int32_t precision, uint64_t shots,
                                           uint64_t seed, const std::vector<std::pair<uint32_t, uint32_t>>& targets) {
    const std::vector<double> u = sorted_uniforms(seed, shots);
    std::vector<uint64_t> basis(shots);
    sample_sorted(state, n, precision, u.data(), shots, basis.data());
    return fold_counts(basis, targets);
}

ExecutionResult HipDenseEngine::execute(const ExecutionRequest& r) const {
    using clock = std::chrono::steady_clock;
    const auto t0 = clock::now();
    amp_bytes(r.precision);
    const Plan p = make_plan(r);
    const uint32_t n = r.circuit->num_qubits;
    if (r.bitstring_width > 64) fail(MH_ERR_INVALID, "bitstring_width above 64");
    for (const auto& [q, c] : r.measure_targets)
        if (q >= n || c >= std::max<uint32_t>(r.bitstring_width, 1)) fail(MH_ERR_INVALID, "measure target out of range");

    DeviceLedger& ledger = DeviceLedger::current();
    ledger.reset();
    const uint64_t sb = state_bytes(n, r.precision);
    const uint64_t need = sb + scratch_bytes(n, r.shots, r.terms.size());
    if (r.budget_bytes != 0 && need > r.budget_bytes)
        fail(MH_ERR_RESOURCE, "the dense state needs " + std::to_string(need) + " device bytes, budget is " +
                                  std::to_string(r.budget_bytes));
    context();

    ExecutionResult result;
    result.num_blocks = static_cast<uint32_t>(p.blocks.size());
    result.fusion_used = p.fusion_used;
    result.width_used = p.width_used;
    result.fused_blocks = p.fusion.fused_blocks;
    result.gates_absorbed = p.fusion.gates_absorbed;
    {
        ChargedBuffer state(sb);
        const std::complex<double> phase = host_phase(r.global_phase);
        launch_init(state.get(), n, r.precision, phase.real(), phase.imag());

        hipEvent_t g0 = nullptr, g1 = nullptr;
        struct EventPair {
            hipEvent_t& a;
            hipEvent_t& b;
            ~EventPair() {
                if (a) (void)hipEventDestroy(a);
                if (b) (void)hipEventDestroy(b);
            }
        } events{g0, g1};
        MH_HIP(hipEventCreate(&g0));
        MH_HIP(hipEventCreate(&g1));
        MH_HIP(hipEventRecord(g0, streams().main));
        PassStats stats;
        if (r.monitor) {
            const mh_monitor_fn fn = r.monitor;
            void* const data = r.monitor_data;
            stats.monitor = [fn, data](uint64_t crossed) { return fn(data, crossed) != 0; };
        }
        apply_blocks(state.get(), n, r.precision, p.blocks, window_for(r), r.record_pass_times, &stats,
                     r.max_pass_cost);
        MH_HIP(hipEventRecord(g1, streams().main));
        synchronize();
        float gate_ms = 0;
        MH_HIP(hipEventElapsedTime(&gate_ms, g0, g1));
        result.gate_seconds = gate_ms * 1e-3;
        result.num_passes = stats.passes;
        result.pass_ms = std::move(stats.pass_ms);

        if (r.shots > 0) {
            const auto s0 = clock::now();
            result.counts = sample_counts(state.get(), n, r.precision, r.shots, r.seed, r.measure_targets);
            result.sample_seconds = std::chrono::duration<double>(clock::now() - s0).count();
        }
        if (!r.terms.empty()) {
            const auto e0 = clock::now();
            std::vector<uint64_t> xs, zs;
            for (const auto& [x, z] : r.terms) {
                xs.push_back(x);
                zs.push_back(z);
            }
            std::vector<cplx> out(r.terms.size());
            expectation_terms(state.get(), n, r.precision, xs.data(), zs.data(), static_cast<uint32_t>(xs.size()),
                              out.data());
            result.expectation_terms = std::move(out);
            result.expectation_seconds = std::chrono::duration<double>(clock::now() - e0).count();
        }
        if (r.want_statevector) {
            std::vector<unsigned char> host(sb);
            download(host.data(), state.get(), sb);
            result.statevector = std::move(host);
        }
        result.actual_bytes = sb;
        result.peak_device_bytes = std::max(ledger.peak(), sb);
        // Last, so that nothing after it can fail and leak the state.
        if (r.want_region_sampler) result.region_sampler = make_region_sampler(state.release(), n, r.precision);
    }
    result.execute_seconds = std::chrono::duration<double>(clock::now() - t0).count();
    return result;
}

}  // namespace mh

namespace mh {

ExecutionRequest to_request(const mh_execution_request* q) {
    if (q == nullptr || q->circuit == nullptr) fail(MH_ERR_INVALID, "null request or circuit");
    ExecutionRequest r;
    r.circuit = q->circuit;
    r.precision = q->precision;
    r.fusion = q->fusion;
    r.fusion_width = q->fusion_width;
    r.window_qubits = q->window_qubits;
    r.global_phase = q->global_phase;
    r.want_statevector = q->want_statevector != 0;
    r.shots = q->shots;
    r.seed = q->seed;
    for (uint32_t t = 0; t < q->num_targets; ++t) r.measure_targets.emplace_back(q->target_qubits[t], q->target_clbits[t]);
    r.bitstring_width = q->bitstring_width;
    for (uint32_t t = 0; t < q->num_terms; ++t) r.terms.emplace_back(q->x_masks[t], q->z_masks[t]);
    r.budget_bytes = q->budget_bytes;
    r.record_pass_times = q->record_pass_times != 0;
    r.max_pass_cost = q->max_pass_cost;
    r.monitor = q->monitor;
    r.monitor_data = q->monitor_data;
    r.want_region_sampler = q->want_region_sampler != 0;
    return r;
}

void write_result(ExecutionResult& res, mh_execution_result* out) {
    out->region_sampler = res.region_sampler;
    res.region_sampler = nullptr;
    if (res.statevector && out->statevector) std::memcpy(out->statevector, res.statevector->data(), res.statevector->size());
    out->num_outcomes = 0;
    if (res.counts && out->outcomes && out->counts) {
        uint64_t at = 0;
        for (const auto& [k, v] : *res.counts) {
            out->outcomes[at] = k;
            out->counts[at] = v;
            ++at;
        }
        out->num_outcomes = at;
    }
    if (res.expectation_terms && out->expectation_terms)
        std::memcpy(out->expectation_terms, res.expectation_terms->data(), res.expectation_terms->size() * sizeof(cplx));
    if (out->pass_ms) {
        const std::size_t m = std::min<std::size_t>(out->pass_ms_capacity, res.pass_ms.size());
        std::copy(res.pass_ms.begin(), res.pass_ms.begin() + static_cast<std::ptrdiff_t>(m), out->pass_ms);
    }
    out->num_blocks = res.num_blocks;
    out->num_passes = res.num_passes;
    out->fusion_used = res.fusion_used;
    out->fusion_width_used = res.width_used;
    out->fused_blocks = res.fused_blocks;
    out->gates_absorbed = res.gates_absorbed;
    out->actual_bytes = res.actual_bytes;
    out->peak_device_bytes = res.peak_device_bytes;
    out->execute_seconds = res.execute_seconds;
    out->gate_seconds = res.gate_seconds;
    out->sample_seconds = res.sample_seconds;
    out->expectation_seconds = res.expectation_seconds;
}

}  // namespace mh

using namespace mh;

extern "C" {

int mh_engine_estimate(const mh_execution_request* request, mh_cost_estimate* out) {
    return guarded([&] {
        if (out == nullptr) fail(MH_ERR_INVALID, "null output");
        std::memset(out, 0, sizeof(*out));
        CostEstimate e;
        try {
            e = HipDenseEngine{}.estimate(to_request(request));
        } catch (const std::exception& ex) {
            e.feasible = false;
            e.infeasible_reason = ex.what();
        }
        out->feasible = e.feasible ? 1 : 0;
        std::strncpy(out->infeasible_reason, e.infeasible_reason.c_str(), sizeof(out->infeasible_reason) - 1);
        out->device_bytes = e.device_bytes;
        out->host_bytes = e.host_bytes;
        out->work_units = e.work_units;
        out->blocks = e.blocks;
        out->passes = e.passes;
        out->launches = e.launches;
        out->h2d_bytes = e.h2d_bytes;
        out->d2h_bytes = e.d2h_bytes;
        out->seconds = e.seconds;
        out->fusion_used = e.fusion_used;
        out->fusion_width_used = e.width_used;
    });
}

int mh_engine_execute(const mh_execution_request* request, mh_execution_result* out) {
    return guarded([&] {
        if (out == nullptr) fail(MH_ERR_INVALID, "null result");
        out->region_sampler = nullptr;
        const ExecutionRequest r = to_request(request);
        ExecutionResult res = HipDenseEngine{}.execute(r);
        write_result(res, out);
    });
}

}  // extern "C"

namespace mh {
void set_calibration(const mh_calibration& c) { g_calibration = c; }
}  // namespace mh

extern "C" {

int mh_set_calibration(const mh_calibration* profile) {
    return guarded([&] {
        if (profile == nullptr) fail(MH_ERR_INVALID, "null profile");
        const double rates[] = {profile->hbm_bytes_per_second, profile->pass_bytes_per_second,
                                profile->h2d_bytes_per_second, profile->d2h_bytes_per_second,
                                profile->fp64_ops_per_second, profile->mfma_flops_per_second};
        for (double r : rates)
            if (!(r > 0.0) || !std::isfinite(r)) fail(MH_ERR_INVALID, "profile rates must be finite and positive");
        if (!(profile->launch_seconds >= 0.0) || !std::isfinite(profile->launch_seconds))
            fail(MH_ERR_INVALID, "launch_seconds must be finite and non-negative");
        set_calibration(*profile);
    });
}

int mh_get_calibration(mh_calibration* out) {
    return guarded([&] {
        if (out == nullptr) fail(MH_ERR_INVALID, "null output");
        *out = g_calibration;
    });
}

}  // extern "C"
