// The SimulationEngine adapter over libmenos_hip.so (see hip_dense_engine.h).
// Reference counterpart: CudaDenseEngine, src/engines/dense_cuda/cuda_dense_engine.cpp:182-412.
#include "hip_dense_engine.h"

#include <algorithm>
#include <complex>
#include <exception>
#include <map>
#include <memory>
#include <optional>
#include <string>
#include <type_traits>
#include <utility>
#include <variant>

#include "common/error.h"
#include "common/memory_ledger.h"
#include "common/rng.h"
#include "ir/noise.h"

namespace menos::engines::dense_hip {

namespace {

constexpr std::int32_t kOpUnitary = 21;
constexpr std::int32_t kOpMeasure = 22;

// The ABI status as the reference's exception types (C11: ResourceExceeded
// names the bytes; a HIP failure is InternalError).
[[noreturn]] void throw_status(int status, const char* what) {
    const std::string message = std::string(what) + ": " + mh_last_error();
    switch (status) {
        case MH_ERR_RESOURCE: throw common::ResourceExceeded(message);
        case MH_ERR_INVALID: throw common::InvalidCircuit(message);
        default: throw common::InternalError(message);
    }
}

// The request's output and term fields, shared by estimate and execute.
struct Outputs {
    std::vector<std::uint64_t> x, z;
    std::vector<std::uint32_t> target_qubits, target_clbits;
};

// Reasons this engine cannot take a region, before anything is costed; empty
// when it can. The ABI packs an outcome and a Pauli mask in 64-bit words.
std::optional<std::string> unsupported(std::uint32_t num_qubits, std::span<const Operation> operations,
                                       bool want_expectation, const Observable* observable,
                                       std::uint32_t bitstring_width) {
    if (const std::optional<std::string> reason = ir::first_noise_reason(operations); reason.has_value())
        return *reason;
    if (num_qubits > 40) return "HIP dense region of " + std::to_string(num_qubits) + " qubits: at most 40";
    if (bitstring_width > 64) return "HIP dense: bit strings wider than 64 bits are not supported";
    if (want_expectation && observable != nullptr) {
        for (const auto& [pauli, coefficient] : observable->terms) {
            (void)coefficient;
            for (std::size_t w = 1; w < pauli.x_words.size(); ++w)
                if (pauli.x_words[w] != 0 || pauli.z_words[w] != 0)
                    return std::string("HIP dense: a Pauli term beyond qubit 63");
        }
    }
    return std::nullopt;
}

Outputs outputs_of(bool want_expectation, const Observable* observable,
                   std::span<const std::pair<QubitId, ClbitId>> targets) {
    Outputs o;
    if (want_expectation && observable != nullptr) {
        for (const auto& [pauli, coefficient] : observable->terms) {
            (void)coefficient;
            o.x.push_back(pauli.x_words.empty() ? 0 : pauli.x_words[0]);
            o.z.push_back(pauli.z_words.empty() ? 0 : pauli.z_words[0]);
        }
    }
    for (const auto& [q, c] : targets) {
        o.target_qubits.push_back(q);
        o.target_clbits.push_back(c);
    }
    return o;
}

// The monitor across the C ABI: the reference abandons by throwing
// (simulation_engine.h:54-116); the ABI takes a nonzero return instead. The
// bridge keeps the monitor's exception and the adapter rethrows it.
struct MonitorBridge {
    const ExecutionMonitor* monitor = nullptr;
    std::exception_ptr thrown;

    static int call(void* self, std::uint64_t crossed) {
        auto* b = static_cast<MonitorBridge*>(self);
        try {
            b->monitor->at_safe_boundary(crossed);
            return 0;
        } catch (...) {
            b->thrown = std::current_exception();
            return 1;
        }
    }
};

BitString bitstring_of(std::uint64_t word, std::uint32_t width) {
    BitString key(width);
    for (std::uint32_t b = 0; b < width; ++b)
        if ((word >> b) & 1u) key.set(b, true);
    return key;
}

// A non-sole region's sampler over the final state kept on the device
// (mh_region_sampler_*), in place of the reference's CumulativeDistribution over
// a host copy (cumulative_distribution.cpp:14-92; the CUDA engine copies the
// state back for it, cuda_dense_engine.cpp:386-399). sample_one draws exactly
// one word from the region's stream per call, as the reference does, and
// returns the same outcome bit for bit. Draws go to the device in batches: on
// the first call (and whenever a batch is used up) a copy of the caller's Rng
// runs ahead for the shots the request asked for, and each call then checks
// that its own word is the one predicted (a caller that draws from the stream
// between calls gets a single-word draw instead). The device memory is charged
// to the ledger of the thread that executed the region while the sampler lives.
class HipRegionSampler final : public RegionCountsSampler {
public:
    HipRegionSampler(mh_region_sampler* sampler, std::span<const std::pair<std::uint32_t, std::uint32_t>> targets,
                     std::uint32_t width, std::uint64_t expected_draws)
        : sampler_(sampler), targets_(targets.begin(), targets.end()), width_(width),
          expected_(std::max<std::uint64_t>(expected_draws, 1)) {
        std::uint64_t bytes = 0;
        if (mh_region_sampler_info(sampler_, nullptr, &bytes) == MH_OK)
            charge_.emplace(bytes, common::MemoryLedger::Space::Device);
    }
    ~HipRegionSampler() override { (void)mh_region_sampler_destroy(sampler_); }
    HipRegionSampler(const HipRegionSampler&) = delete;
    HipRegionSampler& operator=(const HipRegionSampler&) = delete;

    [[nodiscard]] BitString sample_one(common::Rng& rng) const override {
        if (next_ == words_.size()) refill(rng);
        const std::uint64_t word = rng.next_word();
        std::uint64_t basis = 0;
        if (word == words_[next_]) {
            basis = basis_[next_++];
        } else {
            words_.clear();
            basis_.clear();
            next_ = 0;
            const double u = common::Rng::word_to_uniform(word);
            if (const int st = mh_region_sampler_draw(sampler_, &u, 1, &basis); st != MH_OK)
                throw_status(st, "HipRegionSampler::sample_one");
        }
        ++drawn_;
        BitString outcome(width_);
        for (const auto& [qubit, clbit] : targets_)
            if ((basis >> qubit) & 1u) outcome.set(clbit, true);
        return outcome;
    }

private:
    static constexpr std::uint64_t kMaxBatch = std::uint64_t{1} << 20;

    void refill(const common::Rng& rng) const {
        const std::uint64_t left = expected_ > drawn_ ? expected_ - drawn_ : expected_;
        const std::uint64_t count = std::min(left, kMaxBatch);
        common::Rng ahead = rng;
        words_.resize(count);
        std::vector<double> u(count);
        for (std::uint64_t i = 0; i < count; ++i) {
            words_[i] = ahead.next_word();
            u[i] = common::Rng::word_to_uniform(words_[i]);
        }
        basis_.resize(count);
        if (const int st = mh_region_sampler_draw(sampler_, u.data(), count, basis_.data()); st != MH_OK)
            throw_status(st, "HipRegionSampler::sample_one");
        next_ = 0;
    }

    mh_region_sampler* sampler_;
    std::vector<std::pair<std::uint32_t, std::uint32_t>> targets_;
    std::uint32_t width_;
    std::uint64_t expected_;
    std::optional<common::MemoryLedger::Charge> charge_;
    mutable std::vector<std::uint64_t> words_, basis_;
    mutable std::size_t next_ = 0;
    mutable std::uint64_t drawn_ = 0;
};

}  // namespace

bool hip_device_available() {
    mh_device_info info{};
    return mh_device_query(&info) == MH_OK;
}

FlatCircuit::FlatCircuit(std::uint32_t num_qubits, std::span<const Operation> operations) {
    for (const Operation& op : operations) {
        std::visit(
            [&](const auto& o) {
                using T = std::decay_t<decltype(o)>;
                if constexpr (std::is_same_v<T, GateOp>) {
                    opcodes_.push_back(static_cast<std::int32_t>(o.opcode));
                    qubits_.insert(qubits_.end(), o.qubits.begin(), o.qubits.end());
                    params_.insert(params_.end(), o.params.begin(), o.params.end());
                } else if constexpr (std::is_same_v<T, UnitaryOp>) {
                    opcodes_.push_back(kOpUnitary);
                    qubits_.insert(qubits_.end(), o.qubits.begin(), o.qubits.end());
                    for (const std::complex<double>& e : o.matrix) {
                        params_.push_back(e.real());
                        params_.push_back(e.imag());
                    }
                } else if constexpr (std::is_same_v<T, MeasureOp>) {
                    opcodes_.push_back(kOpMeasure);
                    qubits_.push_back(o.qubit);
                    qubits_.push_back(o.clbit);
                } else {
                    throw common::UnsupportedFeature("HIP dense: noise channels are not supported");
                }
            },
            op);
        qubit_offsets_.push_back(static_cast<std::uint32_t>(qubits_.size()));
        param_offsets_.push_back(static_cast<std::uint32_t>(params_.size()));
    }
    view_.num_qubits = num_qubits;
    view_.num_ops = static_cast<std::uint32_t>(opcodes_.size());
    view_.opcodes = opcodes_.data();
    view_.qubit_offsets = qubit_offsets_.data();
    view_.qubits = qubits_.data();
    view_.param_offsets = param_offsets_.data();
    view_.params = params_.data();
}

EngineCapabilities HipDenseEngine::capabilities() const {
    return EngineCapabilities{/*counts=*/true, /*statevector=*/true, /*expectation=*/true};
}

CostEstimate HipDenseEngine::estimate(const EngineEstimateRequest& request) const {
    try {
        if (!hip_device_available()) return CostEstimate{false, 0, 0, "no HIP device is available", {}};
        if (const auto reason = unsupported(request.num_qubits, *request.operations, request.want_expectation,
                                            request.observable, request.bitstring_width);
            reason.has_value())
            return CostEstimate{false, 0, 0, *reason, {}};
        const FlatCircuit flat(request.num_qubits, *request.operations);
        const Outputs o = outputs_of(request.want_expectation, request.observable, {});
        mh_execution_request q{};
        q.circuit = &flat.view();
        q.precision = MH_C128;
        q.fusion = request.fuse_gates ? MH_FUSE_CONTAINMENT : MH_FUSE_NONE;
        q.shots = request.want_counts ? request.shots : 0;
        q.want_statevector = request.want_statevector ? 1 : 0;
        q.num_terms = static_cast<std::uint32_t>(o.x.size());
        q.x_masks = o.x.data();
        q.z_masks = o.z.data();
        mh_cost_estimate e{};
        if (const int st = mh_engine_estimate(&q, &e); st != MH_OK) return CostEstimate{false, 0, 0, mh_last_error(), {}};
        if (!e.feasible) return CostEstimate{false, 0, 0, e.infeasible_reason, {}};
        std::uint64_t available = 0;
        if (mh_device_available_bytes(&available) != MH_OK) return CostEstimate{false, 0, 0, mh_last_error(), {}};
        if (e.device_bytes > available)
            return CostEstimate{false, 0, 0,
                                "HIP dense region needs " + std::to_string(e.device_bytes) +
                                    " bytes of device memory but only " + std::to_string(available) +
                                    " bytes are available on this GPU",
                                {}};
        CostEstimate out{true, e.host_bytes, e.work_units, "",
                         DeviceCost{e.device_bytes, e.h2d_bytes, e.d2h_bytes, e.launches}};
        // The host prepares every block (fusion plan, block matrices, pass
        // programs): reported as the CPU dense engine's pass component, so the
        // planner prices it at the CPU's per-pass rate as it does for CUDA.
        out.time.counts[static_cast<std::size_t>(Engine::Dense)][0] = static_cast<double>(e.blocks);
        if (request.fuse_gates) {
            std::vector<std::uint32_t> order(flat.view().num_ops + 1), bb(flat.view().num_ops + 2),
                qb(flat.view().num_ops + 2), bq(flat.view().qubit_offsets[flat.view().num_ops] + 1);
            std::uint32_t blocks = 0;
            std::uint64_t fused = 0, absorbed = 0;
            if (mh_plan_fusion(&flat.view(), MH_FUSE_CONTAINMENT, 0, &blocks, order.data(), bb.data(), bq.data(),
                               qb.data(), &fused, &absorbed) == MH_OK)
                out.fusion = FusionSummary{fused, absorbed};
        }
        return out;
    } catch (const std::exception& e) {
        return CostEstimate{false, 0, 0, e.what(), {}};
    }
}

double HipDenseEngine::device_seconds(const EngineEstimateRequest& request) const {
    try {
        const FlatCircuit flat(request.num_qubits, *request.operations);
        const Outputs o = outputs_of(request.want_expectation, request.observable, {});
        mh_execution_request q{};
        q.circuit = &flat.view();
        q.precision = MH_C128;
        q.fusion = request.fuse_gates ? MH_FUSE_CONTAINMENT : MH_FUSE_NONE;
        q.shots = request.want_counts ? request.shots : 0;
        q.want_statevector = request.want_statevector ? 1 : 0;
        q.num_terms = static_cast<std::uint32_t>(o.x.size());
        q.x_masks = o.x.data();
        q.z_masks = o.z.data();
        mh_cost_estimate e{};
        if (mh_engine_estimate(&q, &e) != MH_OK || !e.feasible) return -1.0;
        return e.seconds;
    } catch (const std::exception&) {
        return -1.0;
    }
}

EngineExecutionResult HipDenseEngine::execute(const EngineExecutionRequest& request) const {
    common::ensure(hip_device_available(),
                   "HipDenseEngine::execute: no HIP device is available -- the planner must not select this "
                   "backend without one");
    if (const auto reason = unsupported(request.num_qubits, request.operations, request.want_expectation,
                                        request.observable, request.bitstring_width);
        reason.has_value())
        throw common::UnsupportedFeature(*reason);
    const FlatCircuit flat(request.num_qubits, request.operations);
    const Outputs o = outputs_of(request.want_expectation, request.observable, request.measure_targets);
    const std::uint64_t amplitudes = std::uint64_t{1} << request.num_qubits;
    const bool sole_counts = request.want_counts && request.is_sole_region;
    // A non-sole region's sampler keeps the final state on the device
    // (HipRegionSampler) instead of a host copy.
    const bool region_sampler = request.want_counts && !request.is_sole_region;
    const bool want_host_state = request.want_statevector;

    mh_execution_request q{};
    q.circuit = &flat.view();
    q.precision = MH_C128;
    q.fusion = request.fuse_gates ? MH_FUSE_CONTAINMENT : MH_FUSE_NONE;
    q.global_phase = request.global_phase;
    q.shots = sole_counts ? request.shots : 0;
    q.seed = request.seed;
    q.target_qubits = o.target_qubits.data();
    q.target_clbits = o.target_clbits.data();
    q.num_targets = static_cast<std::uint32_t>(o.target_qubits.size());
    q.bitstring_width = request.bitstring_width;
    q.num_terms = static_cast<std::uint32_t>(o.x.size());
    q.x_masks = o.x.data();
    q.z_masks = o.z.data();
    q.want_statevector = want_host_state ? 1 : 0;
    q.want_region_sampler = region_sampler ? 1 : 0;
    MonitorBridge bridge{request.monitor, nullptr};
    if (request.monitor != nullptr) {
        q.monitor = &MonitorBridge::call;
        q.monitor_data = &bridge;
    }

    std::vector<std::complex<double>> host;
    std::optional<common::MemoryLedger::Charge> host_charge;
    if (want_host_state) {
        const std::uint64_t host_bytes = amplitudes * sizeof(std::complex<double>);
        if (host_bytes > request.budget_bytes)
            throw common::ResourceExceeded("HIP dense: the host copy needs " + std::to_string(host_bytes) +
                                           " bytes, budget is " + std::to_string(request.budget_bytes) + " bytes");
        host.resize(static_cast<std::size_t>(amplitudes));
        host_charge.emplace(common::buffer_bytes(host));
    }
    std::vector<std::uint64_t> outcomes(q.shots ? q.shots : 1), counts(q.shots ? q.shots : 1);
    std::vector<double> terms(2 * o.x.size() + 2);
    mh_execution_result res{};
    res.statevector = want_host_state ? host.data() : nullptr;
    res.outcomes = outcomes.data();
    res.counts = counts.data();
    res.expectation_terms = terms.data();
    const int st = mh_engine_execute(&q, &res);
    if (st == MH_ERR_ABANDONED && bridge.thrown) std::rethrow_exception(bridge.thrown);
    if (st != MH_OK) throw_status(st, "HipDenseEngine::execute");
    std::unique_ptr<RegionCountsSampler> sampler;
    if (res.region_sampler != nullptr)
        sampler = std::make_unique<HipRegionSampler>(res.region_sampler, request.measure_targets,
                                                     request.bitstring_width, request.shots);
    // The device's peak for this call, on the calling thread's ledger.
    { const common::MemoryLedger::Charge device(res.peak_device_bytes, common::MemoryLedger::Space::Device); }

    EngineExecutionResult result;
    result.actual_bytes = res.actual_bytes;
    if (request.want_expectation) {
        std::vector<std::complex<double>> raw(o.x.size());
        for (std::size_t t = 0; t < raw.size(); ++t) raw[t] = {terms[2 * t], terms[2 * t + 1]};
        result.expectation_terms = std::move(raw);
    }
    if (sole_counts) {
        std::map<BitString, std::uint64_t> hist;
        for (std::uint64_t i = 0; i < res.num_outcomes; ++i)
            hist.emplace(bitstring_of(outcomes[i], request.bitstring_width), counts[i]);
        result.counts = std::move(hist);
    } else if (request.want_counts) {
        result.counts_sampler = std::move(sampler);
    }
    if (request.want_statevector) {
        host_charge->release();
        result.statevector = std::move(host);
    }
    return result;
}

// ---- the host-offloaded tier

namespace {

// The tier's refusals before anything is costed (the disk tier's,
// disk_dense_engine.cpp:148-160, and this ABI's).
std::optional<std::string> tier_unsupported(std::uint32_t num_qubits, std::span<const Operation> operations,
                                            bool want_statevector, bool is_sole_region, bool want_expectation,
                                            const Observable* observable, std::uint32_t bitstring_width) {
    if (auto reason = unsupported(num_qubits, operations, want_expectation, observable, bitstring_width)) return reason;
    if (want_statevector)
        return std::string("the host-offloaded tier does not return a statevector: the state would be held twice");
    if (!is_sole_region) return std::string("the host-offloaded tier runs only as a job's sole region");
    return std::nullopt;
}

std::uint64_t device_budget_or_available(std::uint64_t budget) {
    if (budget != 0) return budget;
    std::uint64_t available = 0;
    if (mh_device_available_bytes(&available) != MH_OK) return 0;
    return available;
}

}  // namespace

EngineCapabilities HipHostTierEngine::capabilities() const {
    return EngineCapabilities{/*counts=*/true, /*statevector=*/false, /*expectation=*/true};
}

CostEstimate HipHostTierEngine::estimate(const EngineEstimateRequest& request) const {
    try {
        if (!hip_device_available()) return CostEstimate{false, 0, 0, "no HIP device is available", {}};
        if (const auto reason = tier_unsupported(request.num_qubits, *request.operations, request.want_statevector,
                                                 request.is_sole_region, request.want_expectation, request.observable,
                                                 request.bitstring_width))
            return CostEstimate{false, 0, 0, *reason, {}};
        const FlatCircuit flat(request.num_qubits, *request.operations);
        const Outputs o = outputs_of(request.want_expectation, request.observable, {});
        mh_execution_request q{};
        q.circuit = &flat.view();
        q.precision = MH_C128;
        q.fusion = request.fuse_gates ? MH_FUSE_CONTAINMENT : MH_FUSE_NONE;
        q.shots = request.want_counts ? request.shots : 0;
        q.num_terms = static_cast<std::uint32_t>(o.x.size());
        q.x_masks = o.x.data();
        q.z_masks = o.z.data();
        q.budget_bytes = device_budget_or_available(device_budget_);
        mh_tier_request t{request.tier_chunk_qubits, request.tier_window_qubits, nullptr, 0, 0};
        mh_tier_cost c{};
        if (mh_tier_estimate(&q, &t, &c) != MH_OK) return CostEstimate{false, 0, 0, mh_last_error(), {}};
        if (!c.feasible) return CostEstimate{false, 0, 0, c.infeasible_reason, {}};
        if (c.host_state_bytes > request.budget_bytes)
            return CostEstimate{false, 0, 0,
                                "the host-offloaded state needs " + std::to_string(c.host_state_bytes) +
                                    " bytes of host memory, budget is " + std::to_string(request.budget_bytes),
                                {}};
        // Work units as the in-memory engine counts them (the same plan).
        q.budget_bytes = 0;
        mh_cost_estimate e{};
        const std::uint64_t work = mh_engine_estimate(&q, &e) == MH_OK ? e.work_units : 0;
        CostEstimate out{true, c.host_state_bytes, work, "",
                         DeviceCost{c.device_bytes, c.link_bytes - c.link_write_bytes, c.link_write_bytes, 0}};
        out.tier = TierCost{c.passes, c.chunk_qubits, c.window_qubits, c.link_bytes, /*file_bytes=*/0,
                            c.link_write_bytes};
        return out;
    } catch (const std::exception& e) {
        return CostEstimate{false, 0, 0, e.what(), {}};
    }
}

EngineExecutionResult HipHostTierEngine::execute(const EngineExecutionRequest& request) const {
    common::ensure(hip_device_available(),
                   "HipHostTierEngine::execute: no HIP device is available -- the planner must not select this "
                   "engine without one");
    if (const auto reason = tier_unsupported(request.num_qubits, request.operations, request.want_statevector,
                                             request.is_sole_region, request.want_expectation, request.observable,
                                             request.bitstring_width))
        throw common::UnsupportedFeature(*reason);
    const FlatCircuit flat(request.num_qubits, request.operations);
    const Outputs o = outputs_of(request.want_expectation, request.observable, request.measure_targets);
    mh_execution_request q{};
    q.circuit = &flat.view();
    q.precision = MH_C128;
    q.fusion = request.fuse_gates ? MH_FUSE_CONTAINMENT : MH_FUSE_NONE;
    q.global_phase = request.global_phase;
    q.shots = request.want_counts ? request.shots : 0;
    q.seed = request.seed;
    q.target_qubits = o.target_qubits.data();
    q.target_clbits = o.target_clbits.data();
    q.num_targets = static_cast<std::uint32_t>(o.target_qubits.size());
    q.bitstring_width = request.bitstring_width;
    q.num_terms = static_cast<std::uint32_t>(o.x.size());
    q.x_masks = o.x.data();
    q.z_masks = o.z.data();
    q.budget_bytes = device_budget_or_available(device_budget_);
    MonitorBridge bridge{request.monitor, nullptr};
    if (request.monitor != nullptr) {
        q.monitor = &MonitorBridge::call;
        q.monitor_data = &bridge;
    }
    const std::uint64_t host_bytes = (std::uint64_t{1} << request.num_qubits) * sizeof(std::complex<double>);
    if (host_bytes > request.budget_bytes)
        throw common::ResourceExceeded("the host-offloaded state needs " + std::to_string(host_bytes) +
                                       " bytes of host memory, budget is " + std::to_string(request.budget_bytes));
    // The pinned state, on the host ledger while the engine holds it.
    const common::MemoryLedger::Charge host(host_bytes);
    mh_tier_request t{request.tier_chunk_qubits, request.tier_window_qubits, nullptr, 0, 0};
    std::vector<std::uint64_t> outcomes(q.shots ? q.shots : 1), counts(q.shots ? q.shots : 1);
    std::vector<double> terms(2 * o.x.size() + 2);
    mh_execution_result res{};
    res.outcomes = outcomes.data();
    res.counts = counts.data();
    res.expectation_terms = terms.data();
    const int st = mh_tier_execute(&q, &t, &res, nullptr);
    if (st == MH_ERR_ABANDONED && bridge.thrown) std::rethrow_exception(bridge.thrown);
    if (st != MH_OK) throw_status(st, "HipHostTierEngine::execute");
    { const common::MemoryLedger::Charge device(res.peak_device_bytes, common::MemoryLedger::Space::Device); }

    EngineExecutionResult result;
    result.actual_bytes = res.actual_bytes;
    if (request.want_expectation) {
        std::vector<std::complex<double>> raw(o.x.size());
        for (std::size_t i = 0; i < raw.size(); ++i) raw[i] = {terms[2 * i], terms[2 * i + 1]};
        result.expectation_terms = std::move(raw);
    }
    if (request.want_counts) {
        std::map<BitString, std::uint64_t> hist;
        for (std::uint64_t i = 0; i < res.num_outcomes; ++i)
            hist.emplace(bitstring_of(outcomes[i], request.bitstring_width), counts[i]);
        result.counts = std::move(hist);
    }
    return result;
}

}  // namespace menos::engines::dense_hip
