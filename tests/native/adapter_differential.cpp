// The reference-side adapter (menos_amd/integration/) against the reference's
// own CPU engine, through MENOS's SimulationEngine interface -- the shape of the
// reference's tests/api/cpu_cuda_differential_test.cpp and
// tests/plan/cuda_backend_selection_test.cpp, without gtest (no network here).
// Built by oracle/Makefile (target `adapter`) with the reference's sources.
//
//   adapter_differential --host   the backend choice and registry rules with
//                                 synthetic inputs; needs no device
//   adapter_differential --device the differential cases on the GPU
//
// Statevectors and counts must be bitwise the CPU engine's; raw expectation
// terms within 1e-12 (the reference's own CUDA bar, common/tolerance.h).
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "common/error.h"
#include "common/memory_ledger.h"
#include "common/rng.h"
#include "engines/dense/dense_engine.h"
#include "hip_dense_engine.h"
#include "hip_registry.h"
#include "menos/circuit.h"
#include "menos/observable.h"
#include "menos/simulation_engine.h"

using namespace menos;
using engines::dense::DenseEngine;
using engines::dense_hip::HipDenseEngine;
using engines::dense_hip::kBackendHip;
using C = std::complex<double>;

namespace {

int g_checks = 0;
int g_failures = 0;

void check(bool ok, const std::string& what) {
    ++g_checks;
    if (!ok) {
        ++g_failures;
        std::fprintf(stderr, "FAIL: %s\n", what.c_str());
    }
}

template <typename E, typename F>
bool throws(F&& f) {
    try {
        f();
    } catch (const E&) {
        return true;
    } catch (...) {
        return false;
    }
    return false;
}

GateOp gate(Opcode op, std::vector<QubitId> q, std::vector<double> p = {}) { return GateOp{op, std::move(q), std::move(p)}; }

PauliString pauli(std::uint32_t n, std::uint64_t x, std::uint64_t z) {
    PauliString p;
    p.num_qubits = n;
    p.x_words = {x};
    p.z_words = {z};
    return p;
}

// A random unitary of k qubits (QR of a Gaussian matrix, Gram-Schmidt).
std::vector<C> random_unitary(std::uint32_t k, std::mt19937_64& g) {
    const std::size_t d = std::size_t{1} << k;
    std::normal_distribution<double> nd;
    std::vector<C> m(d * d);
    for (auto& e : m) e = {nd(g), nd(g)};
    for (std::size_t c = 0; c < d; ++c) {
        for (std::size_t p = 0; p < c; ++p) {
            C dot{0, 0};
            for (std::size_t r = 0; r < d; ++r) dot += std::conj(m[r * d + p]) * m[r * d + c];
            for (std::size_t r = 0; r < d; ++r) m[r * d + c] -= dot * m[r * d + p];
        }
        double nrm = 0;
        for (std::size_t r = 0; r < d; ++r) nrm += std::norm(m[r * d + c]);
        nrm = std::sqrt(nrm);
        for (std::size_t r = 0; r < d; ++r) m[r * d + c] /= nrm;
    }
    return m;
}

// A mixed circuit: U3/RY/RZ/H/T on random qubits, CX/CZ/CP/SWAP pairs, CCX,
// and explicit unitaries of 1..4 qubits.
std::vector<Operation> mixed_circuit(std::uint32_t n, std::uint32_t ops, std::uint64_t seed) {
    std::mt19937_64 g(seed);
    std::uniform_real_distribution<double> ang(0.0, 6.283185307179586);
    std::vector<Operation> out;
    auto pick = [&](std::uint32_t k) {
        std::vector<QubitId> q;
        while (q.size() < k) {
            const QubitId c = static_cast<QubitId>(g() % n);
            if (std::find(q.begin(), q.end(), c) == q.end()) q.push_back(c);
        }
        return q;
    };
    for (std::uint32_t i = 0; i < ops; ++i) {
        switch (g() % 9) {
            case 0: out.push_back(gate(Opcode::U, pick(1), {ang(g), ang(g), ang(g)})); break;
            case 1: out.push_back(gate(Opcode::RY, pick(1), {ang(g)})); break;
            case 2: out.push_back(gate(Opcode::RZ, pick(1), {ang(g)})); break;
            case 3: out.push_back(gate(g() % 2 ? Opcode::H : Opcode::T, pick(1))); break;
            case 4: out.push_back(gate(Opcode::CX, pick(2))); break;
            case 5: out.push_back(gate(g() % 2 ? Opcode::CZ : Opcode::SWAP, pick(2))); break;
            case 6: out.push_back(gate(Opcode::CP, pick(2), {ang(g)})); break;
            case 7: out.push_back(gate(Opcode::CCX, pick(3))); break;
            default: {
                const std::uint32_t k = 1 + static_cast<std::uint32_t>(g() % 4);
                out.push_back(UnitaryOp{pick(k), random_unitary(k, g)});
            }
        }
    }
    return out;
}

EngineExecutionRequest request(std::uint32_t n, const std::vector<Operation>& ops, bool sv, bool counts,
                               const Observable* obs, std::uint64_t shots = 0,
                               std::span<const std::pair<QubitId, ClbitId>> targets = {}, std::uint32_t width = 0,
                               double phase = 0.0, bool fuse = true, bool sole = true) {
    EngineExecutionRequest r{
        .num_qubits = n,
        .operations = ops,
        .want_statevector = sv,
        .want_counts = counts,
        .want_expectation = obs != nullptr,
        .shots = shots,
        .measure_targets = targets,
        .bitstring_width = width,
        .seed = 2026,
        .global_phase = phase,
        .budget_bytes = std::uint64_t{1} << 40,
        .is_sole_region = sole,
        .observable = obs,
    };
    r.fuse_gates = fuse;
    return r;
}

bool same_bits(const std::vector<C>& a, const std::vector<C>& b) {
    return a.size() == b.size() && std::memcmp(a.data(), b.data(), a.size() * sizeof(C)) == 0;
}

// ---- host-only: the backend choice and the registry

engines::dense_hip::DenseBackendInputs base_inputs() {
    engines::dense_hip::DenseBackendInputs in;
    in.hip_engine_registered = true;
    in.hip.present = true;
    in.hip.calibrated = true;
    in.hip.device_available_bytes = std::uint64_t{200} << 30;
    in.cpu_calibrated = true;
    in.cpu_seconds = 1.0;
    in.hip_estimate = CostEstimate{true, 0, 100, "", DeviceCost{std::uint64_t{16} << 30, 0, 0, 10}};
    in.hip_seconds = 0.1;
    return in;
}

void host_cases() {
    using engines::dense_hip::BackendRequest;
    using engines::dense_hip::choose_dense_backend;
    {
        auto in = base_inputs();
        check(choose_dense_backend(in).backend == kBackendHip, "a faster device takes the dense region");
    }
    {
        auto in = base_inputs();
        in.hip_seconds = 1.0;
        check(choose_dense_backend(in).backend == Backend::Cpu, "ties go to the CPU");
        in.hip_seconds = 1.5;
        check(choose_dense_backend(in).backend == Backend::Cpu, "a slower device loses");
    }
    {
        auto in = base_inputs();
        in.hip.device_available_bytes = 1024;
        check(choose_dense_backend(in).backend == Backend::Cpu, "a device too small keeps the region on the CPU");
    }
    {
        auto in = base_inputs();
        in.hip.calibrated = false;
        check(choose_dense_backend(in).backend == Backend::Cpu, "an uncalibrated device is never chosen");
        in = base_inputs();
        in.cpu_calibrated = false;
        check(choose_dense_backend(in).backend == Backend::Cpu, "an uncalibrated CPU is never compared");
    }
    {
        auto in = base_inputs();
        in.requested = BackendRequest::Cpu;
        check(choose_dense_backend(in).backend == Backend::Cpu, "backend=cpu never moves to the device");
    }
    {
        auto in = base_inputs();
        in.cpu_tier = Tier::Disk;
        check(choose_dense_backend(in).backend == Backend::Cpu, "a disk-tier candidate is not compared");
    }
    {
        auto in = base_inputs();
        in.observed_hip_seconds = 5.0;
        check(choose_dense_backend(in).backend == Backend::Cpu, "an observed device overrun steers to the CPU");
        in = base_inputs();
        in.hip_seconds = 2.0;
        in.observed_cpu_seconds = 3.0;
        check(choose_dense_backend(in).backend == kBackendHip, "an observed CPU overrun steers to the device");
    }
    {
        auto in = base_inputs();
        in.requested = BackendRequest::Hip;
        in.hip_seconds = 9.0;
        const auto c = choose_dense_backend(in);
        check(c.backend == kBackendHip && c.feasible, "backend=hip is honoured even when slower");
        in.hip_estimate.device.device_bytes = std::uint64_t{300} << 30;
        const auto d = choose_dense_backend(in);
        check(!d.feasible && d.infeasible_reason.find("bytes of device memory") != std::string::npos,
              "backend=hip with a state beyond the device is infeasible, naming the bytes");
        in = base_inputs();
        in.requested = BackendRequest::Hip;
        in.hip_estimate = CostEstimate{false, 0, 0, "too wide", {}};
        const auto e = choose_dense_backend(in);
        check(!e.feasible && e.infeasible_reason == "backend=hip requested: too wide",
              "backend=hip with an infeasible estimate carries its reason");
    }
    {
        auto in = base_inputs();
        in.requested = BackendRequest::Hip;
        in.hip_engine_registered = false;
        check(throws<common::UnsupportedFeature>([&] { (void)choose_dense_backend(in); }),
              "backend=hip without HIP support is refused");
        in = base_inputs();
        in.requested = BackendRequest::Hip;
        in.hip.present = false;
        check(throws<common::UnsupportedFeature>([&] { (void)choose_dense_backend(in); }),
              "backend=hip without a device is refused");
        in = base_inputs();
        in.requested = BackendRequest::Hip;
        in.noise_reason = "depolarizing channel";
        check(throws<common::UnsupportedFeature>([&] { (void)choose_dense_backend(in); }),
              "backend=hip on a noisy circuit is refused");
    }
    {
        // The flat layout of each operation kind.
        std::vector<Operation> ops = {gate(Opcode::U, {2}, {0.1, 0.2, 0.3}), gate(Opcode::CX, {0, 1}),
                                      UnitaryOp{{1}, {C{0, 1}, C{0, 0}, C{0, 0}, C{1, 0}}}, MeasureOp{1, 0}};
        const engines::dense_hip::FlatCircuit flat(3, ops);
        const mh_circuit& c = flat.view();
        check(c.num_ops == 4 && c.opcodes[0] == static_cast<int>(Opcode::U) && c.opcodes[2] == 21 &&
                  c.opcodes[3] == 22,
              "flat opcodes");
        check(c.qubit_offsets[4] == 6 && c.qubits[5] == 0 && c.param_offsets[3] == 11 && c.params[4] == 1.0,
              "flat qubits and params");
    }
    {
        // Registry: without a device the HIP engine is not registered and
        // estimate() reports it, never throwing.
        const DenseEngine cpu;
        const auto engines = engines::dense_hip::with_hip_engine({&cpu});
        check(engines::dense_hip::find_engine(engines, Engine::Dense, Backend::Cpu) == &cpu, "CPU dense resolves");
        const bool device = engines::dense_hip::hip_device_available();
        const SimulationEngine* hip = engines::dense_hip::find_engine(engines, Engine::Dense, kBackendHip);
        check((hip != nullptr) == device, "the HIP entry is present exactly when a device is");
        const HipDenseEngine standalone;
        const std::vector<Operation> ops = {gate(Opcode::H, {0})};
        EngineEstimateRequest er{.num_qubits = 1, .operations = &ops, .want_statevector = true,
                                 .want_counts = false, .want_expectation = false, .observable = nullptr};
        const CostEstimate e = standalone.estimate(er);
        check(e.feasible == device, "estimate feasibility follows the device");
        if (!device) check(e.infeasible_reason == "no HIP device is available", "no-device reason");
    }
}

// ---- device: the differential cases

void tier_cases(const SimulationEngine& cpu);

void device_cases() {
    const DenseEngine cpu_engine;
    const HipDenseEngine hip_engine;
    const SimulationEngine& cpu = cpu_engine;
    const SimulationEngine& hip = hip_engine;
    check(hip.backend() == kBackendHip && hip.engine_id() == Engine::Dense && hip.name() == "dense_hip",
          "engine identity");

    // Single-qubit gates on every qubit after an H layer (the reference's first case).
    for (const Opcode op : {Opcode::H, Opcode::X, Opcode::Y, Opcode::Z, Opcode::S, Opcode::T}) {
        constexpr std::uint32_t n = 11;
        std::vector<Operation> ops;
        for (std::uint32_t q = 0; q < n; ++q) ops.push_back(gate(Opcode::H, {q}));
        for (std::uint32_t q = 0; q < n; ++q) ops.push_back(gate(op, {q}));
        const auto r = request(n, ops, true, false, nullptr);
        check(same_bits(*cpu.execute(r).statevector, *hip.execute(r).statevector),
              "single-qubit gate " + std::to_string(static_cast<int>(op)) + " on every qubit");
    }

    // Mixed circuits, fused and unfused, with a global phase.
    for (std::uint64_t seed = 0; seed < 6; ++seed) {
        const std::uint32_t n = 10 + static_cast<std::uint32_t>(seed) * 2;
        const std::vector<Operation> ops = mixed_circuit(n, 120, 900 + seed);
        for (const bool fuse : {true, false}) {
            const auto r = request(n, ops, true, false, nullptr, 0, {}, 0, 0.3 * static_cast<double>(seed), fuse);
            check(same_bits(*cpu.execute(r).statevector, *hip.execute(r).statevector),
                  "mixed circuit n=" + std::to_string(n) + (fuse ? " fused" : " unfused"));
        }
    }

    // Counts of a sole region, through measure targets into a wider bit string.
    {
        const std::uint32_t n = 16;
        const std::vector<Operation> ops = mixed_circuit(n, 80, 77);
        std::vector<std::pair<QubitId, ClbitId>> targets = {{0, 3}, {5, 0}, {9, 7}, {15, 1}, {3, 9}};
        const auto r = request(n, ops, false, true, nullptr, 4096, targets, 12);
        const auto a = cpu.execute(r);
        const auto b = hip.execute(r);
        check(a.counts.has_value() && b.counts.has_value() && *a.counts == *b.counts, "sole-region counts");
    }

    // A non-sole region's sampler: the same draws from the same Rng stream.
    {
        const std::uint32_t n = 14;
        const std::vector<Operation> ops = mixed_circuit(n, 60, 78);
        std::vector<std::pair<QubitId, ClbitId>> targets;
        for (std::uint32_t q = 0; q < n; ++q) targets.push_back({q, q});
        const auto r = request(n, ops, false, true, nullptr, 1000, targets, n, 0.0, true, /*sole=*/false);
        const auto a = cpu.execute(r);
        const auto b = hip.execute(r);
        bool same = a.counts_sampler && b.counts_sampler;
        common::Rng ra(41), rb(41);
        for (int i = 0; same && i < 2000; ++i) same = a.counts_sampler->sample_one(ra) == b.counts_sampler->sample_one(rb);
        check(same, "non-sole region sampler draws");
        // A caller that also draws from the stream between calls: the HIP
        // sampler's look-ahead no longer matches its words, and every draw
        // stays the reference's.
        for (int i = 0; same && i < 64; ++i) {
            (void)ra.next_word();
            (void)rb.next_word();
            same = a.counts_sampler->sample_one(ra) == b.counts_sampler->sample_one(rb);
        }
        check(same, "non-sole region sampler draws with other draws in between");
    }

    // Raw expectation terms within 1e-12, in term order.
    {
        const std::uint32_t n = 18;
        const std::vector<Operation> ops = mixed_circuit(n, 100, 79);
        Observable obs;
        std::mt19937_64 g(80);
        for (int t = 0; t < 24; ++t) obs.terms.push_back({pauli(n, g() & 0x3ffff, g() & 0x3ffff), C{0.5, -0.25}});
        obs.terms.push_back({pauli(n, 0, 0), C{1, 0}});
        const auto r = request(n, ops, false, false, &obs);
        const auto a = cpu.execute(r);
        const auto b = hip.execute(r);
        double worst = 0;
        for (std::size_t t = 0; t < obs.terms.size(); ++t) worst = std::max(worst, std::abs((*a.expectation_terms)[t] - (*b.expectation_terms)[t]));
        check(b.expectation_terms->size() == obs.terms.size() && worst <= 1e-12,
              "raw expectation terms within 1e-12 (worst " + std::to_string(worst) + ")");
    }

    // The monitor: boundaries counted from 1 and results unchanged; a throwing
    // monitor's own exception reaches the caller.
    {
        struct Counting : ExecutionMonitor {
            mutable std::vector<std::uint64_t> seen;
            void at_safe_boundary(std::uint64_t crossed) const override { seen.push_back(crossed); }
        } counting;
        struct Abandon : ExecutionMonitor {
            void at_safe_boundary(std::uint64_t crossed) const override {
                if (crossed == 2) throw std::range_error("abandoned at boundary 2");
            }
        } abandon;
        const std::uint32_t n = 20;
        const std::vector<Operation> ops = mixed_circuit(n, 200, 81);
        auto r = request(n, ops, true, false, nullptr);
        const auto plain = hip.execute(r);
        r.monitor = &counting;
        const auto watched = hip.execute(r);
        bool ascending = !counting.seen.empty() && counting.seen.front() == 1;
        for (std::size_t i = 1; i < counting.seen.size(); ++i) ascending = ascending && counting.seen[i] == i + 1;
        check(ascending && same_bits(*plain.statevector, *watched.statevector), "monitor boundaries, same state");
        r.monitor = &abandon;
        check(throws<std::range_error>([&] { (void)hip.execute(r); }), "a throwing monitor abandons the run");
    }

    // Budgets and the ledger: a host copy over the budget is refused before
    // it is made; the device peak is charged to the installed ledger.
    {
        const std::uint32_t n = 20;
        const std::vector<Operation> ops = mixed_circuit(n, 30, 82);
        auto r = request(n, ops, true, false, nullptr);
        r.budget_bytes = 1024;
        check(throws<common::ResourceExceeded>([&] { (void)hip.execute(r); }), "host copy over budget");
        r.budget_bytes = std::uint64_t{1} << 40;
        common::MemoryLedger ledger;
        {
            const common::MemoryLedger::Scope scope(ledger);
            (void)hip.execute(r);
        }
        check(ledger.peak_device_bytes() >= (std::uint64_t{16} << n), "device peak on the ledger");
    }

    // estimate(): feasible with device figures; infeasible with a reason for
    // noise; never throws.
    {
        const std::uint32_t n = 22;
        const std::vector<Operation> ops = mixed_circuit(n, 50, 83);
        EngineEstimateRequest er{.num_qubits = n, .operations = &ops, .want_statevector = true,
                                 .want_counts = false, .want_expectation = false, .observable = nullptr};
        const CostEstimate e = hip.estimate(er);
        check(e.feasible && e.device.device_bytes >= (std::uint64_t{16} << n) && e.device.kernel_launches > 0,
              "estimate: device figures");
        check(hip_engine.device_seconds(er) > 0.0, "estimate: calibrated device seconds");
        std::vector<Operation> noisy = ops;
        noisy.push_back(ChannelOp{{0}, {{C{1, 0}, C{0, 0}, C{0, 0}, C{1, 0}}}});
        EngineEstimateRequest nr = er;
        nr.operations = &noisy;
        bool threw = false;
        CostEstimate ne{};
        try {
            ne = hip.estimate(nr);
        } catch (...) {
            threw = true;
        }
        check(!threw && !ne.feasible && !ne.infeasible_reason.empty(), "estimate: noise is infeasible, no throw");
    }

    // The registry resolves (Dense, Hip, InMemory) to the adapter.
    {
        const auto engines = engines::dense_hip::with_hip_engine({&cpu_engine});
        const SimulationEngine* found = engines::dense_hip::find_engine(engines, Engine::Dense, kBackendHip);
        check(found != nullptr && found->name() == "dense_hip", "registry resolves the HIP engine");
    }

    tier_cases(cpu);
}

// The host-offloaded tier through the same interface: a device budget that
// holds two 2^12-amplitude windows and the outputs' tables (and, once,
// explicit seams) makes 16-18-qubit states run chunk by chunk; counts equal
// the CPU engine's and raw terms are within 1e-12.
void tier_cases(const SimulationEngine& cpu) {
    const engines::dense_hip::HipHostTierEngine tier_engine(880000);
    const SimulationEngine& tier = tier_engine;
    check(tier.tier() == Tier::HostOffloaded && tier.backend() == kBackendHip && !tier.capabilities().statevector,
          "tier engine identity");
    for (std::uint64_t seed = 0; seed < 3; ++seed) {
        const std::uint32_t n = 16 + static_cast<std::uint32_t>(seed);
        const std::vector<Operation> ops = mixed_circuit(n, 90, 990 + seed);
        Observable obs;
        std::mt19937_64 g(991 + seed);
        const std::uint64_t all = (std::uint64_t{1} << n) - 1;
        for (int t = 0; t < 12; ++t) obs.terms.push_back({pauli(n, g() & all, g() & all), C{1, 0}});
        std::vector<std::pair<QubitId, ClbitId>> targets;
        for (std::uint32_t q = 0; q < n; ++q) targets.push_back({q, q});
        for (const bool fuse : {true, false}) {
            auto r = request(n, ops, false, true, &obs, 5000, targets, n, 0.2, fuse);
            if (seed == 2) {
                r.tier_chunk_qubits = 6;
                r.tier_window_qubits = 12;
            }
            EngineEstimateRequest er{.num_qubits = n, .operations = &ops, .want_statevector = false,
                                     .want_counts = true, .want_expectation = true, .observable = &obs};
            er.budget_bytes = r.budget_bytes;
            er.shots = 5000;
            er.bitstring_width = n;
            er.is_sole_region = true;
            er.fuse_gates = fuse;
            er.tier_chunk_qubits = r.tier_chunk_qubits;
            er.tier_window_qubits = r.tier_window_qubits;
            const CostEstimate e = tier.estimate(er);
            const std::string name = "tier n=" + std::to_string(n) + (fuse ? " fused" : " unfused");
            check(e.feasible && e.tier.passes > 1 && e.tier.window_qubits == 12u &&
                      e.tier.link_bytes > e.tier.link_write_bytes,
                  name + ": estimate (" + e.infeasible_reason + ")");
            const auto a = cpu.execute(r);
            const auto b = tier.execute(r);
            check(a.counts && b.counts && *a.counts == *b.counts, name + ": counts");
            double worst = 0;
            for (std::size_t t = 0; t < obs.terms.size(); ++t)
                worst = std::max(worst, std::abs((*a.expectation_terms)[t] - (*b.expectation_terms)[t]));
            check(worst <= 1e-12, name + ": raw terms within 1e-12 (worst " + std::to_string(worst) + ")");
        }
    }
    // Refusals: a statevector, a non-sole region, a host state over budget.
    {
        const std::uint32_t n = 16;
        const std::vector<Operation> ops = mixed_circuit(n, 40, 995);
        EngineEstimateRequest er{.num_qubits = n, .operations = &ops, .want_statevector = true,
                                 .want_counts = false, .want_expectation = false, .observable = nullptr};
        er.budget_bytes = std::uint64_t{1} << 40;
        er.is_sole_region = true;
        check(!tier.estimate(er).feasible, "tier: a statevector is infeasible");
        er.want_statevector = false;
        er.is_sole_region = false;
        check(!tier.estimate(er).feasible, "tier: a non-sole region is infeasible");
        er.is_sole_region = true;
        er.budget_bytes = (std::uint64_t{16} << n) - 1;  // one byte short of the state
        check(!tier.estimate(er).feasible, "tier: a host state over budget is infeasible");
        auto r = request(n, ops, false, false, nullptr);
        r.budget_bytes = (std::uint64_t{16} << n) - 1;
        check(throws<common::ResourceExceeded>([&] { (void)tier.execute(r); }), "tier: execute over budget");
    }
    // The monitor at the tier's pass boundaries, and its exception.
    {
        struct Counting : ExecutionMonitor {
            mutable std::vector<std::uint64_t> seen;
            void at_safe_boundary(std::uint64_t crossed) const override { seen.push_back(crossed); }
        } counting;
        struct Abandon : ExecutionMonitor {
            void at_safe_boundary(std::uint64_t crossed) const override {
                if (crossed == 2) throw std::range_error("abandoned at tier pass 2");
            }
        } abandon;
        const std::uint32_t n = 16;
        const std::vector<Operation> ops = mixed_circuit(n, 80, 996);
        Observable obs;
        obs.terms.push_back({pauli(n, 3, 5), C{1, 0}});
        auto r = request(n, ops, false, false, &obs);
        r.monitor = &counting;
        const auto watched = tier.execute(r);
        bool ascending = counting.seen.size() > 2 && counting.seen.front() == 1;
        for (std::size_t i = 1; i < counting.seen.size(); ++i) ascending = ascending && counting.seen[i] == i + 1;
        const auto a = cpu.execute(request(n, ops, false, false, &obs));
        check(ascending && std::abs((*a.expectation_terms)[0] - (*watched.expectation_terms)[0]) <= 1e-12,
              "tier: monitor boundaries, same terms");
        r.monitor = &abandon;
        check(throws<std::range_error>([&] { (void)tier.execute(r); }), "tier: a throwing monitor abandons the run");
    }
}

}  // namespace

int main(int argc, char** argv) {
    const bool device = argc > 1 && std::strcmp(argv[1], "--device") == 0;
    host_cases();
    if (device) device_cases();
    std::printf("%s: %d checks, %d failed\n", device ? "device" : "host", g_checks, g_failures);
    return g_failures == 0 ? 0 : 1;
}
