// A C entry layer over the reference's own dense path, compiled together with
// the reference's sources (oracle/Makefile builds both into oracle/_ref/).
// TEST INFRASTRUCTURE ONLY: it exists to pin the oracle's restatement and to
// time the reference CPU engine as the CPU baseline. Nothing in menos_amd/
// links it.
//
// Every function takes the oracle's flat layouts (oracle.h) and calls the
// reference's real code: DenseEngine::execute, ir::plan_fusion,
// engines::dense::fused_block_matrix, kernels::cpu::apply_matrix(_generic),
// ir::generate_matrix, engines::dense::sample, CumulativeDistribution.
#include <chrono>
#include <complex>
#include <cstdint>
#include <cstring>
#include <map>
#include <optional>
#include <vector>

#include "common/memory_ledger.h"
#include "common/rng.h"
#include "engines/dense/cumulative_distribution.h"
#include "engines/dense/dense_engine.h"
#include "engines/dense/fused_block.h"
#include "engines/dense/sample.h"
#include "engines/dense_tiered/pass_plan.h"
#include "ir/fuse_gates.h"
#include "ir/gate_table.h"
#include "kernels/dense_kernels.h"
#include "menos/circuit.h"
#include "menos/observable.h"
#include "oracle.h"

namespace {

using C = std::complex<double>;

std::vector<menos::Operation> operations_of(const orc_circuit& c) {
    std::vector<menos::Operation> ops;
    ops.reserve(c.num_ops);
    for (uint32_t i = 0; i < c.num_ops; ++i) {
        const auto op = static_cast<menos::Opcode>(c.opcodes[i]);
        std::vector<menos::QubitId> q(c.qubits + c.qubit_offsets[i], c.qubits + c.qubit_offsets[i + 1]);
        std::vector<double> p(c.params + c.param_offsets[i], c.params + c.param_offsets[i + 1]);
        if (op == menos::Opcode::Measure) {
            ops.emplace_back(menos::MeasureOp{q[0], q[1]});
        } else if (op == menos::Opcode::Unitary) {
            std::vector<C> m(p.size() / 2);
            for (std::size_t e = 0; e < m.size(); ++e) m[e] = C{p[2 * e], p[2 * e + 1]};
            ops.emplace_back(menos::UnitaryOp{std::move(q), std::move(m)});
        } else {
            ops.emplace_back(menos::GateOp{op, std::move(q), std::move(p)});
        }
    }
    return ops;
}

menos::ir::FusionPlan plan_of(const orc_circuit& c, const orc_plan& p) {
    menos::ir::FusionPlan plan;
    plan.order.assign(p.order, p.order + c.num_ops);
    plan.block_begin.assign(p.block_begin, p.block_begin + p.num_blocks + 1);
    plan.qubits_begin.assign(p.qubits_begin, p.qubits_begin + p.num_blocks + 1);
    plan.block_qubits.assign(p.block_qubits, p.block_qubits + p.qubits_begin[p.num_blocks]);
    return plan;
}

}  // namespace

extern "C" {

int ref_gate_matrix(int32_t opcode, const double* params, uint32_t num_params, double* out) {
    const std::vector<C> m =
        menos::ir::generate_matrix(static_cast<menos::Opcode>(opcode), std::span<const double>(params, num_params));
    std::memcpy(out, m.data(), m.size() * sizeof(C));
    return static_cast<int>(m.size());
}

void ref_apply_matrix_generic(double* state, uint32_t num_qubits, const double* matrix, const uint32_t* operands,
                              uint32_t k, uint32_t threads) {
    const std::size_t size = std::size_t{1} << num_qubits;
    const std::size_t dim = std::size_t{1} << k;
    menos::kernels::cpu::apply_matrix_generic(std::span<C>(reinterpret_cast<C*>(state), size),
                                              std::span<const C>(reinterpret_cast<const C*>(matrix), dim * dim),
                                              std::span<const uint32_t>(operands, k), threads);
}

void ref_apply_matrix(double* state, uint32_t num_qubits, const double* matrix, const uint32_t* operands,
                      uint32_t k, uint32_t threads) {
    const std::size_t size = std::size_t{1} << num_qubits;
    const std::size_t dim = std::size_t{1} << k;
    menos::kernels::cpu::apply_matrix(std::span<C>(reinterpret_cast<C*>(state), size),
                                      std::span<const C>(reinterpret_cast<const C*>(matrix), dim * dim),
                                      std::span<const uint32_t>(operands, k), threads);
}

uint32_t ref_plan_fusion(const orc_circuit* circuit, uint32_t* order, uint32_t* block_begin, uint32_t* block_qubits,
                         uint32_t* qubits_begin, uint64_t* fused_blocks, uint64_t* gates_absorbed) {
    const std::vector<menos::Operation> ops = operations_of(*circuit);
    const menos::ir::FusionPlan p = menos::ir::plan_fusion(ops);
    std::copy(p.order.begin(), p.order.end(), order);
    std::copy(p.block_begin.begin(), p.block_begin.end(), block_begin);
    std::copy(p.block_qubits.begin(), p.block_qubits.end(), block_qubits);
    std::copy(p.qubits_begin.begin(), p.qubits_begin.end(), qubits_begin);
    *fused_blocks = p.fused_blocks;
    *gates_absorbed = p.gates_absorbed;
    return static_cast<uint32_t>(p.block_count());
}

// The disk tier's pass plan (engines/dense_tiered/pass_plan.h): over the
// operations, or with fuse over ir::plan_fusion's blocks; chunk_qubits 0 with
// choose set asks choose_pass_plan for the chunk. Writes up to `capacity`
// passes; returns the pass count, or UINT32_MAX when no plan fits.
uint32_t ref_tier_pass_plan(const orc_circuit* circuit, uint32_t chunk_qubits, uint32_t window_qubits, int32_t fuse,
                            int32_t choose, uint64_t* masks, uint32_t* first, uint32_t* end, uint32_t capacity,
                            uint32_t* chunk_out, uint64_t* global_gates, uint32_t* max_global, int32_t* too_small) {
    const std::vector<menos::Operation> ops = operations_of(*circuit);
    std::optional<menos::ir::FusionPlan> fusion;
    if (fuse) fusion.emplace(menos::ir::plan_fusion(ops));
    const menos::ir::FusionPlan* f = fusion ? &*fusion : nullptr;
    menos::engines::dense_tiered::PassPlan plan;
    if (choose) {
        const auto chosen =
            menos::engines::dense_tiered::choose_pass_plan(ops, circuit->num_qubits, window_qubits, false, f);
        if (!chosen) return UINT32_MAX;
        plan = *chosen;
    } else {
        plan = menos::engines::dense_tiered::build_pass_plan(ops, circuit->num_qubits, chunk_qubits, window_qubits, f);
    }
    *chunk_out = plan.chunk_qubits;
    *global_gates = plan.global_gates;
    *max_global = plan.max_global_operands;
    *too_small = plan.window_too_small ? 1 : 0;
    const std::size_t m = std::min<std::size_t>(capacity, plan.passes.size());
    for (std::size_t i = 0; i < m; ++i) {
        masks[i] = plan.passes[i].global_mask;
        first[i] = static_cast<uint32_t>(plan.passes[i].first_operation);
        end[i] = static_cast<uint32_t>(plan.passes[i].end_operation);
    }
    return static_cast<uint32_t>(plan.passes.size());
}

void ref_fused_block_matrix(const orc_circuit* circuit, const orc_plan* plan, uint32_t block, double* out) {
    const std::vector<menos::Operation> ops = operations_of(*circuit);
    const std::vector<C> m = menos::engines::dense::fused_block_matrix(plan_of(*circuit, *plan), block, ops);
    std::memcpy(out, m.data(), m.size() * sizeof(C));
}

void ref_expectation_of_term(const double* state, uint32_t num_qubits, uint64_t x_mask, uint64_t z_mask,
                             double* out2) {
    const C e = menos::kernels::cpu::expectation_of_term(
        std::span<const C>(reinterpret_cast<const C*>(state), std::size_t{1} << num_qubits), x_mask, z_mask);
    out2[0] = e.real();
    out2[1] = e.imag();
}

// Initialise, then sweep an explicit block list with the reference's own
// apply_blocks and kernels::cpu::apply_matrix (the block lists of wider
// fusion plans than the reference's containment rule). Returns wall seconds.
double ref_run_plan(const orc_circuit* circuit, const orc_plan* plan, double global_phase, uint32_t threads,
                    double* state_out) {
    const std::vector<menos::Operation> ops = operations_of(*circuit);
    const menos::ir::FusionPlan p = plan_of(*circuit, *plan);
    const std::size_t size = std::size_t{1} << circuit->num_qubits;
    std::span<C> state(reinterpret_cast<C*>(state_out), size);
    const auto t0 = std::chrono::steady_clock::now();
    menos::kernels::cpu::initialize(state, global_phase);
    menos::engines::dense::apply_blocks(p, ops, nullptr,
                                        [&](std::span<const C> matrix, std::span<const menos::QubitId> operands) {
                                            menos::kernels::cpu::apply_matrix(state, matrix, operands, threads);
                                        });
    const auto t1 = std::chrono::steady_clock::now();
    return std::chrono::duration<double>(t1 - t0).count();
}

// The sole-region sampler over the given state, outcome keys as 64-bit words
// (bitstring_width <= 64). Returns the number of distinct outcomes written.
uint64_t ref_sample_counts(const double* state, uint32_t num_qubits, uint64_t shots, uint64_t seed,
                           const uint32_t* target_qubits, const uint32_t* target_clbits, uint32_t num_targets,
                           uint32_t bitstring_width, uint64_t* outcomes, uint64_t* counts) {
    std::vector<std::pair<uint32_t, uint32_t>> targets;
    for (uint32_t t = 0; t < num_targets; ++t) targets.emplace_back(target_qubits[t], target_clbits[t]);
    const auto hist = menos::engines::dense::sample(
        std::span<const C>(reinterpret_cast<const C*>(state), std::size_t{1} << num_qubits), shots, seed, targets,
        bitstring_width);
    uint64_t at = 0;
    for (const auto& [key, count] : hist) {
        uint64_t word = 0;
        for (uint32_t b = 0; b < bitstring_width; ++b)
            if (key.test(b)) word |= uint64_t{1} << b;
        outcomes[at] = word;
        counts[at] = count;
        ++at;
    }
    return at;
}

// A non-sole region's sampler: `draws` outcomes from Rng(seed).
void ref_region_draws(const double* state, uint32_t num_qubits, uint64_t draws, uint64_t seed,
                      uint64_t* basis_of_draw) {
    // Identity targets over every qubit make the drawn key the basis index.
    std::vector<std::pair<uint32_t, uint32_t>> targets;
    for (uint32_t q = 0; q < num_qubits; ++q) targets.emplace_back(q, q);
    const menos::engines::dense::CumulativeDistribution dist(
        std::span<const C>(reinterpret_cast<const C*>(state), std::size_t{1} << num_qubits), targets, num_qubits);
    menos::common::Rng rng(seed);
    for (uint64_t d = 0; d < draws; ++d) {
        const menos::BitString key = dist.sample_one(rng);
        uint64_t word = 0;
        for (uint32_t b = 0; b < num_qubits; ++b)
            if (key.test(b)) word |= uint64_t{1} << b;
        basis_of_draw[d] = word;
    }
}

// DenseEngine::execute on the request. Counts are written as (outcome word,
// count) pairs (bitstring_width <= 64); returns the wall seconds execute took.
double ref_execute(const orc_request* request, const uint32_t* target_qubits, const uint32_t* target_clbits,
                   uint32_t num_targets, uint32_t bitstring_width, double* statevector_out, uint64_t* outcomes,
                   uint64_t* counts, uint64_t* num_outcomes, double* expectation_out) {
    const orc_circuit& c = *request->circuit;
    const std::vector<menos::Operation> ops = operations_of(c);
    std::vector<std::pair<menos::QubitId, menos::ClbitId>> targets;
    for (uint32_t t = 0; t < num_targets; ++t) targets.emplace_back(target_qubits[t], target_clbits[t]);
    menos::Observable observable;
    for (uint32_t t = 0; t < request->num_terms; ++t) {
        menos::PauliString p;
        p.num_qubits = c.num_qubits;
        p.x_words = {request->x_masks[t]};
        p.z_words = {request->z_masks[t]};
        observable.terms.emplace_back(p, C{1.0, 0.0});
    }
    menos::EngineExecutionRequest req{
        .num_qubits = c.num_qubits,
        .operations = ops,
        .want_statevector = statevector_out != nullptr,
        .want_counts = outcomes != nullptr && request->shots > 0,
        .want_expectation = expectation_out != nullptr && request->num_terms > 0,
        .shots = request->shots,
        .measure_targets = targets,
        .bitstring_width = bitstring_width,
        .seed = request->seed,
        .global_phase = request->global_phase,
        .budget_bytes = ~uint64_t{0},
        .is_sole_region = true,
        .observable = &observable,
        .max_threads = std::max<uint32_t>(1, request->threads),
        .fuse_gates = request->fuse_gates != 0,
    };
    menos::engines::dense::DenseEngine engine;
    menos::common::MemoryLedger ledger;
    const menos::common::MemoryLedger::Scope scope(ledger);
    const auto t0 = std::chrono::steady_clock::now();
    menos::EngineExecutionResult result = engine.execute(req);
    const auto t1 = std::chrono::steady_clock::now();
    if (result.statevector.has_value())
        std::memcpy(statevector_out, result.statevector->data(), result.statevector->size() * sizeof(C));
    if (result.counts.has_value()) {
        uint64_t at = 0;
        for (const auto& [key, count] : *result.counts) {
            uint64_t word = 0;
            for (uint32_t b = 0; b < bitstring_width; ++b)
                if (key.test(b)) word |= uint64_t{1} << b;
            outcomes[at] = word;
            counts[at] = count;
            ++at;
        }
        *num_outcomes = at;
    }
    if (result.expectation_terms.has_value()) {
        for (std::size_t t = 0; t < result.expectation_terms->size(); ++t) {
            expectation_out[2 * t] = (*result.expectation_terms)[t].real();
            expectation_out[2 * t + 1] = (*result.expectation_terms)[t].imag();
        }
    }
    return std::chrono::duration<double>(t1 - t0).count();
}

}  // extern "C"
