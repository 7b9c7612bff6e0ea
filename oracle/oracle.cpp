// The oracle: a CPU restatement of the reference's dense statevector path.
// TEST INFRASTRUCTURE ONLY -- see oracle.h for who may load it and how each
// function is pinned. Built with -ffp-contract=off (oracle/Makefile), as the
// reference builds its kernels (reference src/kernels/CMakeLists.txt:14-21).
#include "oracle.h"

#include <algorithm>
#include <array>
#include <bit>
#include <cmath>
#include <complex>
#include <cstring>
#include <limits>
#include <numbers>
#include <random>
#include <stdexcept>
#include <thread>
#include <vector>

namespace {

using C = std::complex<double>;
using namespace std::complex_literals;

// Opcode numbering of the reference (include/menos/circuit.h, enum Opcode).
enum Op : int32_t {
    kX, kY, kZ, kH, kS, kSdg, kT, kTdg, kSX, kSXdg, kRX, kRY, kRZ, kP, kU,
    kCX, kCZ, kSWAP, kCP, kRZZ, kCCX, kUnitary, kMeasure
};

constexpr double kPi = std::numbers::pi_v<double>;

// The named-gate table: reference src/ir/gate_table.cpp:19-206. The same
// std::complex expressions, so libm and the complex operators round the same
// way (the matrix bits depend on them, reference gate_table.cpp:69-98).
std::vector<C> gate_matrix(int32_t op, const double* p) {
    switch (op) {
        case kX: return {0.0, 1.0, 1.0, 0.0};
        case kY: return {0.0, -1.0i, 1.0i, 0.0};
        case kZ: return {1.0, 0.0, 0.0, -1.0};
        case kH: {
            const double s = 1.0 / std::sqrt(2.0);
            return {s, s, s, -s};
        }
        case kS: return {1.0, 0.0, 0.0, 1.0i};
        case kSdg: return {1.0, 0.0, 0.0, -1.0i};
        case kT: return {1.0, 0.0, 0.0, std::exp(1.0i * (kPi / 4.0))};
        case kTdg: return {1.0, 0.0, 0.0, std::exp(-1.0i * (kPi / 4.0))};
        case kSX: return {0.5 * (1.0 + 1.0i), 0.5 * (1.0 - 1.0i), 0.5 * (1.0 - 1.0i), 0.5 * (1.0 + 1.0i)};
        case kSXdg: return {0.5 * (1.0 - 1.0i), 0.5 * (1.0 + 1.0i), 0.5 * (1.0 + 1.0i), 0.5 * (1.0 - 1.0i)};
        case kRX: {
            const double c = std::cos(p[0] / 2.0);
            const double s = std::sin(p[0] / 2.0);
            return {c, -1.0i * s, -1.0i * s, c};
        }
        case kRY: {
            const double c = std::cos(p[0] / 2.0);
            const double s = std::sin(p[0] / 2.0);
            return {c, -s, s, c};
        }
        case kRZ: return {std::exp(-1.0i * (p[0] / 2.0)), 0.0, 0.0, std::exp(1.0i * (p[0] / 2.0))};
        case kP: return {1.0, 0.0, 0.0, std::exp(1.0i * p[0])};
        case kU: {
            const double c = std::cos(p[0] / 2.0);
            const double s = std::sin(p[0] / 2.0);
            return {c, -std::exp(1.0i * p[2]) * s, std::exp(1.0i * p[1]) * s,
                    std::exp(1.0i * (p[1] + p[2])) * c};
        }
        case kCX: {
            std::vector<C> m(16, 0.0);
            m[0] = 1.0; m[7] = 1.0; m[10] = 1.0; m[13] = 1.0;
            return m;
        }
        case kCZ: {
            std::vector<C> m(16, 0.0);
            m[0] = 1.0; m[5] = 1.0; m[10] = 1.0; m[15] = -1.0;
            return m;
        }
        case kSWAP: {
            std::vector<C> m(16, 0.0);
            m[0] = 1.0; m[6] = 1.0; m[9] = 1.0; m[15] = 1.0;
            return m;
        }
        case kCP: {
            std::vector<C> m(16, 0.0);
            m[0] = 1.0; m[5] = 1.0; m[10] = 1.0; m[15] = std::exp(1.0i * p[0]);
            return m;
        }
        case kRZZ: {
            const C same = std::exp(-0.5i * p[0]);
            const C diff = std::exp(0.5i * p[0]);
            std::vector<C> m(16, 0.0);
            m[0] = same; m[5] = diff; m[10] = diff; m[15] = same;
            return m;
        }
        case kCCX: {
            std::vector<C> m(64, 0.0);
            for (int i = 0; i < 8; ++i) m[i * 8 + i] = 1.0;
            m[3 * 8 + 3] = 0.0; m[7 * 8 + 7] = 0.0; m[3 * 8 + 7] = 1.0; m[7 * 8 + 3] = 1.0;
            return m;
        }
        default: return {};
    }
}

struct OpView {
    int32_t opcode;
    const uint32_t* qubits;
    uint32_t nq;
    const double* params;
    uint32_t np;
};

OpView view(const orc_circuit& c, uint32_t i) {
    return {c.opcodes[i], c.qubits + c.qubit_offsets[i], c.qubit_offsets[i + 1] - c.qubit_offsets[i],
            c.params + c.param_offsets[i], c.param_offsets[i + 1] - c.param_offsets[i]};
}

bool is_gate(const OpView& v) { return v.opcode != kMeasure; }

std::vector<C> matrix_of(const OpView& v) {
    if (v.opcode == kUnitary) {
        std::vector<C> m(v.np / 2);
        for (std::size_t e = 0; e < m.size(); ++e) m[e] = C{v.params[2 * e], v.params[2 * e + 1]};
        return m;
    }
    return gate_matrix(v.opcode, v.params);
}

// Every off-diagonal entry exactly zero (reference src/ir/fuse_diagonal.cpp:34-44).
bool matrix_is_diagonal(const std::vector<C>& m, std::size_t dim) {
    for (std::size_t r = 0; r < dim; ++r)
        for (std::size_t c = 0; c < dim; ++c)
            if (r != c && (m[r * dim + c].real() != 0.0 || m[r * dim + c].imag() != 0.0)) return false;
    return true;
}

// The fusion rule's diagonal flag (reference src/ir/fuse_diagonal.cpp:16-32,
// 180-193): a named gate by its opcode, a unitary by its matrix.
bool is_diagonal_operation(const OpView& v) {
    switch (v.opcode) {
        case kZ: case kS: case kSdg: case kT: case kTdg: case kRZ: case kP: case kCZ: case kCP: case kRZZ:
            return true;
        case kUnitary:
            return matrix_is_diagonal(matrix_of(v), std::size_t{1} << v.nq);
        default:
            return false;
    }
}

// Runs body(begin, end) over [0, count) in contiguous chunks on `threads`
// threads; every chunk is independent (reference thread_pool.h:6-45 contract).
template <typename Body>
void parallel_chunks(std::size_t count, uint32_t threads, Body&& body) {
    if (threads <= 1 || count < (std::size_t{1} << 12)) {
        body(std::size_t{0}, count);
        return;
    }
    std::vector<std::thread> team;
    const std::size_t chunk = (count + threads - 1) / threads;
    for (uint32_t t = 0; t < threads; ++t) {
        const std::size_t b = std::min(count, t * chunk);
        const std::size_t e = std::min(count, b + chunk);
        if (b < e) team.emplace_back([&body, b, e] { body(b, e); });
    }
    for (auto& th : team) th.join();
}

// The generic kernel: reference src/kernels/cpu/dense_kernels.cpp:465-524.
// Per group: gather the 2^k amplitudes in operand order, then for each
// output sum = (+0,+0); sum += M[out][in] * g[in] for in ascending; scatter.
void apply_generic(C* state, uint32_t n, const C* m, const uint32_t* ops, uint32_t k, uint32_t threads) {
    const std::size_t dim = std::size_t{1} << k;
    std::vector<uint32_t> sorted(ops, ops + k);
    std::sort(sorted.begin(), sorted.end());
    const std::size_t outer = (std::size_t{1} << n) >> k;
    parallel_chunks(outer, threads, [&](std::size_t begin, std::size_t end) {
        std::vector<std::size_t> idx(dim);
        std::vector<C> g(dim);
        for (std::size_t o = begin; o < end; ++o) {
            // Zero-bit insertion at every operand position, lowest first.
            std::size_t base = o;
            for (uint32_t q : sorted) {
                const std::size_t low = (std::size_t{1} << q) - 1;
                base = ((base & ~low) << 1) | (base & low);
            }
            for (std::size_t l = 0; l < dim; ++l) {
                std::size_t i = base;
                for (uint32_t j = 0; j < k; ++j)
                    if ((l >> j) & 1U) i |= std::size_t{1} << ops[j];
                idx[l] = i;
                g[l] = state[i];
            }
            for (std::size_t out = 0; out < dim; ++out) {
                C sum{0.0, 0.0};
                for (std::size_t in = 0; in < dim; ++in) sum += m[out * dim + in] * g[in];
                state[idx[out]] = sum;
            }
        }
    });
}

// Containment fusion: reference src/ir/fuse_gates.cpp:135-147 (may_join) and
// :190-283 (plan_fusion).
struct Plan {
    std::vector<uint32_t> order, block_begin, block_qubits, qubits_begin;
    uint64_t fused_blocks = 0, gates_absorbed = 0;
};

bool may_join(uint64_t block, bool block_diag, uint64_t gate, bool gate_diag) {
    if (block == gate) return true;
    if ((block & gate) == block) return !gate_diag || block_diag;
    if ((block & gate) == gate) return !block_diag || gate_diag;
    return false;
}

Plan plan_fusion(const orc_circuit& c) {
    constexpr uint32_t kNone = std::numeric_limits<uint32_t>::max();
    const uint32_t count = c.num_ops;
    std::vector<uint32_t> block_of(count);
    std::vector<uint64_t> mask;
    std::vector<bool> diag;
    std::array<uint32_t, 64> newest;
    newest.fill(kNone);
    for (uint32_t i = 0; i < count; ++i) {
        const OpView v = view(c, i);
        if (!is_gate(v)) {
            block_of[i] = static_cast<uint32_t>(mask.size());
            mask.push_back(0);
            diag.push_back(false);
            newest.fill(kNone);
            continue;
        }
        uint64_t gm = 0;
        for (uint32_t j = 0; j < v.nq; ++j) gm |= uint64_t{1} << v.qubits[j];
        const bool gd = is_diagonal_operation(v);
        uint32_t dep = kNone;
        for (uint64_t r = gm; r != 0; r &= r - 1) {
            const uint32_t b = newest[std::countr_zero(r)];
            if (b != kNone && (dep == kNone || b > dep)) dep = b;
        }
        uint32_t blk;
        if (dep != kNone && may_join(mask[dep], diag[dep], gm, gd)) {
            blk = dep;
            mask[blk] |= gm;
            diag[blk] = diag[blk] && gd;
        } else {
            blk = static_cast<uint32_t>(mask.size());
            mask.push_back(gm);
            diag.push_back(gd);
        }
        block_of[i] = blk;
        for (uint64_t r = gm; r != 0; r &= r - 1) newest[std::countr_zero(r)] = blk;
    }
    const std::size_t blocks = mask.size();
    Plan p;
    p.block_begin.assign(blocks + 1, 0);
    for (uint32_t i = 0; i < count; ++i) ++p.block_begin[block_of[i] + 1];
    for (std::size_t b = 0; b < blocks; ++b) p.block_begin[b + 1] += p.block_begin[b];
    p.order.assign(count, 0);
    std::vector<uint32_t> fill(p.block_begin.begin(), p.block_begin.end() - 1);
    for (uint32_t i = 0; i < count; ++i) p.order[fill[block_of[i]]++] = i;
    p.qubits_begin.assign(blocks + 1, 0);
    for (std::size_t b = 0; b < blocks; ++b)
        p.qubits_begin[b + 1] = p.qubits_begin[b] + static_cast<uint32_t>(std::popcount(mask[b]));
    p.block_qubits.assign(p.qubits_begin[blocks], 0);
    for (std::size_t b = 0; b < blocks; ++b) {
        uint32_t at = p.qubits_begin[b];
        for (uint64_t r = mask[b]; r != 0; r &= r - 1) p.block_qubits[at++] = std::countr_zero(r);
        const uint32_t size = p.block_begin[b + 1] - p.block_begin[b];
        if (size >= 2) {
            ++p.fused_blocks;
            p.gates_absorbed += size - 1;
        }
    }
    return p;
}

// The block matrix: reference src/engines/dense/fused_block.cpp:61-99. Column
// c starts as e_c; every member gate is applied to it with the generic kernel
// on its position within the block's ascending operands; then transposed.
std::vector<C> fused_block_matrix(const orc_circuit& c, const orc_plan& p, uint32_t b) {
    const uint32_t* bq = p.block_qubits + p.qubits_begin[b];
    const uint32_t k = p.qubits_begin[b + 1] - p.qubits_begin[b];
    const std::size_t dim = std::size_t{1} << k;
    std::vector<C> cols(dim * dim, C{0.0, 0.0});
    for (std::size_t col = 0; col < dim; ++col) cols[col * dim + col] = C{1.0, 0.0};
    for (uint32_t at = p.block_begin[b]; at < p.block_begin[b + 1]; ++at) {
        const OpView v = view(c, p.order[at]);
        if (!is_gate(v)) throw std::runtime_error("fused block holds a measurement");
        const std::vector<C> m = matrix_of(v);
        std::vector<uint32_t> local;
        for (uint32_t j = 0; j < v.nq; ++j) {
            const uint32_t* pos = std::find(bq, bq + k, v.qubits[j]);
            if (pos == bq + k) throw std::runtime_error("operand outside its block");
            local.push_back(static_cast<uint32_t>(pos - bq));
        }
        for (std::size_t col = 0; col < dim; ++col)
            apply_generic(cols.data() + col * dim, k, m.data(), local.data(), v.nq, 1);
    }
    for (std::size_t r = 0; r < dim; ++r)
        for (std::size_t col = r + 1; col < dim; ++col) std::swap(cols[r * dim + col], cols[col * dim + r]);
    return cols;
}

C i_power(int k) {
    switch (k & 3) {
        case 0: return {1.0, 0.0};
        case 1: return {0.0, 1.0};
        case 2: return {-1.0, 0.0};
        default: return {0.0, -1.0};
    }
}

// reference src/kernels/cpu/dense_kernels.cpp:544-569.
C expectation_of_term(const C* s, std::size_t size, uint64_t x, uint64_t z) {
    C acc{0.0, 0.0};
    for (std::size_t j = 0; j < size; ++j) {
        const uint64_t k = j ^ x;
        const bool negative = (std::popcount(z & k) & 1) != 0;
        const C term = std::conj(s[j]) * s[k];
        acc += negative ? -term : term;
    }
    return i_power(std::popcount(x & z)) * acc;
}

// reference src/common/rng.cpp:8-24: (word >> 11) * 2^-53.
double word_to_uniform(uint64_t w) { return static_cast<double>(w >> 11) * (1.0 / 9007199254740992.0); }

// The sole-region walk: reference src/engines/dense/sample.cpp:24-75.
template <typename Prob>
void sample_walk(std::size_t size, Prob prob, uint64_t shots, uint64_t seed, uint64_t* out) {
    if (shots == 0) return;
    std::mt19937_64 rng(seed);
    std::vector<double> u(shots);
    for (uint64_t i = 0; i < shots; ++i) u[i] = word_to_uniform(rng());
    std::sort(u.begin(), u.end());
    double cumulative = 0.0;
    std::size_t shot = 0;
    for (std::size_t basis = 0; basis < size && shot < u.size(); ++basis) {
        cumulative += prob(basis);
        while (shot < u.size() && u[shot] < cumulative) out[shot++] = basis;
    }
    while (shot < u.size() && size != 0) out[shot++] = size - 1;
}

}  // namespace

extern "C" {

int orc_gate_matrix(int32_t opcode, const double* params, double* out) {
    const std::vector<C> m = gate_matrix(opcode, params);
    if (m.empty()) return -1;
    std::memcpy(out, m.data(), m.size() * sizeof(C));
    return std::countr_zero(m.size()) / 2;
}

// reference src/kernels/cpu/dense_kernels.cpp:21-27.
void orc_initialize(double* state, uint64_t amplitudes, double global_phase) {
    const C phase = std::exp(C{0.0, global_phase});
    C* s = reinterpret_cast<C*>(state);
    for (uint64_t i = 0; i < amplitudes; ++i) s[i] = C{0.0, 0.0};
    s[0] = phase;
}

void orc_apply_generic(double* state, uint32_t num_qubits, const double* matrix,
                       const uint32_t* operands, uint32_t k, uint32_t threads) {
    apply_generic(reinterpret_cast<C*>(state), num_qubits, reinterpret_cast<const C*>(matrix), operands, k,
                  threads);
}

uint32_t orc_plan_fusion(const orc_circuit* circuit, uint32_t* order, uint32_t* block_begin,
                         uint32_t* block_qubits, uint32_t* qubits_begin, uint64_t* fused_blocks,
                         uint64_t* gates_absorbed) {
    const Plan p = plan_fusion(*circuit);
    std::copy(p.order.begin(), p.order.end(), order);
    std::copy(p.block_begin.begin(), p.block_begin.end(), block_begin);
    std::copy(p.block_qubits.begin(), p.block_qubits.end(), block_qubits);
    std::copy(p.qubits_begin.begin(), p.qubits_begin.end(), qubits_begin);
    *fused_blocks = p.fused_blocks;
    *gates_absorbed = p.gates_absorbed;
    return static_cast<uint32_t>(p.block_begin.size() - 1);
}

void orc_fused_block_matrix(const orc_circuit* circuit, const orc_plan* plan, uint32_t block, double* out) {
    const std::vector<C> m = fused_block_matrix(*circuit, *plan, block);
    std::memcpy(out, m.data(), m.size() * sizeof(C));
}

void orc_expectation_of_term(const double* state, uint32_t num_qubits, uint64_t x_mask, uint64_t z_mask,
                             double* out2) {
    const C e = expectation_of_term(reinterpret_cast<const C*>(state), std::size_t{1} << num_qubits, x_mask,
                                    z_mask);
    out2[0] = e.real();
    out2[1] = e.imag();
}

void orc_sample_indices(const double* state, uint32_t num_qubits, uint64_t shots, uint64_t seed,
                        uint64_t* basis_of_shot) {
    const C* s = reinterpret_cast<const C*>(state);
    sample_walk(std::size_t{1} << num_qubits, [s](std::size_t b) { return std::norm(s[b]); }, shots, seed,
                basis_of_shot);
}

void orc_sample_indices_from_probabilities(const double* probabilities, uint64_t size, uint64_t shots,
                                           uint64_t seed, uint64_t* basis_of_shot) {
    sample_walk(size, [probabilities](std::size_t b) { return probabilities[b]; }, shots, seed, basis_of_shot);
}

void orc_cumulative(const double* state, uint32_t num_qubits, double* cumulative) {
    const C* s = reinterpret_cast<const C*>(state);
    double c = 0.0;
    for (std::size_t b = 0; b < (std::size_t{1} << num_qubits); ++b) {
        c += std::norm(s[b]);
        cumulative[b] = c;
    }
}

void orc_region_draw_indices(const double* state, uint32_t num_qubits, uint64_t draws, uint64_t seed,
                             uint64_t* basis_of_draw) {
    const std::size_t size = std::size_t{1} << num_qubits;
    std::vector<double> cum(size);
    orc_cumulative(state, num_qubits, cum.data());
    std::mt19937_64 rng(seed);
    for (uint64_t d = 0; d < draws; ++d) {
        const double u = word_to_uniform(rng());
        const auto it = std::upper_bound(cum.begin(), cum.end(), u);
        basis_of_draw[d] = it == cum.end() ? size - 1 : static_cast<uint64_t>(it - cum.begin());
    }
}

void orc_mt19937_64(uint64_t seed, uint64_t count, uint64_t* out) {
    std::mt19937_64 rng(seed);
    for (uint64_t i = 0; i < count; ++i) out[i] = rng();
}

void orc_execute(const orc_request* request, double* statevector_out, uint64_t* basis_of_shot,
                 double* expectation_out) {
    const orc_circuit& c = *request->circuit;
    const uint32_t n = c.num_qubits;
    const std::size_t size = std::size_t{1} << n;
    const uint32_t threads = std::max<uint32_t>(1, request->threads);
    std::vector<C> state(size);
    orc_initialize(reinterpret_cast<double*>(state.data()), size, request->global_phase);

    auto apply_gate = [&](const OpView& v) {
        if (!is_gate(v)) return;
        const std::vector<C> m = matrix_of(v);
        apply_generic(state.data(), n, m.data(), v.qubits, v.nq, threads);
    };
    Plan own;
    orc_plan plan_view{};
    const orc_plan* plan = request->plan;
    if (plan == nullptr && request->fuse_gates != 0) {
        own = plan_fusion(c);
        plan_view = {static_cast<uint32_t>(own.block_begin.size() - 1), own.order.data(), own.block_begin.data(),
                     own.block_qubits.data(), own.qubits_begin.data()};
        plan = &plan_view;
    }
    if (plan == nullptr) {
        for (uint32_t i = 0; i < c.num_ops; ++i) apply_gate(view(c, i));
    } else {
        // reference src/engines/dense/fused_block.h:59-101.
        for (uint32_t b = 0; b < plan->num_blocks; ++b) {
            const uint32_t members = plan->block_begin[b + 1] - plan->block_begin[b];
            if (members == 1) {
                apply_gate(view(c, plan->order[plan->block_begin[b]]));
            } else {
                const std::vector<C> m = fused_block_matrix(c, *plan, b);
                apply_generic(state.data(), n, m.data(), plan->block_qubits + plan->qubits_begin[b],
                              plan->qubits_begin[b + 1] - plan->qubits_begin[b], threads);
            }
        }
    }
    if (basis_of_shot != nullptr) {
        orc_sample_indices(reinterpret_cast<const double*>(state.data()), n, request->shots, request->seed,
                           basis_of_shot);
    }
    if (expectation_out != nullptr) {
        for (uint32_t t = 0; t < request->num_terms; ++t) {
            const C e = expectation_of_term(state.data(), size, request->x_masks[t], request->z_masks[t]);
            expectation_out[2 * t] = e.real();
            expectation_out[2 * t + 1] = e.imag();
        }
    }
    if (statevector_out != nullptr) std::memcpy(statevector_out, state.data(), size * sizeof(C));
}

uint64_t orc_fnv1a(const void* data, uint64_t size, uint64_t seed_hash) {
    const auto* bytes = static_cast<const unsigned char*>(data);
    uint64_t h = seed_hash;
    for (uint64_t i = 0; i < size; ++i) {
        h ^= bytes[i];
        h *= 0x100000001b3ULL;
    }
    return h;
}

}  // extern "C"
