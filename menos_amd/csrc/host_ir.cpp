// The host side of the execution plan: named-gate matrices, fusion plans and
// block matrices. This is plan compilation, done once per job on the host as
// the reference does it for its CUDA engine (matrices from host libm, reference
// src/ir/gate_table.cpp:69-98; block matrices built before the sweep,
// src/engines/dense/fused_block.cpp:61-99); the state itself is only ever
// touched by the device kernels.
#include <algorithm>
#include <array>
#include <bit>
#include <cmath>
#include <limits>
#include <numbers>

#include "gate_table.h"
#include "mh_internal.h"

namespace mh {

namespace {

enum : int32_t {
    kX, kY, kZ, kH, kS, kSdg, kT, kTdg, kSX, kSXdg, kRX, kRY, kRZ, kP, kU,
    kCX, kCZ, kSWAP, kCP, kRZZ, kCCX
};

struct GateInfo {
    uint32_t arity;
    uint32_t params;
};

GateInfo info_of(int32_t op) {
    static constexpr std::array<GateInfo, 21> kInfo{{
        {1, 0}, {1, 0}, {1, 0}, {1, 0}, {1, 0}, {1, 0}, {1, 0}, {1, 0}, {1, 0}, {1, 0},  // X .. SXdg
        {1, 1}, {1, 1}, {1, 1}, {1, 1}, {1, 3},                                          // RX RY RZ P U
        {2, 0}, {2, 0}, {2, 0}, {2, 1}, {2, 1},                                          // CX CZ SWAP CP RZZ
        {3, 0},                                                                          // CCX
    }};
    if (op < 0 || op > kCCX) fail(MH_ERR_INVALID, "unknown opcode " + std::to_string(op));
    return kInfo[static_cast<std::size_t>(op)];
}

// The generic kernel on a small host vector (the block-matrix build): each
// output from (+0, +0), adding M[out][in] * a[in] in increasing `in`.
void small_generic(cplx* state, uint32_t n, const cplx* m, const uint32_t* ops, uint32_t k) {
    const std::size_t dim = std::size_t{1} << k;
    std::array<uint32_t, 8> sorted{};
    std::copy(ops, ops + k, sorted.begin());
    std::sort(sorted.begin(), sorted.begin() + k);
    std::array<std::size_t, 64> idx{};
    std::array<cplx, 64> g{};
    const std::size_t outer = (std::size_t{1} << n) >> k;
    for (std::size_t o = 0; o < outer; ++o) {
        std::size_t base = o;
        for (uint32_t j = 0; j < k; ++j) {
            const std::size_t low = (std::size_t{1} << sorted[j]) - 1;
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
            double sr = 0.0, si = 0.0;
            for (std::size_t in = 0; in < dim; ++in) {
                const cplx e = m[out * dim + in];
                const double pr = e.real() * g[in].real() - e.imag() * g[in].imag();
                const double pi = e.real() * g[in].imag() + e.imag() * g[in].real();
                sr = sr + pr;
                si = si + pi;
            }
            state[idx[out]] = cplx{sr, si};
        }
    }
}

}  // namespace

// The named-gate matrices come from gate_table.cc, which is compiled with g++
// (see there for why the compiler matters).
std::vector<cplx> gate_matrix(int32_t op, const double* p, uint32_t num_params) {
    const GateInfo info = info_of(op);
    if (num_params != info.params)
        fail(MH_ERR_INVALID, "opcode " + std::to_string(op) + " takes " + std::to_string(info.params) + " parameters");
    std::vector<cplx> m(std::size_t{1} << (2 * info.arity));
    host_gate_matrix(op, p, m.data());
    return m;
}

std::vector<cplx> op_matrix(const CircuitView& cv, uint32_t i) {
    if (cv.opcode(i) == kOpUnitary) {
        const double* p = cv.params(i);
        std::vector<cplx> m(cv.num_params(i) / 2);
        for (std::size_t e = 0; e < m.size(); ++e) m[e] = cplx{p[2 * e], p[2 * e + 1]};
        return m;
    }
    return gate_matrix(cv.opcode(i), cv.params(i), cv.num_params(i));
}

// The fusion rule's diagonal flag (reference src/ir/fuse_diagonal.cpp:16-32,
// 180-193): named gates by opcode, unitaries by their matrix.
bool op_is_diagonal_for_fusion(const CircuitView& cv, uint32_t i) {
    switch (cv.opcode(i)) {
        case kZ: case kS: case kSdg: case kT: case kTdg: case kRZ: case kP: case kCZ: case kCP: case kRZZ:
            return true;
        case kOpUnitary: {
            const std::vector<cplx> m = op_matrix(cv, i);
            return matrix_is_diagonal(m.data(), std::size_t{1} << cv.num_qubits_of(i));
        }
        default:
            return false;
    }
}

// The checks the reference makes once at its IR boundary (src/ir/validate.cpp:102-158).
void validate_circuit(const mh_circuit* c) {
    if (c == nullptr) fail(MH_ERR_INVALID, "null circuit");
    if (c->num_qubits == 0 || c->num_qubits > 40) fail(MH_ERR_INVALID, "num_qubits must be 1..40");
    if (c->num_ops > 0 && (c->opcodes == nullptr || c->qubit_offsets == nullptr || c->param_offsets == nullptr))
        fail(MH_ERR_INVALID, "null circuit arrays");
    const CircuitView cv{c};
    for (uint32_t i = 0; i < c->num_ops; ++i) {
        const int32_t op = cv.opcode(i);
        const uint32_t k = cv.num_qubits_of(i);
        const uint32_t* q = cv.qubits(i);
        if (op == kOpMeasure) {
            if (k != 2 || q[0] >= c->num_qubits) fail(MH_ERR_INVALID, "op " + std::to_string(i) + ": bad measurement");
            continue;
        }
        if (op == kOpUnitary) {
            if (k == 0 || k > 6) fail(MH_ERR_INVALID, "op " + std::to_string(i) + ": a unitary acts on 1..6 qubits");
            if (cv.num_params(i) != 2u * (1u << (2 * k)))
                fail(MH_ERR_INVALID, "op " + std::to_string(i) + ": unitary matrix size mismatch");
        } else {
            const GateInfo info = info_of(op);
            if (k != info.arity) fail(MH_ERR_INVALID, "op " + std::to_string(i) + ": wrong operand count");
            if (cv.num_params(i) != info.params) fail(MH_ERR_INVALID, "op " + std::to_string(i) + ": wrong parameter count");
        }
        uint64_t seen = 0;
        for (uint32_t j = 0; j < k; ++j) {
            if (q[j] >= c->num_qubits) fail(MH_ERR_INVALID, "op " + std::to_string(i) + ": operand out of range");
            if ((seen >> q[j]) & 1ULL) fail(MH_ERR_INVALID, "op " + std::to_string(i) + ": repeated operand");
            seen |= 1ULL << q[j];
        }
    }
}

namespace {

bool may_join(uint64_t block, bool block_diag, uint64_t gate, bool gate_diag) {
    if (block == gate) return true;
    if ((block & gate) == block) return !gate_diag || block_diag;
    if ((block & gate) == gate) return !block_diag || gate_diag;
    return false;
}

uint64_t op_mask(const CircuitView& cv, uint32_t i) {
    uint64_t m = 0;
    for (uint32_t j = 0; j < cv.num_qubits_of(i); ++j) m |= uint64_t{1} << cv.qubits(i)[j];
    return m;
}

FusionPlan finish_plan(const std::vector<uint32_t>& block_of, const std::vector<uint64_t>& mask) {
    const std::size_t count = block_of.size();
    const std::size_t blocks = mask.size();
    FusionPlan p;
    p.block_begin.assign(blocks + 1, 0);
    for (std::size_t i = 0; i < count; ++i) ++p.block_begin[block_of[i] + 1];
    for (std::size_t b = 0; b < blocks; ++b) p.block_begin[b + 1] += p.block_begin[b];
    p.order.assign(count, 0);
    std::vector<uint32_t> fill(p.block_begin.begin(), p.block_begin.end() - 1);
    for (std::size_t i = 0; i < count; ++i) p.order[fill[block_of[i]]++] = static_cast<uint32_t>(i);
    p.qubits_begin.assign(blocks + 1, 0);
    for (std::size_t b = 0; b < blocks; ++b)
        p.qubits_begin[b + 1] = p.qubits_begin[b] + static_cast<uint32_t>(std::popcount(mask[b]));
    p.block_qubits.assign(p.qubits_begin[blocks], 0);
    for (std::size_t b = 0; b < blocks; ++b) {
        uint32_t at = p.qubits_begin[b];
        for (uint64_t r = mask[b]; r != 0; r &= r - 1) p.block_qubits[at++] = static_cast<uint32_t>(std::countr_zero(r));
        const uint32_t size = p.block_begin[b + 1] - p.block_begin[b];
        if (size >= 2) {
            ++p.fused_blocks;
            p.gates_absorbed += size - 1;
        }
    }
    return p;
}

// Containment fusion, the reference's default (plan_fusion,
// src/ir/fuse_gates.cpp:108-201, with may_join at :53-65): a gate joins the newest open block sharing a qubit with it only if
// one qubit set contains the other (diagonal rule aside); a measurement is a
// barrier block that closes every open block; blocks run in opening order.
FusionPlan containment_plan(const CircuitView& cv, bool keep_pairs) {
    constexpr uint32_t kNone = std::numeric_limits<uint32_t>::max();
    const uint32_t count = cv.num_ops();
    std::vector<uint32_t> block_of(count);
    std::vector<uint64_t> mask;
    std::vector<uint8_t> diag;
    // keep_pairs: per block, the control qubit of its controlled NOTs and the
    // qubits its non-diagonal gates other than those act on.
    std::vector<uint64_t> ctrl, dense;
    std::array<uint32_t, 64> newest{};
    newest.fill(kNone);
    for (uint32_t i = 0; i < count; ++i) {
        if (cv.opcode(i) == kOpMeasure) {
            block_of[i] = static_cast<uint32_t>(mask.size());
            mask.push_back(0);
            diag.push_back(0);
            ctrl.push_back(0);
            dense.push_back(0);
            newest.fill(kNone);
            continue;
        }
        const uint64_t gm = op_mask(cv, i);
        const bool gd = op_is_diagonal_for_fusion(cv, i);
        uint32_t dep = kNone;
        for (uint64_t r = gm; r != 0; r &= r - 1) {
            const uint32_t b = newest[static_cast<std::size_t>(std::countr_zero(r))];
            if (b != kNone && (dep == kNone || b > dep)) dep = b;
        }
        const bool cnot = cv.opcode(i) == kCX;
        const uint64_t gc = cnot ? uint64_t{1} << cv.qubits(i)[0] : 0;   // the control
        const uint64_t gdense = gd ? 0 : (cnot ? gm & ~gc : gm);
        // A block that stays block-diagonal in one control qubit applies as a
        // controlled pair: no non-diagonal gate may touch the control.
        auto keeps_pair = [&](uint32_t b) {
            const uint64_t c = ctrl[b] | gc;
            return std::popcount(c) <= 1 && ((dense[b] | gdense) & c) == 0;
        };
        uint32_t blk;
        if (dep != kNone && may_join(mask[dep], diag[dep] != 0, gm, gd) &&
            (!keep_pairs || (ctrl[dep] == 0 && gc == 0) || keeps_pair(dep))) {
            blk = dep;
            mask[blk] |= gm;
            diag[blk] = static_cast<uint8_t>(diag[blk] != 0 && gd);
            if (keep_pairs && keeps_pair(blk)) {
                ctrl[blk] |= gc;
                dense[blk] |= gdense;
            } else {
                ctrl[blk] = 0;
                dense[blk] = ~uint64_t{0};
            }
        } else {
            blk = static_cast<uint32_t>(mask.size());
            mask.push_back(gm);
            diag.push_back(gd ? 1 : 0);
            ctrl.push_back(gc);
            dense.push_back(gdense);
        }
        block_of[i] = blk;
        for (uint64_t r = gm; r != 0; r &= r - 1) newest[static_cast<std::size_t>(std::countr_zero(r))] = blk;
    }
    return finish_plan(block_of, mask);
}

// Width-capped fusion: consecutive gates share a block while the union of
// their qubits has at most `width` members. Order-preserving (every block is
// a contiguous run of the gate list), so the plan's product equals the
// unfused circuit's up to rounding, and the oracle applies the same blocks.
FusionPlan width_plan(const CircuitView& cv, uint32_t width) {
    const uint32_t count = cv.num_ops();
    std::vector<uint32_t> block_of(count);
    std::vector<uint64_t> mask;
    bool open = false;
    for (uint32_t i = 0; i < count; ++i) {
        if (cv.opcode(i) == kOpMeasure) {
            block_of[i] = static_cast<uint32_t>(mask.size());
            mask.push_back(0);
            open = false;
            continue;
        }
        const uint64_t gm = op_mask(cv, i);
        if (open && static_cast<uint32_t>(std::popcount(mask.back() | gm)) <= width) {
            mask.back() |= gm;
        } else {
            mask.push_back(gm);
            open = true;
        }
        block_of[i] = static_cast<uint32_t>(mask.size() - 1);
    }
    return finish_plan(block_of, mask);
}

FusionPlan unfused_plan(const CircuitView& cv) {
    const uint32_t count = cv.num_ops();
    std::vector<uint32_t> block_of(count);
    std::vector<uint64_t> mask(count);
    for (uint32_t i = 0; i < count; ++i) {
        block_of[i] = i;
        mask[i] = cv.opcode(i) == kOpMeasure ? 0 : op_mask(cv, i);
    }
    return finish_plan(block_of, mask);
}

}  // namespace

FusionPlan plan_fusion(const CircuitView& cv, int32_t fusion, uint32_t width) {
    switch (fusion) {
        case MH_FUSE_NONE: return unfused_plan(cv);
        case MH_FUSE_CONTAINMENT: return containment_plan(cv, false);
        case MH_FUSE_PAIRS: return containment_plan(cv, true);
        case MH_FUSE_AUTO:
            fail(MH_ERR_INVALID, "MH_FUSE_AUTO is resolved by the engine (mh_engine_estimate reports the plan it takes)");
        case MH_FUSE_WIDTH:
            if (width < 1 || width > 6) fail(MH_ERR_INVALID, "fusion_width must be 1..6");
            return width_plan(cv, width);
        default: fail(MH_ERR_INVALID, "unknown fusion mode");
    }
}

// reference src/engines/dense/fused_block.h:59-101 and fused_block.cpp:61-99:
// a one-gate block keeps the gate's own matrix and operand order; a fused
// block's matrix is the generic kernel's action on each basis column, in
// block order, on the block's ascending operands, then transposed.
std::vector<Block> build_blocks(const CircuitView& cv, const FusionPlan& plan) {
    std::vector<Block> out;
    out.reserve(plan.block_count());
    for (uint32_t b = 0; b < plan.block_count(); ++b) {
        const uint32_t begin = plan.block_begin[b], end = plan.block_begin[b + 1];
        if (end - begin == 1) {
            const uint32_t i = plan.order[begin];
            if (cv.opcode(i) == kOpMeasure) continue;
            Block blk;
            blk.operands.assign(cv.qubits(i), cv.qubits(i) + cv.num_qubits_of(i));
            blk.matrix = op_matrix(cv, i);
            blk.diagonal = matrix_is_diagonal(blk.matrix.data(), std::size_t{1} << blk.operands.size());
            out.push_back(std::move(blk));
            continue;
        }
        const uint32_t* bq = plan.block_qubits.data() + plan.qubits_begin[b];
        const uint32_t k = plan.qubits_begin[b + 1] - plan.qubits_begin[b];
        if (k > 6) fail(MH_ERR_INVALID, "a fused block wider than 6 qubits");
        const std::size_t dim = std::size_t{1} << k;
        std::vector<cplx> cols(dim * dim, cplx{0.0, 0.0});
        for (std::size_t c = 0; c < dim; ++c) cols[c * dim + c] = cplx{1.0, 0.0};
        for (uint32_t at = begin; at < end; ++at) {
            const uint32_t i = plan.order[at];
            if (cv.opcode(i) == kOpMeasure) fail(MH_ERR_INTERNAL, "a measurement inside a fused block");
            const std::vector<cplx> m = op_matrix(cv, i);
            std::array<uint32_t, 8> local{};
            const uint32_t gk = cv.num_qubits_of(i);
            for (uint32_t j = 0; j < gk; ++j) {
                const uint32_t* pos = std::find(bq, bq + k, cv.qubits(i)[j]);
                if (pos == bq + k) fail(MH_ERR_INTERNAL, "a block member outside its block");
                local[j] = static_cast<uint32_t>(pos - bq);
            }
            for (std::size_t c = 0; c < dim; ++c) small_generic(cols.data() + c * dim, k, m.data(), local.data(), gk);
        }
        for (std::size_t r = 0; r < dim; ++r)
            for (std::size_t c = r + 1; c < dim; ++c) std::swap(cols[r * dim + c], cols[c * dim + r]);
        Block blk;
        blk.operands.assign(bq, bq + k);
        blk.matrix = std::move(cols);
        blk.diagonal = matrix_is_diagonal(blk.matrix.data(), dim);
        out.push_back(std::move(blk));
    }
    return out;
}

}  // namespace mh

using namespace mh;

extern "C" {

int mh_gate_matrix(int32_t opcode, const double* params, double* out, uint32_t* k) {
    return guarded([&] {
        const std::vector<cplx> m = gate_matrix(opcode, params, info_of(opcode).params);
        std::copy(m.begin(), m.end(), reinterpret_cast<cplx*>(out));
        *k = info_of(opcode).arity;
    });
}

int mh_plan_fusion(const mh_circuit* circuit, int32_t fusion, uint32_t fusion_width, uint32_t* num_blocks,
                   uint32_t* order, uint32_t* block_begin, uint32_t* block_qubits, uint32_t* qubits_begin,
                   uint64_t* fused_blocks, uint64_t* gates_absorbed) {
    return guarded([&] {
        validate_circuit(circuit);
        const FusionPlan p = plan_fusion(CircuitView{circuit}, fusion, fusion_width);
        *num_blocks = p.block_count();
        std::copy(p.order.begin(), p.order.end(), order);
        std::copy(p.block_begin.begin(), p.block_begin.end(), block_begin);
        std::copy(p.block_qubits.begin(), p.block_qubits.end(), block_qubits);
        std::copy(p.qubits_begin.begin(), p.qubits_begin.end(), qubits_begin);
        *fused_blocks = p.fused_blocks;
        *gates_absorbed = p.gates_absorbed;
    });
}

}  // extern "C"
