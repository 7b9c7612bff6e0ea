// Initialisation, the generic per-gate kernel and the gate entry points.
//
// generic_gate_kernel is the reference's one-thread-per-group kernel
// (reference src/kernels/cuda/dense_kernels.cu:63-107): gather the 2^k
// amplitudes of a group, multiply, scatter. It backs mh_apply_matrix_generic
// (the independent path the window passes are tested against) and blocks wider
// than a window pass's register slots. Its arithmetic is the reference CPU
// generic kernel's (dense_kernels.cpp:465-524): every output starts at (+0, +0)
// and adds M[out][in] * a[in] for in ascending, each operation rounded on its
// own (-ffp-contract=off). complex64 accumulates with FP32 FMA.
#include <algorithm>
#include <cstring>

#include "gate_index.h"
#include "mh_internal.h"
#include "window.h"

namespace mh {
namespace {

constexpr int kThreads = kGroupThreads;
constexpr uint32_t kMaxK = 6;

// ---- complex arithmetic

__device__ __forceinline__ void cmac(double& sr, double& si, double mr, double mi, double ar, double ai) {
    const double pr = mr * ar - mi * ai;
    const double pi = mr * ai + mi * ar;
    sr = sr + pr;
    si = si + pi;
}

__device__ __forceinline__ void cmac(float& sr, float& si, float mr, float mi, float ar, float ai) {
    sr = __builtin_fmaf(mr, ar, sr);
    sr = __builtin_fmaf(-mi, ai, sr);
    si = __builtin_fmaf(mr, ai, si);
    si = __builtin_fmaf(mi, ar, si);
}

template <typename R> struct Vec;
template <> struct Vec<double> { using type = double2; };
template <> struct Vec<float> { using type = float2; };

// ---- the generic per-gate kernel

struct GenericArgs {
    void* state;
    const void* matrix;
    uint64_t groups;
    uint32_t k;
    uint32_t ops[kMaxK];     // operand order
    uint32_t sorted[kMaxK];  // ascending
};

template <typename R, uint32_t K>
__device__ __forceinline__ void generic_group(const GenericArgs& a, const typename Vec<R>::type* m, uint64_t g) {
    using V = typename Vec<R>::type;
    constexpr uint32_t dim = 1u << K;
    V* s = static_cast<V*>(a.state);
    const uint64_t base = group_base(g, a.sorted, K);
    V v[dim];
#pragma unroll
    for (uint32_t l = 0; l < dim; ++l) v[l] = s[group_member(base, l, a.ops, K)];
#pragma unroll(K <= 2 ? dim : 1)
    for (uint32_t o = 0; o < dim; ++o) {
        R sr = R(0), si = R(0);
#pragma unroll
        for (uint32_t in = 0; in < dim; ++in) {
            const V e = m[o * dim + in];
            cmac(sr, si, e.x, e.y, v[in].x, v[in].y);
        }
        V r;
        r.x = sr;
        r.y = si;
        s[group_member(base, o, a.ops, K)] = r;
    }
}

// One thread per group of 2^K amplitudes. The matrix stays in global memory:
// every lane of a wave reads the same entry at the same time, so the entries
// come in through scalar loads and the lanes' registers hold amplitudes only.
template <typename R, uint32_t K>
__global__ void __launch_bounds__(kThreads) generic_gate_kernel(GenericArgs a) {
    using V = typename Vec<R>::type;
    const V* m = static_cast<const V*>(a.matrix);
    const uint64_t stride = uint64_t(gridDim.x) * kThreads;
    for (uint64_t g = uint64_t(blockIdx.x) * kThreads + threadIdx.x; g < a.groups; g += stride)
        generic_group<R, K>(a, m, g);
}

// ---- complex64 blocks of K = 4..6 qubits on the matrix cores
//
// A block applied to 2^(n-K) groups of 2^K amplitudes is one product
// OUT = A x IN of real matrices: IN's column j is group j's amplitudes as
// 2^(K+1) reals (re, im of amplitude 0, re, im of amplitude 1, ...), and A is
// the complex matrix M in the same real form, A[2o + p][2m + q] =
//   Mr[o][m] (p = q = 0), -Mi[o][m] (p = 0, q = 1), Mi[o][m] (p = 1, q = 0),
//   Mr[o][m] (p = q = 1).
// A wave takes 32 groups at a time (a tile): per row block rb of 32 rows of
// A and per amplitude s of the group, one v_mfma_f32_32x32x2_f32 adds
// A[rows][2s, 2s + 1] x IN[2s, 2s + 1][32 groups]. Operand layouts (wave64):
// lane l holds A[32 rb + l % 32][2s + l / 32] and IN[2s + l / 32][l % 32];
// output register r of lane l is row 8 (r / 4) + 4 (l / 32) + r % 4 of the
// block, column l % 32. Rows 2o and 2o + 1 (re and im of amplitude o) land in
// consecutive registers of one lane, so each lane stores whole amplitudes.
// The host lays A out per (rb, s) as the 64 lanes read it. K <= 5 keeps A in
// registers across tiles; K = 6 (A is 64 KiB) stages it in LDS.
constexpr uint32_t kMfmaTileGroups = 32;

typedef float MfmaAcc __attribute__((ext_vector_type(16)));

struct MfmaArgs {
    float2* state;
    const float* a;          // (rb * 2^K + s) * 64 + lane
    uint64_t tiles;          // groups / 32
    uint32_t ops[kMaxK];     // operand order
    uint32_t sorted[kMaxK];  // ascending
};

// One tile of 32 groups; a_at(e) is the lane's A entry e = rb * 2^K + s.
template <uint32_t K, typename AEntry>
__device__ __forceinline__ void mfma_tile(const MfmaArgs& a, const AEntry& a_at, uint64_t tile, uint32_t lane) {
    constexpr uint32_t S = 1u << K;               // amplitudes per group
    constexpr uint32_t RBLK = (2u * S) / 32u;      // row blocks of A
    const uint32_t j = lane & 31u, h = lane >> 5;
    const uint64_t base = group_base(tile * kMfmaTileGroups + j, a.sorted, K);
    const float* st = reinterpret_cast<const float*>(a.state);
    float in[S];
#pragma unroll
    for (uint32_t s = 0; s < S; ++s) in[s] = st[2 * group_member(base, s, a.ops, K) + h];
    MfmaAcc acc[RBLK];
#pragma unroll
    for (uint32_t rb = 0; rb < RBLK; ++rb) {
        acc[rb] = MfmaAcc{};
#pragma unroll
        for (uint32_t s = 0; s < S; ++s)
            acc[rb] = __builtin_amdgcn_mfma_f32_32x32x2f32(a_at(rb * S + s), in[s], acc[rb], 0, 0, 0);
    }
#pragma unroll
    for (uint32_t rb = 0; rb < RBLK; ++rb) {
#pragma unroll
        for (uint32_t q = 0; q < 4; ++q) {
            const uint32_t o = 16 * rb + 4 * q + 2 * h;
            a.state[group_member(base, o, a.ops, K)] = make_float2(acc[rb][4 * q], acc[rb][4 * q + 1]);
            a.state[group_member(base, o + 1, a.ops, K)] = make_float2(acc[rb][4 * q + 2], acc[rb][4 * q + 3]);
        }
    }
}

// Waves take tiles wave, wave + waves, ... (every wave exits when the tiles
// run out).
template <uint32_t K>
__global__ void __launch_bounds__(kThreads) mfma_gate_kernel(MfmaArgs a) {
    constexpr uint32_t S = 1u << K;
    constexpr uint32_t A_FLOATS = (2u * S / 32u) * S * 64u;
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t wave = uint64_t(blockIdx.x) * (kThreads / 64) + (threadIdx.x >> 6);
    const uint64_t waves = uint64_t(gridDim.x) * (kThreads / 64);
    if constexpr (K == 6) {
        // A (64 KiB) from LDS, staged once per workgroup; every thread reaches
        // the barrier before any tile.
        __shared__ float am[A_FLOATS];
        for (uint32_t e = threadIdx.x; e < A_FLOATS; e += kThreads) am[e] = a.a[e];
        __syncthreads();
        const auto a_at = [&](uint32_t e) { return am[e * 64 + lane]; };
        for (uint64_t t = wave; t < a.tiles; t += waves) mfma_tile<K>(a, a_at, t, lane);
    } else {
        // A in registers: the lane's own entries, loaded once.
        float am[A_FLOATS / 64];
#pragma unroll
        for (uint32_t e = 0; e < A_FLOATS / 64; ++e) am[e] = a.a[e * 64 + lane];
        const auto a_at = [&](uint32_t e) { return am[e]; };
        for (uint64_t t = wave; t < a.tiles; t += waves) mfma_tile<K>(a, a_at, t, lane);
    }
}

// A's real form in the lane order the kernel reads (see above).
std::vector<float> mfma_matrix(const cplx* m, uint32_t k) {
    const uint32_t S = 1u << k, D = 2u * S, rblk = D / 32u;
    std::vector<float> out(std::size_t(rblk) * S * 64);
    for (uint32_t rb = 0; rb < rblk; ++rb)
        for (uint32_t s = 0; s < S; ++s)
            for (uint32_t l = 0; l < 64; ++l) {
                const uint32_t row = 32 * rb + (l & 31u), col = 2 * s + (l >> 5);
                const uint32_t o = row >> 1, p = row & 1u, mm = col >> 1, q = col & 1u;
                const cplx e = m[std::size_t(o) * S + mm];
                const double v = p == q ? e.real() : (p == 0 ? -e.imag() : e.imag());
                out[(std::size_t(rb) * S + s) * 64 + l] = static_cast<float>(v);
            }
    return out;
}

// Blocks the matrix cores take: complex64, 5 or 6 qubits, at least one tile
// of 32 groups.
bool mfma_applies(int32_t precision, uint32_t n, uint32_t k) {
    return precision == MH_C64 && k >= 5 && k <= 6 && n >= k + 5;
}

void run_mfma(void* state, uint32_t n, const cplx* matrix, const uint32_t* operands, uint32_t k) {
    context();
    const std::vector<float> am = mfma_matrix(matrix, k);
    DeviceBuffer dm(am.size() * sizeof(float));
    upload(dm.get(), am.data(), am.size() * sizeof(float));
    MfmaArgs a{};
    a.state = static_cast<float2*>(state);
    a.a = dm.as<float>();
    if (k < 4 || k > 6 || n < k + 5) fail(MH_ERR_INTERNAL, "matrix-core kernel: block shape out of range");
    a.tiles = ((uint64_t{1} << n) >> k) / kMfmaTileGroups;
    for (uint32_t j = 0; j < k; ++j) a.ops[j] = a.sorted[j] = operands[j];
    std::sort(a.sorted, a.sorted + k);
    for (uint32_t j = 0; j < k; ++j)
        if (a.sorted[j] >= n || (j > 0 && a.sorted[j] == a.sorted[j - 1]))
            fail(MH_ERR_INTERNAL, "matrix-core kernel: operands out of range or repeated");
    // Four waves per workgroup, at most one tile per wave per round of the grid.
    const uint64_t groups_of_waves = (a.tiles + (kThreads / 64) - 1) / (kThreads / 64);
    const dim3 grid(static_cast<uint32_t>(std::min<uint64_t>(groups_of_waves, 4096))), block(kThreads);
    switch (k) {
        case 4: hipLaunchKernelGGL(mfma_gate_kernel<4>, grid, block, 0, streams().main, a); break;
        case 5: hipLaunchKernelGGL(mfma_gate_kernel<5>, grid, block, 0, streams().main, a); break;
        default: hipLaunchKernelGGL(mfma_gate_kernel<6>, grid, block, 0, streams().main, a); break;
    }
    MH_HIP(hipGetLastError());
    synchronize();
}

template <typename R>
__global__ void __launch_bounds__(kThreads) init_kernel(void* state, uint64_t size, R re, R im) {
    using V = typename Vec<R>::type;
    V* s = static_cast<V*>(state);
    const uint64_t stride = uint64_t(gridDim.x) * kThreads;
    for (uint64_t i = uint64_t(blockIdx.x) * kThreads + threadIdx.x; i < size; i += stride) {
        V v;
        v.x = i == 0 ? re : R(0);
        v.y = i == 0 ? im : R(0);
        s[i] = v;
    }
}

// out[i] = state[index[i]] for i < count (indices checked on the host).
template <typename V>
__global__ void __launch_bounds__(kThreads) gather_kernel(const V* state, const uint64_t* index, uint64_t count,
                                                         V* out) {
    const uint64_t stride = uint64_t(gridDim.x) * kThreads;
    for (uint64_t i = uint64_t(blockIdx.x) * kThreads + threadIdx.x; i < count; i += stride) out[i] = state[index[i]];
}

uint32_t grid_for(uint64_t items) { return group_grid(items); }

template <typename R>
void run_generic(void* state, uint32_t n, const cplx* matrix, const uint32_t* operands, uint32_t k) {
    using V = typename Vec<R>::type;
    context();
    const std::size_t dim = std::size_t{1} << k;
    std::vector<V> m(dim * dim);
    for (std::size_t e = 0; e < dim * dim; ++e) {
        m[e].x = static_cast<R>(matrix[e].real());
        m[e].y = static_cast<R>(matrix[e].imag());
    }
    DeviceBuffer dm(m.size() * sizeof(V));
    upload(dm.get(), m.data(), m.size() * sizeof(V));
    GenericArgs a{};
    a.state = state;
    a.matrix = dm.get();
    if (k < 1 || k > kMaxK || k > n) fail(MH_ERR_INTERNAL, "generic kernel: block width out of range");
    a.groups = (uint64_t{1} << n) >> k;
    a.k = k;
    for (uint32_t j = 0; j < k; ++j) a.ops[j] = a.sorted[j] = operands[j];
    std::sort(a.sorted, a.sorted + k);
    for (uint32_t j = 0; j < k; ++j)
        if (a.sorted[j] >= n || (j > 0 && a.sorted[j] == a.sorted[j - 1]))
            fail(MH_ERR_INTERNAL, "generic kernel: operands out of range or repeated");
    const dim3 grid(grid_for(a.groups)), block(kThreads);
    switch (k) {
        case 1: hipLaunchKernelGGL((generic_gate_kernel<R, 1>), grid, block, 0, streams().main, a); break;
        case 2: hipLaunchKernelGGL((generic_gate_kernel<R, 2>), grid, block, 0, streams().main, a); break;
        case 3: hipLaunchKernelGGL((generic_gate_kernel<R, 3>), grid, block, 0, streams().main, a); break;
        case 4: hipLaunchKernelGGL((generic_gate_kernel<R, 4>), grid, block, 0, streams().main, a); break;
        case 5: hipLaunchKernelGGL((generic_gate_kernel<R, 5>), grid, block, 0, streams().main, a); break;
        default: hipLaunchKernelGGL((generic_gate_kernel<R, 6>), grid, block, 0, streams().main, a); break;
    }
    MH_HIP(hipGetLastError());
    synchronize();
}

// Blocks from the C ABI's flat arrays (block b: operands
// [operand_offsets[b], operand_offsets[b + 1]), its matrix at
// matrix_offsets[b] complex entries).
std::vector<Block> blocks_from_arrays(uint32_t num_blocks, const uint32_t* operand_offsets, const uint32_t* operands,
                                      const uint64_t* matrix_offsets, const double* matrices) {
    if (num_blocks > 0 && (operand_offsets == nullptr || operands == nullptr || matrix_offsets == nullptr ||
                           matrices == nullptr))
        fail(MH_ERR_INVALID, "null block arrays");
    std::vector<Block> blocks(num_blocks);
    const cplx* mats = reinterpret_cast<const cplx*>(matrices);
    for (uint32_t b = 0; b < num_blocks; ++b) {
        Block& blk = blocks[b];
        if (operand_offsets[b + 1] < operand_offsets[b] || operand_offsets[b + 1] - operand_offsets[b] > kMaxK)
            fail(MH_ERR_INVALID, "block " + std::to_string(b) + " has more than 6 operands");
        blk.operands.assign(operands + operand_offsets[b], operands + operand_offsets[b + 1]);
        const std::size_t dim = std::size_t{1} << blk.operands.size();
        blk.matrix.assign(mats + matrix_offsets[b], mats + matrix_offsets[b] + dim * dim);
        blk.diagonal = matrix_is_diagonal(blk.matrix.data(), dim);
    }
    return blocks;
}

void check_operands(uint32_t n, const uint32_t* operands, uint32_t k) {
    if (k == 0 || k > kMaxK) fail(MH_ERR_INVALID, "a gate needs 1 to 6 operands");
    if (k > n) fail(MH_ERR_INVALID, "more operands than qubits");
    uint64_t seen = 0;
    for (uint32_t j = 0; j < k; ++j) {
        if (operands[j] >= n) fail(MH_ERR_INVALID, "operand " + std::to_string(operands[j]) + " out of range");
        if ((seen >> operands[j]) & 1ull) fail(MH_ERR_INVALID, "repeated operand");
        seen |= 1ull << operands[j];
    }
}

}  // namespace

bool matrix_is_diagonal(const cplx* m, std::size_t dim) {
    for (std::size_t r = 0; r < dim; ++r)
        for (std::size_t c = 0; c < dim; ++c)
            if (r != c && (m[r * dim + c].real() != 0.0 || m[r * dim + c].imag() != 0.0)) return false;
    return true;
}

void launch_init_on(void* state, uint32_t n, int32_t precision, double re, double im, hipStream_t stream) {
    context();
    const uint64_t size = uint64_t{1} << n;
    if (precision == MH_C128)
        hipLaunchKernelGGL(init_kernel<double>, dim3(grid_for(size)), dim3(kThreads), 0, stream, state, size, re, im);
    else
        hipLaunchKernelGGL(init_kernel<float>, dim3(grid_for(size)), dim3(kThreads), 0, stream, state, size,
                           static_cast<float>(re), static_cast<float>(im));
    MH_HIP(hipGetLastError());
}

void launch_init(void* state, uint32_t n, int32_t precision, double re, double im) {
    launch_init_on(state, n, precision, re, im, streams().main);
}

void apply_blocks(void* state, uint32_t n, int32_t precision, const std::vector<Block>& blocks,
                  uint32_t window_qubits, bool time_passes, PassStats* stats, uint32_t max_cost) {
    const uint32_t wmax = window_width(precision);
    const uint32_t w = window_qubits == 0 ? wmax : window_qubits;
    if (w > wmax) fail(MH_ERR_INVALID, "window_qubits above " + std::to_string(wmax));
    if (w < 7) fail(MH_ERR_INVALID, "window_qubits below 7");
    for (const Block& b : blocks) check_operands(n, b.operands.data(), static_cast<uint32_t>(b.operands.size()));
    run_window_passes(state, n, precision, blocks, w, max_cost == 0 ? kDefaultPassCost : max_cost, time_passes, stats);
}

uint32_t window_default(int32_t precision) noexcept { return window_width(precision); }

uint32_t count_passes(uint32_t n, int32_t precision, const std::vector<Block>& blocks, uint32_t window_qubits,
                      uint32_t max_cost) {
    return static_cast<uint32_t>(plan_window_passes(n, precision, blocks, window_qubits == 0 ? window_width(precision) : window_qubits,
                                                    max_cost == 0 ? kDefaultPassCost : max_cost)
                                     .size());
}

void launch_generic(void* state, uint32_t n, int32_t precision, const cplx* matrix, const uint32_t* operands,
                    uint32_t k) {
    if (precision == MH_C128)
        run_generic<double>(state, n, matrix, operands, k);
    else if (mfma_applies(precision, n, k))
        run_mfma(state, n, matrix, operands, k);
    else
        run_generic<float>(state, n, matrix, operands, k);
}

void apply_gate(void* state, uint32_t n, int32_t precision, const cplx* matrix, const uint32_t* operands, uint32_t k,
                bool generic_only) {
    check_operands(n, operands, k);
    if (generic_only) {
        // The one-thread-per-group kernel itself, at either precision (the
        // independent path the others are checked against).
        if (precision == MH_C128)
            run_generic<double>(state, n, matrix, operands, k);
        else
            run_generic<float>(state, n, matrix, operands, k);
        return;
    }
    Block b;
    b.operands.assign(operands, operands + k);
    const std::size_t dim = std::size_t{1} << k;
    b.matrix.assign(matrix, matrix + dim * dim);
    b.diagonal = matrix_is_diagonal(matrix, dim);
    apply_blocks(state, n, precision, std::vector<Block>{b}, 0, false, nullptr, 0);
}

}  // namespace mh

using namespace mh;

extern "C" {

int mh_state_init(void* state, uint32_t num_qubits, int32_t precision, double phase_re, double phase_im) {
    return guarded([&] {
        amp_bytes(precision);
        if (num_qubits > 40) fail(MH_ERR_INVALID, "num_qubits above 40");
        launch_init(state, num_qubits, precision, phase_re, phase_im);
        synchronize();
    });
}

int mh_apply_matrix(void* state, uint32_t num_qubits, int32_t precision, const double* matrix,
                    const uint32_t* operands, uint32_t k) {
    return guarded([&] {
        amp_bytes(precision);
        apply_gate(state, num_qubits, precision, reinterpret_cast<const cplx*>(matrix), operands, k, false);
    });
}

int mh_apply_matrix_generic(void* state, uint32_t num_qubits, int32_t precision, const double* matrix,
                            const uint32_t* operands, uint32_t k) {
    return guarded([&] {
        amp_bytes(precision);
        apply_gate(state, num_qubits, precision, reinterpret_cast<const cplx*>(matrix), operands, k, true);
    });
}

int mh_apply_blocks(void* state, uint32_t num_qubits, int32_t precision, uint32_t num_blocks,
                    const uint32_t* operand_offsets, const uint32_t* operands, const uint64_t* matrix_offsets,
                    const double* matrices, uint32_t window_qubits, uint32_t* passes) {
    return guarded([&] {
        amp_bytes(precision);
        if (num_qubits > 40) fail(MH_ERR_INVALID, "num_qubits above 40");
        const std::vector<Block> blocks =
            blocks_from_arrays(num_blocks, operand_offsets, operands, matrix_offsets, matrices);
        PassStats stats;
        apply_blocks(state, num_qubits, precision, blocks, window_qubits, false, &stats, 0);
        if (passes) *passes = stats.passes;
    });
}

int mh_plan_passes(uint32_t num_qubits, int32_t precision, uint32_t num_blocks, const uint32_t* operand_offsets,
                   const uint32_t* operands, const uint64_t* matrix_offsets, const double* matrices,
                   uint32_t window_qubits, uint32_t max_pass_cost, uint32_t* pass_begin, uint32_t* num_passes) {
    return guarded([&] {
        amp_bytes(precision);
        if (num_qubits > 40) fail(MH_ERR_INVALID, "num_qubits above 40");
        if (pass_begin == nullptr || num_passes == nullptr) fail(MH_ERR_INVALID, "null output");
        const std::vector<Block> blocks =
            blocks_from_arrays(num_blocks, operand_offsets, operands, matrix_offsets, matrices);
        for (const Block& b : blocks) check_operands(num_qubits, b.operands.data(), static_cast<uint32_t>(b.operands.size()));
        const uint32_t wmax = window_width(precision);
        const uint32_t w = window_qubits == 0 ? wmax : window_qubits;
        if (w > wmax || w < 7) fail(MH_ERR_INVALID, "window_qubits must be 7.." + std::to_string(wmax));
        const std::vector<WindowPass> passes =
            plan_window_passes(num_qubits, precision, blocks, w, max_pass_cost == 0 ? kDefaultPassCost : max_pass_cost);
        // Pass p covers blocks [pass_begin[p], pass_begin[p + 1]); a block the
        // planner skips (no operands) stays in the pass before it.
        for (std::size_t p = 0; p < passes.size(); ++p) pass_begin[p] = p == 0 ? 0 : passes[p].blocks.front();
        pass_begin[passes.size()] = num_blocks;
        *num_passes = static_cast<uint32_t>(passes.size());
    });
}

int mh_gather(const void* state, uint32_t num_qubits, int32_t precision, const uint64_t* indices, uint64_t count,
              void* out) {
    return guarded([&] {
        const std::size_t amp = amp_bytes(precision);
        if (num_qubits > 40) fail(MH_ERR_INVALID, "num_qubits above 40");
        if (count == 0) return;
        if (state == nullptr || indices == nullptr || out == nullptr) fail(MH_ERR_INVALID, "null argument");
        const uint64_t size = uint64_t{1} << num_qubits;
        for (uint64_t i = 0; i < count; ++i)
            if (indices[i] >= size) fail(MH_ERR_INVALID, "gather index " + std::to_string(indices[i]) + " beyond the state");
        context();
        DeviceBuffer didx(count * sizeof(uint64_t)), dout(count * amp);
        upload(didx.get(), indices, count * sizeof(uint64_t));
        const dim3 grid(grid_for(count)), block(kThreads);
        if (precision == MH_C128)
            hipLaunchKernelGGL(gather_kernel<double2>, grid, block, 0, streams().main, static_cast<const double2*>(state),
                               didx.as<uint64_t>(), count, dout.as<double2>());
        else
            hipLaunchKernelGGL(gather_kernel<float2>, grid, block, 0, streams().main, static_cast<const float2*>(state),
                               didx.as<uint64_t>(), count, dout.as<float2>());
        MH_HIP(hipGetLastError());
        download(out, dout.get(), count * amp);
    });
}

}  // extern "C"
