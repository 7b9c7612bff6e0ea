// Gate application kernels for gfx950.
//
// Two kernels do all gate work:
//
//  * window_pass_kernel -- one HBM read and one HBM write of the state for a
//    whole run of consecutive plan blocks. Each workgroup stages a window of
//    2^W amplitudes in LDS: the low qubits 0..L-1 (contiguous runs of 2^L
//    amplitudes, so every global load and store is coalesced) plus up to W-L
//    higher qubits (strided runs). Every dense block of the run acts on window
//    qubits only; diagonal blocks may act anywhere (their entry depends only
//    on the amplitude's own index). Blocks are applied in plan order with a
//    workgroup barrier between them.
//
//  * generic_gate_kernel -- the reference's one-thread-per-group kernel
//    (reference src/kernels/cuda/dense_kernels.cu:63-107), kept as the
//    independent generic path the window kernel is tested against.
//
// Arithmetic (complex128): every output amplitude starts at (+0, +0) and adds
// M[out][in] * a[in] for in = 0 .. 2^k - 1 in increasing order; a complex
// product is (mr*ar - mi*ai, mr*ai + mi*ar); each multiply and add is rounded
// on its own (this file is compiled with -ffp-contract=off, so no FMA is
// formed). That is the reference CPU generic kernel's arithmetic
// (dense_kernels.cpp:465-524), so results are bitwise equal. A diagonal
// block computes (+0,+0) + d[local] * a, which equals the generic sum bit for
// bit (reference dense_kernels.cpp:299-313).
//
// complex64 stores float2 amplitudes and accumulates with FP32 FMA.
#include <algorithm>
#include <cstring>

#include "mh_internal.h"

namespace mh {
namespace {

constexpr int kThreads = 256;
constexpr uint32_t kMaxWindowC128 = 12;  // 2^12 * 16 B = 64 KiB of LDS
constexpr uint32_t kMaxWindowC64 = 13;   // 2^13 * 8 B = 64 KiB of LDS
constexpr uint32_t kMaxK = 6;
constexpr uint32_t kMaxHigh = 10;        // window qubits above the low run

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

// ---- a pass program

// One block inside a pass.
struct OpDesc {
    uint32_t kind;        // 0 dense on window bits, 1 diagonal
    uint32_t k;
    uint32_t pos[kMaxK];  // dense: window-local bit of operand j; diagonal: global qubit of operand j
    uint32_t sorted[kMaxK];  // dense: window-local bits ascending
    uint64_t mat;         // offset of the matrix (dense, 4^k entries) or diagonal (2^k entries)
};

struct PassArgs {
    void* state;
    const OpDesc* ops;
    const void* mats;      // double2 (c128) or float2 (c64) entries
    uint64_t window_mask;  // global qubits in the window
    uint32_t n;
    uint32_t w;            // window qubits
    uint32_t low;          // qubits 0..low-1 are in the window
    uint32_t num_ops;
    uint32_t mat_slots;    // LDS entries reserved for a staged k >= 3 matrix
    uint32_t high_q[kMaxHigh];  // global qubit of window bit low + j
};

// The groups of one dense block inside the LDS window. `m` points at the
// block's matrix: global memory (scalar loads) for k <= 2, an LDS copy for
// k >= 3, so a wide matrix is never hoisted into vector registers.
template <typename R, uint32_t K>
__device__ __forceinline__ void dense_groups(typename Vec<R>::type* lds, const OpDesc& op,
                                             const typename Vec<R>::type* m, uint32_t w) {
    using V = typename Vec<R>::type;
    constexpr uint32_t dim = 1u << K;
    const uint32_t groups = (1u << w) >> K;
    uint32_t bit[K];
    uint32_t sorted[K];
#pragma unroll
    for (uint32_t j = 0; j < K; ++j) {
        bit[j] = 1u << op.pos[j];
        sorted[j] = op.sorted[j];
    }
    for (uint32_t g = threadIdx.x; g < groups; g += kThreads) {
        uint32_t base = g;
#pragma unroll
        for (uint32_t j = 0; j < K; ++j) {
            const uint32_t low = (1u << sorted[j]) - 1u;
            base = ((base & ~low) << 1) | (base & low);
        }
        V a[dim];
#pragma unroll
        for (uint32_t l = 0; l < dim; ++l) {
            uint32_t idx = base;
#pragma unroll
            for (uint32_t j = 0; j < K; ++j) idx |= ((l >> j) & 1u) ? bit[j] : 0u;
            a[l] = lds[idx];
        }
        // Rows one at a time for k >= 3: a whole unrolled matrix would be
        // hoisted into registers.
#pragma unroll(K <= 2 ? dim : 1)
        for (uint32_t o = 0; o < dim; ++o) {
            R sr = R(0), si = R(0);
#pragma unroll
            for (uint32_t in = 0; in < dim; ++in) {
                const V e = m[o * dim + in];
                cmac(sr, si, e.x, e.y, a[in].x, a[in].y);
            }
            uint32_t idx = base;
#pragma unroll
            for (uint32_t j = 0; j < K; ++j) idx |= ((o >> j) & 1u) ? bit[j] : 0u;
            V r;
            r.x = sr;
            r.y = si;
            lds[idx] = r;
        }
    }
}

// KMAX bounds the widest dense block the kernel instance handles, so a pass
// of 1- and 2-qubit blocks does not pay the registers of a 6-qubit one.
template <typename R, uint32_t KMAX>
__device__ __forceinline__ void dense_op(typename Vec<R>::type* lds, const OpDesc& op,
                                         const typename Vec<R>::type* m, uint32_t w) {
    if (op.k == 1) {
        dense_groups<R, 1>(lds, op, m, w);
    } else if (op.k == 2 || KMAX < 3) {
        if constexpr (KMAX >= 2) dense_groups<R, 2>(lds, op, m, w);
    } else if (op.k == 3) {
        if constexpr (KMAX >= 3) dense_groups<R, 3>(lds, op, m, w);
    } else if (op.k == 4 || KMAX < 5) {
        if constexpr (KMAX >= 4) dense_groups<R, 4>(lds, op, m, w);
    } else if (op.k == 5) {
        if constexpr (KMAX >= 5) dense_groups<R, 5>(lds, op, m, w);
    } else {
        if constexpr (KMAX >= 6) dense_groups<R, 6>(lds, op, m, w);
    }
}

// The non-window bits of `o` deposited at the non-window qubit positions.
__device__ __forceinline__ uint64_t outer_base(uint64_t o, uint64_t window_mask, uint32_t n) {
    uint64_t base = 0;
    uint32_t src = 0;
    for (uint32_t p = 0; p < n; ++p) {
        if ((window_mask >> p) & 1ull) continue;
        base |= ((o >> src) & 1ull) << p;
        ++src;
    }
    return base;
}

template <typename R, uint32_t KMAX>
__global__ void __launch_bounds__(kThreads) window_pass_kernel(PassArgs args) {
    using V = typename Vec<R>::type;
    extern __shared__ unsigned char lds_raw[];
    const uint32_t w = args.w;
    const uint32_t low = args.low;
    const uint32_t high = w - low;
    // LDS: the window's amplitudes, then the staged matrix (both 16-byte
    // aligned for ds_read/write_b128), then the offset table.
    V* lds = reinterpret_cast<V*>(lds_raw);
    V* lds_mat = lds + (1u << w);
    uint64_t* high_off = reinterpret_cast<uint64_t*>(lds_mat + args.mat_slots);

    // Offsets of the high window bits: entry h is sum_j bit_j(h) << high_q[j].
    for (uint32_t h = threadIdx.x; h < (1u << high); h += kThreads) {
        uint64_t off = 0;
        for (uint32_t j = 0; j < high; ++j) off |= uint64_t((h >> j) & 1u) << args.high_q[j];
        high_off[h] = off;
    }
    const uint64_t base = outer_base(blockIdx.x, args.window_mask, args.n);
    V* state = static_cast<V*>(args.state);
    const uint32_t lowmask = (1u << low) - 1u;
    __syncthreads();

    const uint32_t size = 1u << w;
    for (uint32_t l = threadIdx.x; l < size; l += kThreads)
        lds[l] = state[base + (l & lowmask) + high_off[l >> low]];
    __syncthreads();

    const V* mats = static_cast<const V*>(args.mats);
    for (uint32_t i = 0; i < args.num_ops; ++i) {
        const OpDesc op = args.ops[i];
        if (op.kind == 0) {
            if (KMAX >= 3 && op.k >= 3) {
                const uint32_t entries = 1u << (2 * op.k);
                for (uint32_t e = threadIdx.x; e < entries; e += kThreads) lds_mat[e] = mats[op.mat + e];
                __syncthreads();
                dense_op<R, KMAX>(lds, op, lds_mat, w);
            } else {
                dense_op<R, KMAX>(lds, op, mats + op.mat, w);
            }
        } else {
            const V* d = mats + op.mat;
            for (uint32_t l = threadIdx.x; l < size; l += kThreads) {
                const uint64_t gi = base + (l & lowmask) + high_off[l >> low];
                uint32_t local = 0;
#pragma unroll
                for (uint32_t j = 0; j < kMaxK; ++j)
                    if (j < op.k) local |= uint32_t((gi >> op.pos[j]) & 1ull) << j;
                const V e = d[local];
                const V a = lds[l];
                R sr = R(0), si = R(0);
                cmac(sr, si, e.x, e.y, a.x, a.y);
                V r;
                r.x = sr;
                r.y = si;
                lds[l] = r;
            }
        }
        __syncthreads();
    }

    for (uint32_t l = threadIdx.x; l < size; l += kThreads)
        state[base + (l & lowmask) + high_off[l >> low]] = lds[l];
}

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
    uint64_t base = g;
#pragma unroll
    for (uint32_t j = 0; j < K; ++j) {
        const uint64_t lowm = (1ull << a.sorted[j]) - 1ull;
        base = ((base & ~lowm) << 1) | (base & lowm);
    }
    V v[dim];
#pragma unroll
    for (uint32_t l = 0; l < dim; ++l) {
        uint64_t idx = base;
#pragma unroll
        for (uint32_t j = 0; j < K; ++j) idx |= ((l >> j) & 1u) ? (1ull << a.ops[j]) : 0ull;
        v[l] = s[idx];
    }
#pragma unroll(K <= 2 ? dim : 1)
    for (uint32_t o = 0; o < dim; ++o) {
        R sr = R(0), si = R(0);
#pragma unroll
        for (uint32_t in = 0; in < dim; ++in) {
            const V e = m[o * dim + in];
            cmac(sr, si, e.x, e.y, v[in].x, v[in].y);
        }
        uint64_t idx = base;
#pragma unroll
        for (uint32_t j = 0; j < K; ++j) idx |= ((o >> j) & 1u) ? (1ull << a.ops[j]) : 0ull;
        V r;
        r.x = sr;
        r.y = si;
        s[idx] = r;
    }
}

// One thread per group of 2^K amplitudes; the matrix is staged in LDS.
template <typename R, uint32_t K>
__global__ void __launch_bounds__(kThreads) generic_gate_kernel(GenericArgs a) {
    using V = typename Vec<R>::type;
    extern __shared__ unsigned char lds_raw[];
    V* m = reinterpret_cast<V*>(lds_raw);
    const V* src = static_cast<const V*>(a.matrix);
    for (uint32_t e = threadIdx.x; e < (1u << (2 * K)); e += kThreads) m[e] = src[e];
    __syncthreads();
    const uint64_t stride = uint64_t(gridDim.x) * kThreads;
    for (uint64_t g = uint64_t(blockIdx.x) * kThreads + threadIdx.x; g < a.groups; g += stride)
        generic_group<R, K>(a, m, g);
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

uint32_t grid_for(uint64_t items) {
    const uint64_t blocks = (items + kThreads - 1) / kThreads;
    return static_cast<uint32_t>(std::min<uint64_t>(std::max<uint64_t>(blocks, 1), uint64_t{1} << 20));
}

// The widest window whose amplitudes and offset table fit the device's LDS.
uint32_t max_window(int32_t precision) {
    const std::size_t amp = precision == MH_C128 ? 16 : 8;
    uint32_t w = precision == MH_C128 ? kMaxWindowC128 : kMaxWindowC64;
    const std::size_t limit = context().props.sharedMemPerBlock;
    while (w > 6 && (amp << w) + (sizeof(uint64_t) << (w - 5)) > limit) --w;
    return w;
}

// ---- the pass schedule

struct Pass {
    uint64_t window_mask = 0;  // dense operand qubits (the low run is added at launch)
    uint32_t kmax = 0;         // widest dense block
    std::vector<uint32_t> blocks;
};

// The low run: qubits 0..low-1 always sit in the window so global accesses are
// runs of 2^low amplitudes (512 B at c128 for low = 5).
uint32_t low_run(uint32_t n, uint32_t w) { return std::min(std::min(n, w), 5u); }

uint64_t mask_of(const std::vector<uint32_t>& q) {
    uint64_t m = 0;
    for (uint32_t x : q) m |= 1ull << x;
    return m;
}

// Greedy: consecutive blocks join the current pass while the union of dense
// operands and the low run fits the window; a diagonal block always joins.
std::vector<Pass> schedule(uint32_t n, const std::vector<Block>& blocks, uint32_t w) {
    const uint32_t low = low_run(n, w);
    const uint64_t low_mask = low >= 64 ? ~0ull : ((1ull << low) - 1ull);
    std::vector<Pass> passes;
    Pass cur;
    bool open = false;
    for (uint32_t b = 0; b < blocks.size(); ++b) {
        const Block& blk = blocks[b];
        if (blk.operands.empty()) continue;
        if (blk.diagonal) {
            if (!open) { cur = Pass{}; open = true; }
            cur.blocks.push_back(b);
            continue;
        }
        const uint32_t k = static_cast<uint32_t>(blk.operands.size());
        const uint64_t want = cur.window_mask | mask_of(blk.operands);
        if (open && __builtin_popcountll(want | low_mask) <= static_cast<int>(w)) {
            cur.window_mask = want;
            cur.kmax = std::max(cur.kmax, k);
            cur.blocks.push_back(b);
            continue;
        }
        if (open) passes.push_back(cur);
        if (__builtin_popcountll(mask_of(blk.operands) | low_mask) > static_cast<int>(w))
            fail(MH_ERR_INVALID, "a dense block does not fit the window beside the low run");
        cur = Pass{};
        cur.window_mask = mask_of(blk.operands);
        cur.kmax = k;
        cur.blocks.push_back(b);
        open = true;
    }
    if (open) passes.push_back(cur);
    return passes;
}

// Lets each window kernel instance take up to the device's per-workgroup LDS.
template <typename R, uint32_t KMAX>
void allow_lds() {
    static const bool done = [] {
        Context& ctx = context();
        MH_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&window_pass_kernel<R, KMAX>),
                                   hipFuncAttributeMaxDynamicSharedMemorySize,
                                   static_cast<int>(ctx.props.sharedMemPerBlock)));
        return true;
    }();
    (void)done;
}

template <typename R>
void run_passes(void* state, uint32_t n, const std::vector<Block>& blocks, uint32_t w, bool time_passes,
                PassStats* stats) {
    using V = typename Vec<R>::type;
    Context& ctx = context();
    const std::vector<Pass> passes = schedule(n, blocks, w);
    if (stats) stats->passes = static_cast<uint32_t>(passes.size());
    if (passes.empty()) return;

    // One program for the whole call: op descriptors and matrices.
    std::vector<OpDesc> ops;
    std::vector<V> mats;
    std::vector<uint32_t> op_begin;
    std::vector<PassArgs> args(passes.size());
    const uint32_t low = low_run(n, w);
    for (std::size_t p = 0; p < passes.size(); ++p) {
        const Pass& pass = passes[p];
        uint64_t wmask = pass.window_mask | ((1ull << low) - 1ull);
        // Fill the window with further low qubits so it always holds w qubits.
        for (uint32_t q = 0; q < n && static_cast<uint32_t>(__builtin_popcountll(wmask)) < w; ++q) wmask |= 1ull << q;
        std::vector<uint32_t> wq;
        for (uint32_t q = 0; q < n; ++q)
            if ((wmask >> q) & 1ull) wq.push_back(q);
        uint32_t run = 0;
        while (run < wq.size() && wq[run] == run) ++run;
        PassArgs& a = args[p];
        std::memset(&a, 0, sizeof(a));
        a.n = n;
        a.w = static_cast<uint32_t>(wq.size());
        a.low = run;
        a.window_mask = wmask;
        if (a.w - a.low > kMaxHigh) fail(MH_ERR_INTERNAL, "window schedule exceeds the high-bit table");
        for (uint32_t j = a.low; j < a.w; ++j) a.high_q[j - a.low] = wq[j];
        op_begin.push_back(static_cast<uint32_t>(ops.size()));
        for (uint32_t b : pass.blocks) {
            const Block& blk = blocks[b];
            OpDesc d{};
            d.k = static_cast<uint32_t>(blk.operands.size());
            if (d.k > kMaxK) fail(MH_ERR_INVALID, "a block wider than 6 qubits");
            d.mat = mats.size();
            const std::size_t dim = std::size_t{1} << d.k;
            if (blk.diagonal) {
                d.kind = 1;
                for (uint32_t j = 0; j < d.k; ++j) d.pos[j] = blk.operands[j];
                for (std::size_t r = 0; r < dim; ++r) {
                    V e;
                    e.x = static_cast<R>(blk.matrix[r * dim + r].real());
                    e.y = static_cast<R>(blk.matrix[r * dim + r].imag());
                    mats.push_back(e);
                }
            } else {
                d.kind = 0;
                for (uint32_t j = 0; j < d.k; ++j) {
                    const auto it = std::find(wq.begin(), wq.end(), blk.operands[j]);
                    if (it == wq.end()) fail(MH_ERR_INTERNAL, "a dense operand outside its window");
                    d.pos[j] = static_cast<uint32_t>(it - wq.begin());
                    d.sorted[j] = d.pos[j];
                }
                std::sort(d.sorted, d.sorted + d.k);
                for (std::size_t e = 0; e < dim * dim; ++e) {
                    V v;
                    v.x = static_cast<R>(blk.matrix[e].real());
                    v.y = static_cast<R>(blk.matrix[e].imag());
                    mats.push_back(v);
                }
            }
            ops.push_back(d);
        }
        a.num_ops = static_cast<uint32_t>(pass.blocks.size());
    }
    DeviceBuffer dops(ops.size() * sizeof(OpDesc));
    DeviceBuffer dmats(std::max<std::size_t>(mats.size(), 1) * sizeof(V));
    upload(dops.get(), ops.data(), ops.size() * sizeof(OpDesc));
    upload(dmats.get(), mats.data(), mats.size() * sizeof(V));

    std::vector<hipEvent_t> events;
    if (time_passes) {
        events.resize(passes.size() + 1);
        for (auto& e : events) MH_HIP(hipEventCreate(&e));
        MH_HIP(hipEventRecord(events[0], ctx.stream));
    }
    for (std::size_t p = 0; p < passes.size(); ++p) {
        PassArgs a = args[p];
        a.state = state;
        a.ops = dops.as<OpDesc>() + op_begin[p];
        a.mats = dmats.get();
        const uint32_t kmat = passes[p].kmax >= 3 ? passes[p].kmax : 0;
        a.mat_slots = kmat ? (1u << (2 * kmat)) : 0u;
        const std::size_t lds = (sizeof(V) << a.w) + sizeof(V) * a.mat_slots + (sizeof(uint64_t) << (a.w - a.low));
        const uint32_t grid = 1u << (n - a.w);
        const uint32_t kmax = passes[p].kmax;
        if (lds > ctx.props.sharedMemPerBlock)
            fail(MH_ERR_INVALID, "a pass needs " + std::to_string(lds) + " bytes of LDS; the device has " +
                                     std::to_string(ctx.props.sharedMemPerBlock));
        if (kmax <= 2) allow_lds<R, 2>();
        else if (kmax <= 4) allow_lds<R, 4>();
        else allow_lds<R, 6>();
        if (kmax <= 2)
            hipLaunchKernelGGL((window_pass_kernel<R, 2>), dim3(grid), dim3(kThreads), lds, ctx.stream, a);
        else if (kmax <= 4)
            hipLaunchKernelGGL((window_pass_kernel<R, 4>), dim3(grid), dim3(kThreads), lds, ctx.stream, a);
        else
            hipLaunchKernelGGL((window_pass_kernel<R, 6>), dim3(grid), dim3(kThreads), lds, ctx.stream, a);
        MH_HIP(hipGetLastError());
        if (time_passes) MH_HIP(hipEventRecord(events[p + 1], ctx.stream));
    }
    synchronize();
    if (time_passes) {
        stats->pass_ms.resize(passes.size());
        for (std::size_t p = 0; p < passes.size(); ++p)
            MH_HIP(hipEventElapsedTime(&stats->pass_ms[p], events[p], events[p + 1]));
        for (auto& e : events) (void)hipEventDestroy(e);
    }
}

template <typename R>
void run_generic(void* state, uint32_t n, const cplx* matrix, const uint32_t* operands, uint32_t k) {
    using V = typename Vec<R>::type;
    Context& ctx = context();
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
    a.groups = (uint64_t{1} << n) >> k;
    a.k = k;
    for (uint32_t j = 0; j < k; ++j) a.ops[j] = a.sorted[j] = operands[j];
    std::sort(a.sorted, a.sorted + k);
    const std::size_t lds = sizeof(V) << (2 * k);
    const dim3 grid(grid_for(a.groups)), block(kThreads);
    switch (k) {
        case 1: hipLaunchKernelGGL((generic_gate_kernel<R, 1>), grid, block, lds, ctx.stream, a); break;
        case 2: hipLaunchKernelGGL((generic_gate_kernel<R, 2>), grid, block, lds, ctx.stream, a); break;
        case 3: hipLaunchKernelGGL((generic_gate_kernel<R, 3>), grid, block, lds, ctx.stream, a); break;
        case 4: hipLaunchKernelGGL((generic_gate_kernel<R, 4>), grid, block, lds, ctx.stream, a); break;
        case 5: hipLaunchKernelGGL((generic_gate_kernel<R, 5>), grid, block, lds, ctx.stream, a); break;
        default: hipLaunchKernelGGL((generic_gate_kernel<R, 6>), grid, block, lds, ctx.stream, a); break;
    }
    MH_HIP(hipGetLastError());
    synchronize();
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

void launch_init(void* state, uint32_t n, int32_t precision, double re, double im) {
    Context& ctx = context();
    const uint64_t size = uint64_t{1} << n;
    if (precision == MH_C128)
        hipLaunchKernelGGL(init_kernel<double>, dim3(grid_for(size)), dim3(kThreads), 0, ctx.stream, state, size, re, im);
    else
        hipLaunchKernelGGL(init_kernel<float>, dim3(grid_for(size)), dim3(kThreads), 0, ctx.stream, state, size,
                           static_cast<float>(re), static_cast<float>(im));
    MH_HIP(hipGetLastError());
}

void apply_blocks(void* state, uint32_t n, int32_t precision, const std::vector<Block>& blocks,
                  uint32_t window_qubits, bool time_passes, PassStats* stats) {
    const uint32_t wmax = max_window(precision);
    uint32_t w = window_qubits == 0 ? wmax : window_qubits;
    if (w > wmax) fail(MH_ERR_INVALID, "window_qubits above " + std::to_string(wmax));
    if (window_qubits != 0 && window_qubits < 7) fail(MH_ERR_INVALID, "window_qubits below 7");
    w = std::min(w, n);
    for (const Block& b : blocks) {
        if (b.operands.size() > w && !b.diagonal) fail(MH_ERR_INVALID, "a dense block wider than the window");
        check_operands(n, b.operands.data(), static_cast<uint32_t>(b.operands.size()));
    }
    if (precision == MH_C128)
        run_passes<double>(state, n, blocks, w, time_passes, stats);
    else
        run_passes<float>(state, n, blocks, w, time_passes, stats);
}

uint32_t window_default(int32_t precision) noexcept {
    try {
        return max_window(precision);
    } catch (...) {
        return precision == MH_C128 ? kMaxWindowC128 : kMaxWindowC64;
    }
}

uint32_t count_passes(uint32_t n, const std::vector<Block>& blocks, uint32_t window_qubits) {
    return static_cast<uint32_t>(schedule(n, blocks, std::min(window_qubits, n)).size());
}

void apply_gate(void* state, uint32_t n, int32_t precision, const cplx* matrix, const uint32_t* operands, uint32_t k,
                bool generic_only) {
    check_operands(n, operands, k);
    if (generic_only) {
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
    apply_blocks(state, n, precision, std::vector<Block>{b}, 0, false, nullptr);
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
        std::vector<Block> blocks(num_blocks);
        const cplx* mats = reinterpret_cast<const cplx*>(matrices);
        for (uint32_t b = 0; b < num_blocks; ++b) {
            Block& blk = blocks[b];
            blk.operands.assign(operands + operand_offsets[b], operands + operand_offsets[b + 1]);
            const std::size_t dim = std::size_t{1} << blk.operands.size();
            blk.matrix.assign(mats + matrix_offsets[b], mats + matrix_offsets[b] + dim * dim);
            blk.diagonal = matrix_is_diagonal(blk.matrix.data(), dim);
        }
        PassStats stats;
        apply_blocks(state, num_qubits, precision, blocks, window_qubits, false, &stats);
        if (passes) *passes = stats.passes;
    });
}

}  // extern "C"
