// Exact shot sampling on the device.
//
// The reference sampler (src/engines/dense/sample.cpp:24-75) walks the basis in
// index order with `c = fl(c + p_i)`, p_i = x_i*x_i + y_i*y_i (std::norm, no
// FMA), and sends each sorted uniform u to the first basis with u < c (the last
// basis if none). Counts must be identical, so the device reproduces that
// rounding chain bit for bit, in parallel:
//
//  * While c stays in one binade it is an integer multiple K of ulp = 2^s
//    (s = max(exponent(c) - 52, -1074)), and fl(K*2^s + p) = (K + r)*2^s with
//    r = p / 2^s rounded to the nearest integer, ties going to the r that
//    makes K + r even. So a run of steps is an integer sum whose only state is
//    the parity of K: each element is a function of that parity, and these
//    functions compose associatively (exact integer arithmetic, any order).
//  * The binade each block of 2^13 elements starts in is predicted from an
//    FP64 tree prefix of block sums. The chain then runs serially over blocks
//    (and over groups of blocks), applying a composed transfer only when the
//    exact running value is in the predicted binade and the result stays
//    below the binade's top; otherwise (the block where c crosses a power of
//    two, or a mispredicted one) it walks that block with plain FP64 adds,
//    sub-block by sub-block. The result is the exact c at every block start.
//  * Each shot binary-searches the block starts, then walks its block with
//    plain FP64 adds from the exact start value.
//
// Prediction only decides speed: every transfer is checked against the exact
// running value before it is used.
#include <algorithm>
#include <climits>
#include <numeric>

#include "mh_internal.h"

namespace mh {
namespace {

constexpr uint32_t kBlockLog = 13;           // elements per block
constexpr uint32_t kGroup = 256;             // blocks per group
constexpr uint32_t kSub = 256;               // elements per sub-block of a dirty walk
constexpr int kThreads = 256;
constexpr uint64_t kSat = uint64_t{1} << 62;  // saturation: "leaves the binade"
constexpr uint64_t kTop = (uint64_t{1} << 53) - 1;
constexpr int32_t kNonUniform = INT_MIN;

struct Xfer {
    uint64_t d0, d1;  // increment of K for input parity 0 / 1
    int32_t s;        // the binade's ulp exponent (kNonUniform: blocks disagree)
    uint32_t q;       // bit 0: output parity for input 0; bit 1: for input 1
};

__device__ __forceinline__ uint64_t sat_add(uint64_t a, uint64_t b) {
    const uint64_t c = a + b;
    return (c < a || c > kSat) ? kSat : c;
}

__device__ __forceinline__ Xfer identity_xfer(int32_t s) { return Xfer{0, 0, s, 0b10u}; }

// f then g.
__device__ __forceinline__ Xfer compose(const Xfer& f, const Xfer& g) {
    const uint32_t f0 = f.q & 1u, f1 = (f.q >> 1) & 1u;
    Xfer h;
    h.d0 = sat_add(f.d0, f0 ? g.d1 : g.d0);
    h.d1 = sat_add(f.d1, f1 ? g.d1 : g.d0);
    const uint32_t g0 = g.q & 1u, g1 = (g.q >> 1) & 1u;
    h.q = (f0 ? g1 : g0) | ((f1 ? g1 : g0) << 1);
    h.s = f.s == g.s ? f.s : kNonUniform;
    return h;
}

// The ulp exponent of the binade holding c >= 0.
__device__ __forceinline__ int32_t scale_of(double c) {
    const uint64_t bits = __double_as_longlong(c);
    const int32_t e = static_cast<int32_t>((bits >> 52) & 0x7FF);
    if (e <= 1) return -1074;  // subnormals and [2^-1022, 2^-1021) share ulp 2^-1074
    return e - 1075;
}

// c = K * 2^s exactly.
__device__ __forceinline__ uint64_t k_of(double c, int32_t s) {
    const uint64_t bits = __double_as_longlong(c);
    const uint64_t e = (bits >> 52) & 0x7FF;
    const uint64_t frac = bits & ((uint64_t{1} << 52) - 1);
    if (e == 0) return frac;                      // s == -1074
    const uint64_t m = frac | (uint64_t{1} << 52);
    return m << ((static_cast<int32_t>(e) - 1075) - s);  // 0 when normal binade; 1 for [2^-1022, 2^-1021)
}

__device__ __forceinline__ double value_of(uint64_t k, int32_t s) { return ldexp(static_cast<double>(k), s); }

// The step function of adding p at ulp 2^s.
__device__ __forceinline__ Xfer element_xfer(double p, int32_t s) {
    Xfer x = identity_xfer(s);
    if (p == 0.0) return x;
    const uint64_t bits = __double_as_longlong(p);
    const int32_t e = static_cast<int32_t>((bits >> 52) & 0x7FF);
    const uint64_t frac = bits & ((uint64_t{1} << 52) - 1);
    const uint64_t m = e == 0 ? frac : (frac | (uint64_t{1} << 52));
    const int32_t q = e == 0 ? -1074 : e - 1075;
    const int32_t shift = s - q;
    uint64_t r;
    bool tie = false;
    if (shift <= 0) {
        r = (-shift > 9) ? kSat : (m << (-shift));
    } else if (shift >= 64) {
        r = 0;
    } else {
        r = m >> shift;
        const uint64_t rem = m & ((uint64_t{1} << shift) - 1);
        const uint64_t half = uint64_t{1} << (shift - 1);
        if (rem > half) r += 1;
        tie = rem == half;
    }
    // Input parity P: d = r (+1 on a tie when P + r is odd); output parity (P + d) & 1.
    const uint64_t d0 = r + ((tie && (r & 1u)) ? 1u : 0u);
    const uint64_t d1 = r + ((tie && !(r & 1u)) ? 1u : 0u);
    x.d0 = d0;
    x.d1 = d1;
    x.q = static_cast<uint32_t>(d0 & 1u) | (static_cast<uint32_t>((1u + d1) & 1u) << 1);
    return x;
}

template <typename V>
__device__ __forceinline__ double prob(const V* s, uint64_t i) {
    const V v = s[i];
    const double x = static_cast<double>(v.x);
    const double y = static_cast<double>(v.y);
    return x * x + y * y;
}

struct SampleArgs {
    const void* state;
    uint64_t size;        // 2^n
    uint64_t block;       // elements per block
    uint32_t blocks;
    uint32_t groups;
    double* sums;         // [blocks]
    double* approx;       // [blocks + 1] exclusive prefix of sums
    Xfer* xfer;           // [blocks]
    Xfer* gxfer;          // [groups]
    double* cstart;       // [blocks + 1] exact c before each block
    uint32_t* jumped;     // [groups]
    // The exact running value before the first element (device; null: 0),
    // for a state that arrives in consecutive chunks.
    const double* carry;
};

__device__ __forceinline__ double carry_of(const SampleArgs& a) { return a.carry ? *a.carry : 0.0; }

template <typename V>
__global__ void __launch_bounds__(kThreads) block_sums_kernel(SampleArgs a) {
    __shared__ double red[kThreads];
    const V* s = static_cast<const V*>(a.state);
    const uint64_t b0 = uint64_t(blockIdx.x) * a.block;
    double acc = 0.0;
    for (uint64_t i = threadIdx.x; i < a.block; i += kThreads) acc = acc + prob(s, b0 + i);
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int st = kThreads / 2; st > 0; st >>= 1) {
        if (threadIdx.x < st) red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) a.sums[blockIdx.x] = red[0];
}

// One workgroup: an exclusive FP64 prefix of the block sums (a prediction only).
__global__ void __launch_bounds__(1024) approx_prefix_kernel(SampleArgs a) {
    __shared__ double part[1024];
    const uint32_t per = (a.blocks + 1023) / 1024;
    const uint32_t b0 = min(a.blocks, threadIdx.x * per);
    const uint32_t b1 = min(a.blocks, b0 + per);
    double acc = 0.0;
    for (uint32_t b = b0; b < b1; ++b) acc = acc + a.sums[b];
    part[threadIdx.x] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double run = carry_of(a);
        for (int t = 0; t < 1024; ++t) {
            const double v = part[t];
            part[t] = run;
            run = run + v;
        }
    }
    __syncthreads();
    double run = part[threadIdx.x];
    for (uint32_t b = b0; b < b1; ++b) {
        a.approx[b] = run;
        run = run + a.sums[b];
    }
    if (threadIdx.x == 1023) a.approx[a.blocks] = run;
}

// Amplitudes of a block staged in LDS at a time by block_xfer_kernel: 16 KiB,
// so that several workgroups share a CU (a whole block of 2^13, 64 KiB, left
// room for two and the kernel latency-bound).
constexpr uint32_t kXferStage = 2048;

template <typename V>
__global__ void __launch_bounds__(kThreads) block_xfer_kernel(SampleArgs a) {
    __shared__ double p_sh[kXferStage];
    __shared__ Xfer red[kThreads];
    const V* st = static_cast<const V*>(a.state);
    const uint32_t b = blockIdx.x;
    const int32_t s = scale_of(a.approx[b]);
    const uint64_t i0 = uint64_t(b) * a.block;
    const uint32_t stage = static_cast<uint32_t>(min(a.block, uint64_t{kXferStage}));
    Xfer total = identity_xfer(s);  // thread 0's: the stages composed in order
    for (uint64_t base = 0; base < a.block; base += stage) {
        for (uint32_t i = threadIdx.x; i < stage; i += kThreads) p_sh[i] = prob(st, i0 + base + i);
        __syncthreads();
        // Each thread composes a contiguous run in order; then a fixed tree
        // (the composition is exact, so the order only has to respect the
        // sequence). The tree's barriers also end every read of p_sh and red
        // before the next stage writes them.
        const uint32_t per = (stage + kThreads - 1) / kThreads;
        const uint32_t j0 = min(stage, threadIdx.x * per);
        const uint32_t j1 = min(stage, j0 + per);
        Xfer acc = identity_xfer(s);
        for (uint32_t j = j0; j < j1; ++j) acc = compose(acc, element_xfer(p_sh[j], s));
        red[threadIdx.x] = acc;
        __syncthreads();
        for (int stride = 1; stride < kThreads; stride <<= 1) {
            if ((threadIdx.x % (2 * stride)) == 0) red[threadIdx.x] = compose(red[threadIdx.x], red[threadIdx.x + stride]);
            __syncthreads();
        }
        if (threadIdx.x == 0) total = compose(total, red[0]);
    }
    if (threadIdx.x == 0) a.xfer[b] = total;
}

__global__ void __launch_bounds__(kThreads) group_xfer_kernel(SampleArgs a) {
    __shared__ Xfer red[kThreads];
    const uint32_t g = blockIdx.x;
    const uint32_t b = g * kGroup + threadIdx.x;
    red[threadIdx.x] = b < a.blocks ? a.xfer[b] : identity_xfer(a.xfer[g * kGroup].s);
    __syncthreads();
    for (int stride = 1; stride < kThreads; stride <<= 1) {
        if ((threadIdx.x % (2 * stride)) == 0) red[threadIdx.x] = compose(red[threadIdx.x], red[threadIdx.x + stride]);
        __syncthreads();
    }
    if (threadIdx.x == 0) a.gxfer[g] = red[0];
}

// Tries a transfer from the exact value c; true and c advanced when valid.
__device__ __forceinline__ bool try_xfer(double& c, const Xfer& x) {
    if (x.s == kNonUniform) return false;
    const int32_t s = scale_of(c);
    if (s != x.s) return false;
    const uint64_t k = k_of(c, s);
    const uint64_t kout = sat_add(k, (k & 1u) ? x.d1 : x.d0);
    if (kout > kTop) return false;
    c = value_of(kout, s);
    return true;
}

enum Mode : uint32_t { kModeGroup = 0, kModeBlock = 1, kModeDirty = 2, kModeDone = 3 };

// One workgroup: the exact chain over all blocks.
template <typename V>
__global__ void __launch_bounds__(kThreads) chain_kernel(SampleArgs a) {
    __shared__ double c_sh;
    __shared__ uint32_t b_sh, mode_sh, sub_sh;
    __shared__ Xfer gx[kGroup];
    __shared__ double p_sh[1u << kBlockLog];
    __shared__ Xfer sx[(1u << kBlockLog) / kSub];
    const V* st = static_cast<const V*>(a.state);
    const uint32_t subs = static_cast<uint32_t>(a.block / kSub) == 0 ? 1 : static_cast<uint32_t>(a.block / kSub);
    const uint32_t sub_len = static_cast<uint32_t>(a.block / subs);
    if (threadIdx.x == 0) {
        c_sh = carry_of(a);
        b_sh = 0;
        mode_sh = kModeGroup;
        a.cstart[0] = c_sh;
    }
    __syncthreads();
    while (true) {
        const uint32_t mode = mode_sh;
        const uint32_t b = b_sh;
        // Every thread has read the shared state before thread 0 changes it.
        __syncthreads();
        if (mode == kModeDone) break;
        if (mode == kModeGroup) {
            if (threadIdx.x == 0) {
                const uint32_t g = b / kGroup;
                double c = c_sh;
                const uint32_t end = min(a.blocks, (g + 1) * kGroup);
                if (try_xfer(c, a.gxfer[g])) {
                    a.jumped[g] = 1;
                    c_sh = c;
                    a.cstart[end] = c;
                    b_sh = end;
                    mode_sh = end >= a.blocks ? kModeDone : kModeGroup;
                } else {
                    a.jumped[g] = 0;
                    mode_sh = kModeBlock;
                }
            }
            __syncthreads();
            if (mode_sh == kModeBlock) {
                const uint32_t g = b / kGroup;
                const uint32_t gb = g * kGroup + threadIdx.x;
                if (threadIdx.x < kGroup && gb < a.blocks) gx[threadIdx.x] = a.xfer[gb];
            }
            __syncthreads();
            continue;
        }
        if (mode == kModeBlock) {
            if (threadIdx.x == 0) {
                double c = c_sh;
                uint32_t bb = b;
                const uint32_t end = min(a.blocks, (b / kGroup + 1) * kGroup);
                uint32_t next = kModeGroup;
                while (bb < end) {
                    if (!try_xfer(c, gx[bb % kGroup])) {
                        next = kModeDirty;
                        break;
                    }
                    ++bb;
                    a.cstart[bb] = c;
                }
                c_sh = c;
                b_sh = bb;
                if (next == kModeGroup && bb >= a.blocks) next = kModeDone;
                mode_sh = next;
                sub_sh = 0;
            }
            __syncthreads();
            continue;
        }
        // kModeDirty: walk block b exactly.
        {
            const uint64_t i0 = uint64_t(b) * a.block;
            for (uint32_t i = threadIdx.x; i < a.block; i += kThreads) p_sh[i] = prob(st, i0 + i);
            __syncthreads();
            while (true) {
                const uint32_t sub = sub_sh;
                __syncthreads();
                if (sub >= subs) break;
                // Sub-block transfers at the current binade, for the sub-blocks still ahead.
                const int32_t s = scale_of(c_sh);
                for (uint32_t j = sub + threadIdx.x; j < subs; j += kThreads) {
                    Xfer acc = identity_xfer(s);
                    for (uint32_t i = 0; i < sub_len; ++i) acc = compose(acc, element_xfer(p_sh[j * sub_len + i], s));
                    sx[j] = acc;
                }
                __syncthreads();
                if (threadIdx.x == 0) {
                    double c = c_sh;
                    uint32_t j = sub;
                    while (j < subs && try_xfer(c, sx[j])) ++j;
                    if (j < subs) {
                        // This sub-block leaves the binade: plain FP64 adds, in order.
                        for (uint32_t i = 0; i < sub_len; ++i) c = c + p_sh[j * sub_len + i];
                        ++j;
                    }
                    c_sh = c;
                    sub_sh = j;
                }
                __syncthreads();
            }
            __syncthreads();
            if (threadIdx.x == 0) {
                const uint32_t nb = b + 1;
                a.cstart[nb] = c_sh;
                b_sh = nb;
                sub_sh = 0;
                if (nb >= a.blocks)
                    mode_sh = kModeDone;
                else
                    mode_sh = (nb % kGroup) == 0 ? kModeGroup : kModeBlock;
            }
            __syncthreads();
        }
    }
}

// Block starts inside the groups the chain jumped over.
__global__ void __launch_bounds__(kThreads) expand_kernel(SampleArgs a) {
    __shared__ Xfer gx[kGroup];
    const uint32_t g = blockIdx.x;
    if (!a.jumped[g]) return;
    const uint32_t b0 = g * kGroup;
    const uint32_t end = min(a.blocks, b0 + kGroup);
    if (b0 + threadIdx.x < end) gx[threadIdx.x] = a.xfer[b0 + threadIdx.x];
    __syncthreads();
    if (threadIdx.x == 0) {
        double c = a.cstart[b0];
        for (uint32_t b = b0; b + 1 < end; ++b) {
            try_xfer(c, gx[b - b0]);  // valid: the whole group was
            a.cstart[b + 1] = c;
        }
    }
}

template <typename V>
__global__ void __launch_bounds__(kThreads) shots_kernel(SampleArgs a, const double* u, uint64_t shots,
                                                        uint64_t* out) {
    const uint64_t t = uint64_t(blockIdx.x) * kThreads + threadIdx.x;
    if (t >= shots) return;
    const double x = u[t];
    // First block whose end value exceeds x.
    uint32_t lo = 0, hi = a.blocks;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) / 2;
        if (x < a.cstart[mid + 1])
            hi = mid;
        else
            lo = mid + 1;
    }
    if (lo == a.blocks) {
        out[t] = a.size - 1;
        return;
    }
    const V* st = static_cast<const V*>(a.state);
    double c = a.cstart[lo];
    const uint64_t i0 = uint64_t(lo) * a.block;
    uint64_t hit = i0 + a.block - 1;
    for (uint64_t i = 0; i < a.block; ++i) {
        c = c + prob(st, i0 + i);
        if (x < c) {
            hit = i0 + i;
            break;
        }
    }
    out[t] = hit;
}

// The draws of one chunk of a state that arrives in chunks (index order): a
// uniform is this chunk's when the running value before the chunk is at most
// it and the value after the chunk exceeds it (those ranges partition [0,
// total)); its index is the chunk's first index plus the hit inside it, found
// as shots_kernel finds it. Other uniforms are left alone.
template <typename V>
__global__ void __launch_bounds__(kThreads) chunk_shots_kernel(SampleArgs a, const double* u, uint64_t shots,
                                                              uint64_t* out, uint64_t first_index) {
    const uint64_t t = uint64_t(blockIdx.x) * kThreads + threadIdx.x;
    if (t >= shots) return;
    const double x = u[t];
    if (!(x >= a.cstart[0] && x < a.cstart[a.blocks])) return;
    uint32_t lo = 0, hi = a.blocks;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) / 2;
        if (x < a.cstart[mid + 1])
            hi = mid;
        else
            lo = mid + 1;
    }
    const V* st = static_cast<const V*>(a.state);
    double c = a.cstart[lo];
    const uint64_t i0 = uint64_t(lo) * a.block;
    uint64_t hit = i0 + a.block - 1;
    for (uint64_t i = 0; i < a.block; ++i) {
        c = c + prob(st, i0 + i);
        if (x < c) {
            hit = i0 + i;
            break;
        }
    }
    out[t] = first_index + hit;
}

__global__ void __launch_bounds__(kThreads) fill_index_kernel(uint64_t* out, uint64_t count, uint64_t value) {
    const uint64_t t = uint64_t(blockIdx.x) * kThreads + threadIdx.x;
    if (t < count) out[t] = value;
}

// The first half of the sampler over one state: the exact running value at
// every block start (block sums, predicted binades, composed transfers, the
// chain, the starts inside jumped groups), in device tables kept for draws.
struct SamplerTables {
    DeviceBuffer buf;
    SampleArgs a{};
};

uint64_t r16(uint64_t bytes) { return (bytes + 15) / 16 * 16; }

// Allocates the tables of a state of 2^n amplitudes at `state`.
void layout_tables(SamplerTables& t, const void* state, uint32_t n) {
    context();
    SampleArgs& a = t.a;
    a.state = state;
    a.size = uint64_t{1} << n;
    a.block = uint64_t{1} << std::min<uint32_t>(n, kBlockLog);
    a.blocks = static_cast<uint32_t>(a.size / a.block);
    a.groups = (a.blocks + kGroup - 1) / kGroup;
    const uint64_t nb = a.blocks;
    t.buf = DeviceBuffer(r16(nb * sizeof(Xfer)) + r16(uint64_t(a.groups) * sizeof(Xfer)) + r16(nb * sizeof(double)) +
                         2 * r16((nb + 1) * sizeof(double)) + r16(uint64_t(a.groups) * sizeof(uint32_t)));
    char* p = t.buf.as<char>();
    auto take = [&](uint64_t bytes) {
        char* r = p;
        p += r16(bytes);
        return r;
    };
    a.xfer = reinterpret_cast<Xfer*>(take(nb * sizeof(Xfer)));
    a.gxfer = reinterpret_cast<Xfer*>(take(uint64_t(a.groups) * sizeof(Xfer)));
    a.sums = reinterpret_cast<double*>(take(nb * sizeof(double)));
    a.approx = reinterpret_cast<double*>(take((nb + 1) * sizeof(double)));
    a.cstart = reinterpret_cast<double*>(take((nb + 1) * sizeof(double)));
    a.jumped = reinterpret_cast<uint32_t*>(take(uint64_t(a.groups) * sizeof(uint32_t)));
    a.carry = nullptr;
}

// Queues the table kernels for the state and carry in `a` on `stream`.
template <typename V>
void launch_tables(const SampleArgs& a, hipStream_t stream) {
    hipLaunchKernelGGL(block_sums_kernel<V>, dim3(a.blocks), dim3(kThreads), 0, stream, a);
    MH_HIP(hipGetLastError());
    hipLaunchKernelGGL(approx_prefix_kernel, dim3(1), dim3(1024), 0, stream, a);
    MH_HIP(hipGetLastError());
    hipLaunchKernelGGL(block_xfer_kernel<V>, dim3(a.blocks), dim3(kThreads), 0, stream, a);
    MH_HIP(hipGetLastError());
    hipLaunchKernelGGL(group_xfer_kernel, dim3(a.groups), dim3(kThreads), 0, stream, a);
    MH_HIP(hipGetLastError());
    hipLaunchKernelGGL(chain_kernel<V>, dim3(1), dim3(kThreads), 0, stream, a);
    MH_HIP(hipGetLastError());
    hipLaunchKernelGGL(expand_kernel, dim3(a.groups), dim3(kThreads), 0, stream, a);
    MH_HIP(hipGetLastError());
}

template <typename V>
void prepare_tables(SamplerTables& t, const void* state, uint32_t n) {
    layout_tables(t, state, n);
    launch_tables<V>(t.a, streams().main);
}

// The second half: each uniform's basis index (any order), waited for.
template <typename V>
void draw_basis(const SampleArgs& a, const double* uniforms, uint64_t count, uint64_t* basis_out) {
    if (count == 0) return;
    context();
    DeviceBuffer io(2 * r16(count * sizeof(double)));
    double* du = io.as<double>();
    uint64_t* dout = reinterpret_cast<uint64_t*>(io.as<char>() + r16(count * sizeof(double)));
    upload(du, uniforms, count * sizeof(double));
    const uint32_t grid = static_cast<uint32_t>((count + kThreads - 1) / kThreads);
    hipLaunchKernelGGL(shots_kernel<V>, dim3(grid), dim3(kThreads), 0, streams().main, a,
                       static_cast<const double*>(du), count, dout);
    MH_HIP(hipGetLastError());
    download(basis_out, dout, count * sizeof(uint64_t));
}

template <typename V>
void run_sampler(const void* state, uint32_t n, const double* sorted_u, uint64_t shots, uint64_t* basis_out) {
    SamplerTables t;
    prepare_tables<V>(t, state, n);
    draw_basis<V>(t.a, sorted_u, shots, basis_out);
}

}  // namespace

void sample_sorted(const void* state, uint32_t n, int32_t precision, const double* sorted_u, uint64_t shots,
                   uint64_t* basis_out) {
    amp_bytes(precision);
    if (n > 40) fail(MH_ERR_INVALID, "sample: num_qubits above 40");
    if (shots == 0) return;
    for (uint64_t i = 1; i < shots; ++i)
        if (!(sorted_u[i - 1] <= sorted_u[i])) fail(MH_ERR_INVALID, "sample: uniforms are not sorted");
    if (precision == MH_C128)
        run_sampler<double2>(state, n, sorted_u, shots, basis_out);
    else
        run_sampler<float2>(state, n, sorted_u, shots, basis_out);
}

struct ChunkedSampler::Impl {
    uint32_t n = 0, chunk_log = 0;
    int32_t precision = MH_C128;
    uint64_t shots = 0, next = 0;
    hipStream_t stream = nullptr;
    SamplerTables tables[2];
    DeviceBuffer io;  // uniforms, basis indices, the zero carry of chunk 0
    double* du = nullptr;
    uint64_t* dout = nullptr;
    double* zero = nullptr;
};

ChunkedSampler::ChunkedSampler(uint32_t n, uint32_t chunk_log, int32_t precision, const double* sorted_u,
                               uint64_t shots, hipStream_t stream)
    : impl_(std::make_unique<Impl>()) {
    amp_bytes(precision);
    if (n > 40 || chunk_log > n) fail(MH_ERR_INVALID, "chunked sampler: chunk above the state or above 40 qubits");
    for (uint64_t i = 1; i < shots; ++i)
        if (!(sorted_u[i - 1] <= sorted_u[i])) fail(MH_ERR_INVALID, "sample: uniforms are not sorted");
    Impl& m = *impl_;
    m.n = n;
    m.chunk_log = chunk_log;
    m.precision = precision;
    m.shots = shots;
    m.stream = stream;
    for (SamplerTables& t : m.tables) layout_tables(t, nullptr, chunk_log);
    m.io = DeviceBuffer(r16(std::max<uint64_t>(shots, 1) * sizeof(double)) +
                        r16(std::max<uint64_t>(shots, 1) * sizeof(uint64_t)) + 16);
    m.du = m.io.as<double>();
    m.dout = reinterpret_cast<uint64_t*>(m.io.as<char>() + r16(std::max<uint64_t>(shots, 1) * sizeof(double)));
    m.zero = reinterpret_cast<double*>(reinterpret_cast<char*>(m.dout) + r16(std::max<uint64_t>(shots, 1) * sizeof(uint64_t)));
    MH_HIP(hipMemsetAsync(m.zero, 0, sizeof(double), stream));
    if (shots > 0) {
        MH_HIP(hipMemcpyAsync(m.du, sorted_u, shots * sizeof(double), hipMemcpyHostToDevice, stream));
        // A uniform at or above the total goes to the last index.
        hipLaunchKernelGGL(fill_index_kernel, dim3(static_cast<uint32_t>((shots + kThreads - 1) / kThreads)),
                           dim3(kThreads), 0, stream, m.dout, shots, (uint64_t{1} << n) - 1);
        MH_HIP(hipGetLastError());
        // The uniforms are the caller's memory: the copy is complete before it returns.
        MH_HIP(hipStreamSynchronize(stream));
    }
}

ChunkedSampler::~ChunkedSampler() {
    if (impl_) (void)hipStreamSynchronize(impl_->stream);
}

uint64_t ChunkedSampler::device_bytes() const {
    return impl_->tables[0].buf.bytes() + impl_->tables[1].buf.bytes() + impl_->io.bytes();
}

void ChunkedSampler::add(const void* chunk) {
    Impl& m = *impl_;
    const uint64_t c = m.next;
    if (c >= (uint64_t{1} << (m.n - m.chunk_log))) fail(MH_ERR_INTERNAL, "chunked sampler: more chunks than the state has");
    m.next = c + 1;
    SampleArgs a = m.tables[c % 2].a;
    a.state = chunk;
    const SampleArgs& before = m.tables[(c + 1) % 2].a;
    a.carry = c == 0 ? m.zero : before.cstart + before.blocks;
    // The tables of chunk c + 2 take this slot: the stream orders them after
    // this chunk's draws and after chunk c + 1 has read the carry.
    m.tables[c % 2].a = a;
    if (m.precision == MH_C128)
        launch_tables<double2>(a, m.stream);
    else
        launch_tables<float2>(a, m.stream);
    if (m.shots == 0) return;
    const uint32_t grid = static_cast<uint32_t>((m.shots + kThreads - 1) / kThreads);
    if (m.precision == MH_C128)
        hipLaunchKernelGGL(chunk_shots_kernel<double2>, dim3(grid), dim3(kThreads), 0, m.stream, a,
                           static_cast<const double*>(m.du), m.shots, m.dout, c << m.chunk_log);
    else
        hipLaunchKernelGGL(chunk_shots_kernel<float2>, dim3(grid), dim3(kThreads), 0, m.stream, a,
                           static_cast<const double*>(m.du), m.shots, m.dout, c << m.chunk_log);
    MH_HIP(hipGetLastError());
}

std::vector<uint64_t> ChunkedSampler::basis() {
    Impl& m = *impl_;
    if (m.next != (uint64_t{1} << (m.n - m.chunk_log))) fail(MH_ERR_INTERNAL, "chunked sampler: chunks missing");
    std::vector<uint64_t> out(m.shots);
    if (m.shots > 0) MH_HIP(hipMemcpyAsync(out.data(), m.dout, m.shots * sizeof(uint64_t), hipMemcpyDeviceToHost, m.stream));
    MH_HIP(hipStreamSynchronize(m.stream));
    return out;
}

}  // namespace mh

// A final state kept on the device with its sampler tables.
struct mh_region_sampler {
    mh::DeviceBuffer state;
    mh::SamplerTables tables;
    uint32_t n = 0;
    int32_t precision = MH_C128;
};

namespace mh {

mh_region_sampler* make_region_sampler(DeviceBuffer&& state, uint32_t n, int32_t precision) {
    amp_bytes(precision);
    if (n > 40) fail(MH_ERR_INVALID, "region sampler: num_qubits above 40");
    auto s = std::make_unique<mh_region_sampler>();
    s->state = std::move(state);
    s->n = n;
    s->precision = precision;
    if (precision == MH_C128)
        prepare_tables<double2>(s->tables, s->state.get(), n);
    else
        prepare_tables<float2>(s->tables, s->state.get(), n);
    synchronize();
    return s.release();
}

}  // namespace mh

using namespace mh;

extern "C" {

int mh_sample(const void* state, uint32_t num_qubits, int32_t precision, const double* sorted_uniforms,
              uint64_t shots, uint64_t* basis_out) {
    return guarded([&] { sample_sorted(state, num_qubits, precision, sorted_uniforms, shots, basis_out); });
}

int mh_region_sampler_draw(const mh_region_sampler* sampler, const double* uniforms, uint64_t count,
                           uint64_t* basis_out) {
    return guarded([&] {
        if (sampler == nullptr) fail(MH_ERR_INVALID, "null region sampler");
        if (count == 0) return;
        if (uniforms == nullptr || basis_out == nullptr) fail(MH_ERR_INVALID, "null uniforms or output");
        for (uint64_t i = 0; i < count; ++i)
            if (!(uniforms[i] >= 0.0 && uniforms[i] < 1.0)) fail(MH_ERR_INVALID, "region sampler: a uniform outside [0, 1)");
        if (sampler->precision == MH_C128)
            draw_basis<double2>(sampler->tables.a, uniforms, count, basis_out);
        else
            draw_basis<float2>(sampler->tables.a, uniforms, count, basis_out);
    });
}

int mh_region_sampler_info(const mh_region_sampler* sampler, uint32_t* num_qubits, uint64_t* device_bytes) {
    return guarded([&] {
        if (sampler == nullptr) fail(MH_ERR_INVALID, "null region sampler");
        if (num_qubits) *num_qubits = sampler->n;
        if (device_bytes) *device_bytes = sampler->state.bytes() + sampler->tables.buf.bytes();
    });
}

int mh_region_sampler_destroy(mh_region_sampler* sampler) {
    return guarded([&] {
        if (sampler == nullptr) return;
        synchronize();
        delete sampler;
    });
}

int mh_sample_unsorted(const void* state, uint32_t num_qubits, int32_t precision, const double* uniforms,
                       uint64_t count, uint64_t* basis_out) {
    return guarded([&] {
        std::vector<uint64_t> order(count);
        std::iota(order.begin(), order.end(), uint64_t{0});
        std::stable_sort(order.begin(), order.end(), [&](uint64_t x, uint64_t y) { return uniforms[x] < uniforms[y]; });
        std::vector<double> sorted(count);
        for (uint64_t i = 0; i < count; ++i) sorted[i] = uniforms[order[i]];
        std::vector<uint64_t> hits(count);
        sample_sorted(state, num_qubits, precision, sorted.data(), count, hits.data());
        for (uint64_t i = 0; i < count; ++i) basis_out[order[i]] = hits[i];
    });
}

}  // extern "C"
