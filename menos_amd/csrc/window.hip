// The gate-pass kernel: many plan blocks per HBM read and write of the state.
//
// Each wavefront takes one window of W qubits (the other qubits fixed by the
// wave's index) and holds its 2^W amplitudes in registers: lane bit i of the
// wave's 64 lanes and register bit j of a lane's 2^B amplitudes (B = 4 at
// complex128, 5 at complex64; W = 6 + B) each stand for one window qubit. A
// dense block acts inside every lane once its operands sit on register bits.
// Moving a qubit between a lane bit and a register bit is an exchange between
// partner lanes done in registers (v_permlane32_swap / v_permlane16_swap for
// lane bits 5 and 4, DPP moves for lane bits 0..3), so the kernel uses no LDS
// and no barrier: the waves of a CU run independently, and while some wait on
// HBM the others compute.
//
// The host compiles every pass into a list of ops (a dense block on given
// register bits, a diagonal block, a lane<->register exchange, a swap of two
// register bits) plus the layouts at the load and at the store. Both keep the
// low qubits 0..L-1 (L = 3 at complex128, 4 at complex64: 128 contiguous
// bytes) on lane bits 0..L-1, so every global access covers whole 128-byte
// lines.
//
// Arithmetic (complex128). The reference generic kernel (reference
// src/kernels/cpu/dense_kernels.cpp:465-524) computes every output as
// (+0, +0) + sum over `in` ascending of M[out][in] * a[in], each product
// (mr*ar - mi*ai, mr*ai + mi*ar), every operation rounded on its own; a
// diagonal block as (+0) + d * a (dense_kernels.cpp:299-313). Write x ~ y when
// x and y are equal or are both zeros (of any signs). On finite values + - *
// respect ~, and such a sum that starts at +0 is never -0, so each reference
// block maps ~-equal inputs to identical outputs. This kernel evaluates every
// block only up to ~: it drops the leading (+0) +, skips products with an
// entry part that is a zero (mi * ai ~ 0), and applies a diagonal whose
// entries are all (+-1, +-0) as sign flips; the sums keep the reference's
// order, so by induction over the blocks the state stays ~-equal to the
// reference's. The reference's state after any block has no -0 component, so
// adding +0 to every component at the end of the last pass gives its bits
// exactly. Complex64 passes use fused multiply-adds and carry a stated error.
#include <algorithm>
#include <cstring>

#include "mh_internal.h"
#include "window.h"

namespace mh {
namespace {

constexpr uint32_t kLaneBits = 6;
constexpr uint32_t kWavesPerGroup = 4;
constexpr uint32_t kNone = 255;
constexpr uint32_t kMaxRuns = 16;
constexpr uint32_t kMaxWindowDense = 2;  // widest dense block a window pass applies

// Op codes.
constexpr uint32_t kCodeDense1 = 0;     // + register bit J: a 1-qubit block
constexpr uint32_t kCodeDense1Re = 8;   // + J: its entry (0, 0) has a zero imaginary part
constexpr uint32_t kCodeDenseK = 16;    // + k: operands on register bits 0..k-1 in operand order
constexpr uint32_t kCodeDiag = 24;
constexpr uint32_t kCodeDiagSign = 25;  // every entry (+-1, +-0)
constexpr uint32_t kCodeLaneSwap = 32;  // + 8 * lane bit + register bit
constexpr uint32_t kCodeRegSwap = 96;   // + 8 * j1 + j2, j1 < j2

template <typename R> struct Shape;
template <> struct Shape<double> {
    static constexpr uint32_t kRegBits = 4, kLow = 3;
    using V = double2;
};
template <> struct Shape<float> {
    static constexpr uint32_t kRegBits = 5, kLow = 4;
    using V = float2;
};

struct WOp {
    uint32_t code;
    uint32_t mat;            // first entry in the matrix table
    uint64_t neg_mask;       // sign diagonal: bit e set when entry e is -1
    uint8_t lane_bit[8];     // diagonal: operand bit that lane bit i supplies, or kNone
    uint8_t outer_q[8];      // diagonal: global qubit of the m-th operand outside the window
    uint8_t outer_b[8];      // ... and its operand bit
    uint8_t rpart[32];       // diagonal: operand bits that register index r supplies
    uint32_t num_outer;
    uint32_t pad;
};

struct WPassArgs {
    void* state;
    const WOp* ops;
    const void* mats;
    uint32_t num_ops;
    uint32_t normalize;       // add +0 to every component before the store
    uint32_t num_runs;        // runs of consecutive qubits outside the window
    uint8_t run_start[kMaxRuns], run_len[kMaxRuns];
    uint8_t load_lane_q[8], load_reg_q[8];    // global qubit of each lane / register bit at the load
    uint8_t store_lane_q[8], store_reg_q[8];  // ... and at the store
};

template <uint32_t N> struct Words { uint32_t w[N]; };

// The qubits outside the window, taken from the wave index (a bit deposit).
__device__ __forceinline__ uint64_t outer_base(uint32_t wave, const WPassArgs& a) {
    uint64_t base = 0;
    uint32_t src = 0;
    for (uint32_t i = 0; i < a.num_runs; ++i) {
        const uint32_t len = a.run_len[i];
        base |= (uint64_t(wave >> src) & ((1ull << len) - 1ull)) << a.run_start[i];
        src += len;
    }
    return base;
}

__device__ __forceinline__ uint64_t lane_offset(uint32_t lane, const uint8_t* q) {
    uint64_t off = 0;
#pragma unroll
    for (uint32_t i = 0; i < kLaneBits; ++i) off |= uint64_t((lane >> i) & 1u) << q[i];
    return off;
}

// Exchange across lane bit I: afterwards the lanes with bit I clear hold
// (x, partner's x) and those with it set (partner's y, y), i.e. lane bit I and
// the register bit that tells x from y trade places.
template <uint32_t I>
__device__ __forceinline__ void xchg(uint32_t& x, uint32_t& y, bool upper) {
    if constexpr (I == 5) {
        const auto p = __builtin_amdgcn_permlane32_swap(x, y, false, false);
        x = p[0];
        y = p[1];
    } else if constexpr (I == 4) {
        const auto p = __builtin_amdgcn_permlane16_swap(x, y, false, false);
        x = p[0];
        y = p[1];
    } else {
        // Lanes with bit I set take y from lane l - 2^I, the others x from
        // lane l + 2^I: quad_perm lane ^ 1 (0xB1) or lane ^ 2 (0x4E) for
        // I < 2, row_shr / row_shl by 4 or 8 otherwise (the partner is in the
        // same row of 16 lanes).
        constexpr int from_lower = I == 0 ? 0xB1 : I == 1 ? 0x4E : I == 2 ? 0x114 : 0x118;
        constexpr int from_upper = I == 0 ? 0xB1 : I == 1 ? 0x4E : I == 2 ? 0x104 : 0x108;
        const uint32_t fy = __builtin_amdgcn_update_dpp(0u, y, from_lower, 0xF, 0xF, false);
        const uint32_t fx = __builtin_amdgcn_update_dpp(0u, x, from_upper, 0xF, 0xF, false);
        const uint32_t nx = upper ? fy : x;
        const uint32_t ny = upper ? y : fx;
        x = nx;
        y = ny;
    }
}

// Lane bit I <-> register bit J.
template <typename V, uint32_t E, uint32_t I, uint32_t J>
__device__ __forceinline__ void lane_swap(V* v, uint32_t lane) {
    if constexpr ((1u << J) < E) {
        constexpr uint32_t D = sizeof(V) / 4;
        const bool upper = (lane >> I) & 1u;
#pragma unroll
        for (uint32_t r = 0; r < E; ++r) {
            if (r & (1u << J)) continue;
            Words<D> x = __builtin_bit_cast(Words<D>, v[r]);
            Words<D> y = __builtin_bit_cast(Words<D>, v[r | (1u << J)]);
#pragma unroll
            for (uint32_t d = 0; d < D; ++d) xchg<I>(x.w[d], y.w[d], upper);
            v[r] = __builtin_bit_cast(V, x);
            v[r | (1u << J)] = __builtin_bit_cast(V, y);
        }
    }
}

// Register bit J1 <-> register bit J2 (a renaming of the amplitudes).
template <typename V, uint32_t E, uint32_t J1, uint32_t J2>
__device__ __forceinline__ void reg_swap(V* v) {
    if constexpr ((1u << J2) < E) {
#pragma unroll
        for (uint32_t r = 0; r < E; ++r) {
            if (!((r >> J1) & 1u) || ((r >> J2) & 1u)) continue;
            const uint32_t s = r ^ (1u << J1) ^ (1u << J2);
            const V t = v[r];
            v[r] = v[s];
            v[s] = t;
        }
    }
}

// One 1-qubit block on register bit J.
template <uint32_t E, uint32_t J, bool kRe>
__device__ __forceinline__ void dense1(double2* v, const double2* m) {
    if constexpr ((1u << J) < E) {
        const double2 m00 = m[0], m01 = m[1], m10 = m[2], m11 = m[3];
#pragma unroll
        for (uint32_t r = 0; r < E; ++r) {
            if (r & (1u << J)) continue;
            const double2 a0 = v[r], a1 = v[r | (1u << J)];
            double s0r, s0i;
            if constexpr (kRe) {
                s0r = m00.x * a0.x;
                s0i = m00.x * a0.y;
            } else {
                s0r = m00.x * a0.x - m00.y * a0.y;
                s0i = m00.x * a0.y + m00.y * a0.x;
            }
            s0r = s0r + (m01.x * a1.x - m01.y * a1.y);
            s0i = s0i + (m01.x * a1.y + m01.y * a1.x);
            const double s1r = (m10.x * a0.x - m10.y * a0.y) + (m11.x * a1.x - m11.y * a1.y);
            const double s1i = (m10.x * a0.y + m10.y * a0.x) + (m11.x * a1.y + m11.y * a1.x);
            v[r] = make_double2(s0r, s0i);
            v[r | (1u << J)] = make_double2(s1r, s1i);
        }
    }
}

__device__ __forceinline__ void cmac(float& sr, float& si, float2 m, float2 a) {
    sr = __builtin_fmaf(m.x, a.x, sr);
    sr = __builtin_fmaf(-m.y, a.y, sr);
    si = __builtin_fmaf(m.x, a.y, si);
    si = __builtin_fmaf(m.y, a.x, si);
}

template <uint32_t E, uint32_t J, bool kRe>
__device__ __forceinline__ void dense1(float2* v, const float2* m) {
    if constexpr ((1u << J) < E) {
        const float2 m00 = m[0], m01 = m[1], m10 = m[2], m11 = m[3];
#pragma unroll
        for (uint32_t r = 0; r < E; ++r) {
            if (r & (1u << J)) continue;
            const float2 a0 = v[r], a1 = v[r | (1u << J)];
            float s0r = m00.x * a0.x, s0i = m00.x * a0.y;
            if constexpr (!kRe) {
                s0r = __builtin_fmaf(-m00.y, a0.y, s0r);
                s0i = __builtin_fmaf(m00.y, a0.x, s0i);
            }
            cmac(s0r, s0i, m01, a1);
            float s1r = 0.0f, s1i = 0.0f;
            cmac(s1r, s1i, m10, a0);
            cmac(s1r, s1i, m11, a1);
            v[r] = make_float2(s0r, s0i);
            v[r | (1u << J)] = make_float2(s1r, s1i);
        }
    }
}

// A k-qubit block whose operands are register bits 0..K-1 in operand order.
template <uint32_t E, uint32_t K>
__device__ __forceinline__ void densek(double2* v, const double2* m) {
    if constexpr ((1u << K) <= E) {
        constexpr uint32_t dim = 1u << K;
#pragma unroll
        for (uint32_t hi = 0; hi < (E >> K); ++hi) {
            double2 a[dim];
#pragma unroll
            for (uint32_t in = 0; in < dim; ++in) a[in] = v[in | (hi << K)];
#pragma unroll
            for (uint32_t o = 0; o < dim; ++o) {
                const double2 e0 = m[o * dim];
                double sr = e0.x * a[0].x - e0.y * a[0].y;
                double si = e0.x * a[0].y + e0.y * a[0].x;
#pragma unroll
                for (uint32_t in = 1; in < dim; ++in) {
                    const double2 e = m[o * dim + in];
                    sr = sr + (e.x * a[in].x - e.y * a[in].y);
                    si = si + (e.x * a[in].y + e.y * a[in].x);
                }
                v[o | (hi << K)] = make_double2(sr, si);
            }
        }
    }
}

template <uint32_t E, uint32_t K>
__device__ __forceinline__ void densek(float2* v, const float2* m) {
    if constexpr ((1u << K) <= E) {
        constexpr uint32_t dim = 1u << K;
#pragma unroll
        for (uint32_t hi = 0; hi < (E >> K); ++hi) {
            float2 a[dim];
#pragma unroll
            for (uint32_t in = 0; in < dim; ++in) a[in] = v[in | (hi << K)];
#pragma unroll
            for (uint32_t o = 0; o < dim; ++o) {
                float sr = 0.0f, si = 0.0f;
#pragma unroll
                for (uint32_t in = 0; in < dim; ++in) cmac(sr, si, m[o * dim + in], a[in]);
                v[o | (hi << K)] = make_float2(sr, si);
            }
        }
    }
}

__device__ __forceinline__ double2 diag_mul(double2 d, double2 a) {
    return make_double2(d.x * a.x - d.y * a.y, d.x * a.y + d.y * a.x);
}

__device__ __forceinline__ float2 diag_mul(float2 d, float2 a) {
    return make_float2(__builtin_fmaf(d.x, a.x, -d.y * a.y), __builtin_fmaf(d.x, a.y, d.y * a.x));
}

__device__ __forceinline__ void flip_sign(double2& a, uint32_t s) {
    const uint64_t m = uint64_t(s) << 63;
    a.x = __longlong_as_double(static_cast<long long>(static_cast<uint64_t>(__double_as_longlong(a.x)) ^ m));
    a.y = __longlong_as_double(static_cast<long long>(static_cast<uint64_t>(__double_as_longlong(a.y)) ^ m));
}

__device__ __forceinline__ void flip_sign(float2& a, uint32_t s) {
    const uint32_t m = s << 31;
    a.x = __uint_as_float(__float_as_uint(a.x) ^ m);
    a.y = __uint_as_float(__float_as_uint(a.y) ^ m);
}

// The index bits of a diagonal's entry that come from outside the window and
// from the lane: the register index adds op.rpart[r].
__device__ __forceinline__ uint32_t diag_base(const WOp& op, uint64_t obase, uint32_t lane) {
    uint32_t t = 0;
    for (uint32_t m = 0; m < op.num_outer; ++m) t |= uint32_t((obase >> op.outer_q[m]) & 1ull) << op.outer_b[m];
#pragma unroll
    for (uint32_t i = 0; i < kLaneBits; ++i)
        if (op.lane_bit[i] != kNone) t |= ((lane >> i) & 1u) << op.lane_bit[i];
    return t;
}

#define MH_CASES_J(BASE, CALL) \
    case BASE + 0: CALL(0); break; \
    case BASE + 1: CALL(1); break; \
    case BASE + 2: CALL(2); break; \
    case BASE + 3: CALL(3); break; \
    case BASE + 4: CALL(4); break;

// Three waves per SIMD: each wave holds its window in about 130 VGPRs.
template <typename R>
__global__ void __launch_bounds__(kWavesPerGroup * 64, 3) window_kernel(WPassArgs a) {
    using V = typename Shape<R>::V;
    constexpr uint32_t RB = Shape<R>::kRegBits;
    constexpr uint32_t E = 1u << RB;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
    const uint64_t obase = outer_base(wave, a);
    V* const st = static_cast<V*>(a.state) + obase;
    const V* const mats = static_cast<const V*>(a.mats);

    V v[E];
    {
        const V* src = st + lane_offset(lane, a.load_lane_q);
        uint64_t rq[RB];
#pragma unroll
        for (uint32_t j = 0; j < RB; ++j) rq[j] = 1ull << a.load_reg_q[j];
#pragma unroll
        for (uint32_t r = 0; r < E; ++r) {
            uint64_t h = 0;
#pragma unroll
            for (uint32_t j = 0; j < RB; ++j)
                if ((r >> j) & 1u) h |= rq[j];
            v[r] = src[h];
        }
    }

    for (uint32_t i = 0; i < a.num_ops; ++i) {
        const WOp& op = a.ops[i];
        const V* m = mats + op.mat;
        switch (op.code) {
#define MH_D1(J) dense1<E, J, false>(v, m)
            MH_CASES_J(kCodeDense1, MH_D1)
#undef MH_D1
#define MH_D1R(J) dense1<E, J, true>(v, m)
            MH_CASES_J(kCodeDense1Re, MH_D1R)
#undef MH_D1R
            case kCodeDenseK + 2: densek<E, 2>(v, m); break;
            case kCodeDiag: {
                // Entries gathered four at a time (bounds the registers they hold).
                const uint32_t t = diag_base(op, obase, lane);
#pragma unroll
                for (uint32_t r = 0; r < E; ++r) {
                    v[r] = diag_mul(m[t | op.rpart[r]], v[r]);
                    if ((r & 3u) == 3u) __builtin_amdgcn_sched_barrier(0);
                }
                break;
            }
            case kCodeDiagSign: {
                const uint64_t sh = op.neg_mask >> diag_base(op, obase, lane);
#pragma unroll
                for (uint32_t r = 0; r < E; ++r) flip_sign(v[r], uint32_t(sh >> op.rpart[r]) & 1u);
                break;
            }
#define MH_LS0(J) lane_swap<V, E, 0, J>(v, lane)
            MH_CASES_J(kCodeLaneSwap + 0, MH_LS0)
#undef MH_LS0
#define MH_LS1(J) lane_swap<V, E, 1, J>(v, lane)
            MH_CASES_J(kCodeLaneSwap + 8, MH_LS1)
#undef MH_LS1
#define MH_LS2(J) lane_swap<V, E, 2, J>(v, lane)
            MH_CASES_J(kCodeLaneSwap + 16, MH_LS2)
#undef MH_LS2
#define MH_LS3(J) lane_swap<V, E, 3, J>(v, lane)
            MH_CASES_J(kCodeLaneSwap + 24, MH_LS3)
#undef MH_LS3
#define MH_LS4(J) lane_swap<V, E, 4, J>(v, lane)
            MH_CASES_J(kCodeLaneSwap + 32, MH_LS4)
#undef MH_LS4
#define MH_LS5(J) lane_swap<V, E, 5, J>(v, lane)
            MH_CASES_J(kCodeLaneSwap + 40, MH_LS5)
#undef MH_LS5
            case kCodeRegSwap + 8 * 0 + 1: reg_swap<V, E, 0, 1>(v); break;
            case kCodeRegSwap + 8 * 0 + 2: reg_swap<V, E, 0, 2>(v); break;
            case kCodeRegSwap + 8 * 0 + 3: reg_swap<V, E, 0, 3>(v); break;
            case kCodeRegSwap + 8 * 0 + 4: reg_swap<V, E, 0, 4>(v); break;
            case kCodeRegSwap + 8 * 1 + 2: reg_swap<V, E, 1, 2>(v); break;
            case kCodeRegSwap + 8 * 1 + 3: reg_swap<V, E, 1, 3>(v); break;
            case kCodeRegSwap + 8 * 1 + 4: reg_swap<V, E, 1, 4>(v); break;
            case kCodeRegSwap + 8 * 2 + 3: reg_swap<V, E, 2, 3>(v); break;
            case kCodeRegSwap + 8 * 2 + 4: reg_swap<V, E, 2, 4>(v); break;
            case kCodeRegSwap + 8 * 3 + 4: reg_swap<V, E, 3, 4>(v); break;
            default: break;
        }
    }

    if (a.normalize) {
#pragma unroll
        for (uint32_t r = 0; r < E; ++r) {
            v[r].x = v[r].x + R(0);
            v[r].y = v[r].y + R(0);
        }
    }
    {
        // Addresses recomputed from an opaque copy of the lane index rather
        // than kept live across the whole pass.
        uint32_t lane2 = lane;
        asm volatile("" : "+v"(lane2));
        V* dst = st + lane_offset(lane2, a.store_lane_q);
        uint64_t rq[RB];
#pragma unroll
        for (uint32_t j = 0; j < RB; ++j) rq[j] = 1ull << a.store_reg_q[j];
#pragma unroll
        for (uint32_t r = 0; r < E; ++r) {
            uint64_t h = 0;
#pragma unroll
            for (uint32_t j = 0; j < RB; ++j)
                if ((r >> j) & 1u) h |= rq[j];
            dst[h] = v[r];
        }
    }
}

#undef MH_CASES_J

// ---- host: passes

template <typename R>
constexpr uint32_t width_of() {
    return kLaneBits + Shape<R>::kRegBits;
}

bool sign_diagonal(const Block& b) {
    const std::size_t dim = std::size_t{1} << b.operands.size();
    for (std::size_t r = 0; r < dim; ++r) {
        const cplx e = b.matrix[r * dim + r];
        if (!((e.real() == 1.0 || e.real() == -1.0) && e.imag() == 0.0)) return false;
    }
    return true;
}

// FP operations per amplitude a block costs in a pass (the pass cap unit),
// with an allowance for the exchanges that bring its operands to registers.
uint32_t block_cost(const Block& b) {
    const uint32_t k = static_cast<uint32_t>(b.operands.size());
    if (b.diagonal) return sign_diagonal(b) ? 2 : 6;
    if (k == 1) return (b.matrix[0].imag() == 0.0 ? 12 : 14) + 2;
    return 2 * (3 * (1u << k) - 1) + 4;
}

uint64_t mask_of(const std::vector<uint32_t>& q) {
    uint64_t m = 0;
    for (uint32_t x : q) m |= 1ull << x;
    return m;
}

// Compiles one pass: where every window qubit lives (lane bit or register
// bit) through the pass, the exchanges that bring each dense block's operands
// to register bits, and the ops.
template <typename R>
class PassCompiler {
public:
    using V = typename Shape<R>::V;
    static constexpr uint32_t RB = Shape<R>::kRegBits, L = Shape<R>::kLow, W = width_of<R>(), E = 1u << RB;

    PassCompiler(std::vector<WOp>& ops, std::vector<V>& mats) : ops_(ops), mats_(mats) {}

    WPassArgs compile(uint32_t n, const std::vector<Block>& blocks, const WindowPass& pass, bool normalize) {
        WPassArgs a;
        std::memset(&a, 0, sizeof(a));
        // Window qubits: the low run, the dense operands, then the lowest others.
        uint64_t wmask = pass.dense_mask | ((1ull << L) - 1ull);
        for (uint32_t q = 0; q < n && static_cast<uint32_t>(__builtin_popcountll(wmask)) < W; ++q) wmask |= 1ull << q;
        if (static_cast<uint32_t>(__builtin_popcountll(wmask)) != W) fail(MH_ERR_INTERNAL, "window pass shape");
        std::memset(pos_of_, kNone, sizeof(pos_of_));
        uint32_t count = 0;
        for (uint32_t q = 0; q < n; ++q)
            if ((wmask >> q) & 1ull) {
                wq_[count] = static_cast<uint8_t>(q);
                pos_of_[q] = static_cast<uint8_t>(count++);
            }
        for (uint32_t q = 0; q < n;) {
            if ((wmask >> q) & 1ull) {
                ++q;
                continue;
            }
            const uint32_t s = q;
            while (q < n && !((wmask >> q) & 1ull)) ++q;
            if (a.num_runs >= kMaxRuns) fail(MH_ERR_INTERNAL, "window pass shape");
            a.run_start[a.num_runs] = static_cast<uint8_t>(s);
            a.run_len[a.num_runs] = static_cast<uint8_t>(q - s);
            ++a.num_runs;
        }
        blocks_.clear();
        for (uint32_t b : pass.blocks) blocks_.push_back(&blocks[b]);
        initial_layout();
        for (uint32_t i = 0; i < kLaneBits; ++i) a.load_lane_q[i] = wq_[lane_[i]];
        for (uint32_t j = 0; j < RB; ++j) a.load_reg_q[j] = wq_[reg_[j]];

        for (std::size_t i = 0; i < blocks_.size(); ++i) {
            const Block& b = *blocks_[i];
            const uint32_t k = static_cast<uint32_t>(b.operands.size());
            if (b.diagonal) {
                emit_diag(b);
                continue;
            }
            if (k == 1) {
                const uint8_t p = pos_of_[b.operands[0]];
                int j = reg_index(p);
                if (j < 0) {
                    j = victim(i, p);
                    lane_swap(static_cast<uint32_t>(lane_index(p)), static_cast<uint32_t>(j));
                }
                emit_dense1(b, static_cast<uint32_t>(j));
            } else {
                for (uint32_t t = 0; t < k; ++t) {
                    const uint8_t p = pos_of_[b.operands[t]];
                    if (reg_[t] == p) continue;
                    const int j = reg_index(p);
                    if (j >= 0)
                        reg_swap(t, static_cast<uint32_t>(j));
                    else
                        lane_swap(static_cast<uint32_t>(lane_index(p)), t);
                }
                emit_densek(b);
            }
        }
        // The low qubits back on the low lane bits for the store.
        for (uint32_t l = 0; l < L; ++l) {
            if (lane_[l] < L) continue;
            uint8_t p = kNone;
            for (uint8_t c = 0; c < L && p == kNone; ++c) {
                const int li = lane_index(c);
                if (li < 0 || static_cast<uint32_t>(li) >= L) p = c;
            }
            int j = reg_index(p);
            if (j < 0) {
                for (uint32_t r = 0; r < RB && j < 0; ++r)
                    if (reg_[r] >= L) j = static_cast<int>(r);
                lane_swap(static_cast<uint32_t>(lane_index(p)), static_cast<uint32_t>(j));
            }
            lane_swap(l, static_cast<uint32_t>(j));
        }
        for (uint32_t i = 0; i < kLaneBits; ++i) a.store_lane_q[i] = wq_[lane_[i]];
        for (uint32_t j = 0; j < RB; ++j) a.store_reg_q[j] = wq_[reg_[j]];
        a.normalize = normalize ? 1u : 0u;
        return a;
    }

private:
    bool uses(std::size_t i, uint8_t p) const {
        const Block& b = *blocks_[i];
        if (b.diagonal) return false;
        for (uint32_t q : b.operands)
            if (pos_of_[q] == p) return true;
        return false;
    }

    std::size_t next_use(uint8_t p, std::size_t from) const {
        for (std::size_t i = from; i < blocks_.size(); ++i)
            if (uses(i, p)) return i;
        return SIZE_MAX;
    }

    int reg_index(uint8_t p) const {
        for (uint32_t j = 0; j < RB; ++j)
            if (reg_[j] == p) return static_cast<int>(j);
        return -1;
    }

    int lane_index(uint8_t p) const {
        for (uint32_t i = 0; i < kLaneBits; ++i)
            if (lane_[i] == p) return static_cast<int>(i);
        return -1;
    }

    // Registers: the first non-low qubits the dense blocks use. Lane bits:
    // the low run on 0..L-1, the qubits needed soonest on the cheapest
    // exchange (lane bits 5, 4, ...), unused qubits lowest-first from bit L.
    void initial_layout() {
        std::vector<uint8_t> regs;
        for (const Block* b : blocks_) {
            if (b->diagonal) continue;
            for (uint32_t q : b->operands) {
                const uint8_t p = pos_of_[q];
                if (p >= L && regs.size() < RB && std::find(regs.begin(), regs.end(), p) == regs.end())
                    regs.push_back(p);
            }
        }
        for (int p = static_cast<int>(W) - 1; regs.size() < RB && p >= static_cast<int>(L); --p)
            if (std::find(regs.begin(), regs.end(), p) == regs.end() && next_use(static_cast<uint8_t>(p), 0) == SIZE_MAX)
                regs.push_back(static_cast<uint8_t>(p));
        for (int p = static_cast<int>(W) - 1; regs.size() < RB && p >= static_cast<int>(L); --p)
            if (std::find(regs.begin(), regs.end(), p) == regs.end()) regs.push_back(static_cast<uint8_t>(p));
        for (uint32_t j = 0; j < RB; ++j) reg_[j] = regs[j];
        for (uint32_t i = 0; i < L; ++i) lane_[i] = static_cast<uint8_t>(i);
        std::vector<std::pair<std::size_t, int>> rest;
        for (uint32_t p = L; p < W; ++p) {
            if (std::find(regs.begin(), regs.end(), p) != regs.end()) continue;
            const std::size_t u = next_use(static_cast<uint8_t>(p), 0);
            rest.emplace_back(u, u == SIZE_MAX ? -static_cast<int>(p) : static_cast<int>(p));
        }
        std::sort(rest.begin(), rest.end());
        uint32_t lb = kLaneBits;
        for (const auto& r : rest) lane_[--lb] = static_cast<uint8_t>(r.second < 0 ? -r.second : r.second);
        if (lb != L) fail(MH_ERR_INTERNAL, "window layout");
    }

    // The register whose qubit is needed furthest ahead (never, first).
    int victim(std::size_t i, uint8_t incoming) const {
        int best = -1;
        std::size_t best_use = 0;
        for (uint32_t j = 0; j < RB; ++j) {
            if (reg_[j] == incoming) continue;
            const std::size_t u = next_use(reg_[j], i + 1);
            if (best < 0 || u >= best_use) {
                best = static_cast<int>(j);
                best_use = u;
            }
        }
        return best;
    }

    void push_op(const WOp& op) { ops_.push_back(op); }

    void lane_swap(uint32_t i, uint32_t j) {
        WOp op;
        std::memset(&op, 0, sizeof(op));
        op.code = kCodeLaneSwap + 8 * i + j;
        push_op(op);
        std::swap(lane_[i], reg_[j]);
    }

    void reg_swap(uint32_t j1, uint32_t j2) {
        WOp op;
        std::memset(&op, 0, sizeof(op));
        op.code = kCodeRegSwap + 8 * std::min(j1, j2) + std::max(j1, j2);
        push_op(op);
        std::swap(reg_[j1], reg_[j2]);
    }

    uint32_t push_matrix(const std::vector<cplx>& m) {
        const uint32_t at = static_cast<uint32_t>(mats_.size());
        for (const cplx& e : m) {
            V x;
            x.x = static_cast<R>(e.real());
            x.y = static_cast<R>(e.imag());
            mats_.push_back(x);
        }
        return at;
    }

    void emit_dense1(const Block& b, uint32_t j) {
        WOp op;
        std::memset(&op, 0, sizeof(op));
        op.code = (b.matrix[0].imag() == 0.0 ? kCodeDense1Re : kCodeDense1) + j;
        op.mat = push_matrix(b.matrix);
        push_op(op);
    }

    void emit_densek(const Block& b) {
        WOp op;
        std::memset(&op, 0, sizeof(op));
        op.code = kCodeDenseK + static_cast<uint32_t>(b.operands.size());
        op.mat = push_matrix(b.matrix);
        push_op(op);
    }

    void emit_diag(const Block& b) {
        const uint32_t k = static_cast<uint32_t>(b.operands.size());
        const std::size_t dim = std::size_t{1} << k;
        WOp op;
        std::memset(&op, 0, sizeof(op));
        op.code = sign_diagonal(b) ? kCodeDiagSign : kCodeDiag;
        std::memset(op.lane_bit, kNone, sizeof(op.lane_bit));
        for (uint32_t t = 0; t < k; ++t) {
            const uint32_t q = b.operands[t];
            const uint8_t p = pos_of_[q];
            if (p == kNone) {
                op.outer_q[op.num_outer] = static_cast<uint8_t>(q);
                op.outer_b[op.num_outer] = static_cast<uint8_t>(t);
                ++op.num_outer;
                continue;
            }
            const int li = lane_index(p);
            if (li >= 0) {
                op.lane_bit[li] = static_cast<uint8_t>(t);
                continue;
            }
            const int j = reg_index(p);
            for (uint32_t r = 0; r < E; ++r)
                if ((r >> j) & 1u) op.rpart[r] |= static_cast<uint8_t>(1u << t);
        }
        std::vector<cplx> d(dim);
        for (std::size_t e = 0; e < dim; ++e) {
            d[e] = b.matrix[e * dim + e];
            if (d[e].real() == -1.0) op.neg_mask |= 1ull << e;
        }
        op.mat = push_matrix(d);
        push_op(op);
    }

    std::vector<WOp>& ops_;
    std::vector<V>& mats_;
    std::vector<const Block*> blocks_;
    uint8_t wq_[16] = {};
    uint8_t pos_of_[64] = {};
    uint8_t lane_[kLaneBits] = {};
    uint8_t reg_[8] = {};
};

template <typename R>
void launch_window_passes(void* state, uint32_t n, const std::vector<Block>& blocks,
                          const std::vector<WindowPass>& passes, bool time_passes, PassStats* stats) {
    using V = typename Shape<R>::V;
    constexpr uint32_t W = width_of<R>();
    Context& ctx = context();
    std::vector<WPassArgs> args(passes.size());
    std::vector<uint32_t> op_begin(passes.size(), 0);
    std::vector<WOp> ops;
    std::vector<V> mats;
    PassCompiler<R> compiler(ops, mats);
    for (std::size_t p = 0; p < passes.size(); ++p) {
        op_begin[p] = static_cast<uint32_t>(ops.size());
        if (passes[p].generic) continue;
        const bool last = p + 1 == passes.size();
        args[p] = compiler.compile(n, blocks, passes[p], last && sizeof(R) == 8);
        args[p].num_ops = static_cast<uint32_t>(ops.size()) - op_begin[p];
    }
    DeviceBuffer dops(std::max<std::size_t>(ops.size(), 1) * sizeof(WOp));
    DeviceBuffer dmats(std::max<std::size_t>(mats.size(), 1) * sizeof(V));
    upload(dops.get(), ops.data(), ops.size() * sizeof(WOp));
    upload(dmats.get(), mats.data(), mats.size() * sizeof(V));

    std::vector<hipEvent_t> events;
    if (time_passes) {
        events.resize(passes.size() + 1);
        for (auto& e : events) MH_HIP(hipEventCreate(&e));
        MH_HIP(hipEventRecord(events[0], ctx.stream));
    }
    for (std::size_t p = 0; p < passes.size(); ++p) {
        if (passes[p].generic) {
            const Block& blk = blocks[passes[p].blocks[0]];
            launch_generic(state, n, sizeof(R) == 8 ? MH_C128 : MH_C64, blk.matrix.data(), blk.operands.data(),
                           static_cast<uint32_t>(blk.operands.size()));
        } else {
            WPassArgs a = args[p];
            a.state = state;
            a.ops = dops.as<WOp>() + op_begin[p];
            a.mats = dmats.get();
            if (n < W || n - W > 30) fail(MH_ERR_INTERNAL, "window pass shape out of range");
            const uint32_t waves_log = n - W;
            const uint32_t group_log = std::min<uint32_t>(waves_log, 2);
            const dim3 grid(1u << (waves_log - group_log)), block(64u << group_log);
            hipLaunchKernelGGL(window_kernel<R>, grid, block, 0, ctx.stream, a);
            MH_HIP(hipGetLastError());
        }
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

}  // namespace

uint32_t window_width(int32_t precision) noexcept {
    return precision == MH_C128 ? width_of<double>() : width_of<float>();
}

std::vector<WindowPass> plan_window_passes(uint32_t n, int32_t precision, const std::vector<Block>& blocks,
                                           uint32_t w, uint32_t max_cost) {
    const uint32_t width = window_width(precision);
    const uint32_t low = precision == MH_C128 ? Shape<double>::kLow : Shape<float>::kLow;
    const uint64_t low_mask = (1ull << low) - 1ull;
    w = std::min(w, width);
    std::vector<WindowPass> passes;
    WindowPass cur;
    bool open = false;
    uint32_t cost = 0;
    auto close = [&] {
        if (open) passes.push_back(cur);
        cur = WindowPass{};
        open = false;
        cost = 0;
    };
    for (uint32_t b = 0; b < blocks.size(); ++b) {
        const Block& blk = blocks[b];
        if (blk.operands.empty()) continue;
        const uint32_t k = static_cast<uint32_t>(blk.operands.size());
        if (n < width || (!blk.diagonal && k > kMaxWindowDense)) {
            // A state smaller than one wave's window, or a dense block of
            // more than two qubits: a pass of its own on the generic kernel.
            close();
            WindowPass g;
            g.generic = true;
            g.blocks.push_back(b);
            passes.push_back(g);
            continue;
        }
        const uint32_t c = block_cost(blk);
        const uint64_t dense = blk.diagonal ? 0 : mask_of(blk.operands);
        const uint64_t want = cur.dense_mask | dense;
        const bool fits = static_cast<uint32_t>(__builtin_popcountll(want | low_mask)) <= w;
        if (open && fits && cost + c <= max_cost) {
            cur.dense_mask = want;
            if (!blk.diagonal) cur.kmax = std::max(cur.kmax, k);
            cur.blocks.push_back(b);
            cost += c;
            continue;
        }
        close();
        if (static_cast<uint32_t>(__builtin_popcountll(dense | low_mask)) > w)
            fail(MH_ERR_INVALID, "a dense block does not fit the window beside the low run");
        cur.dense_mask = dense;
        cur.kmax = blk.diagonal ? 0 : k;
        cur.blocks.push_back(b);
        cost = c;
        open = true;
    }
    close();
    return passes;
}

void run_window_passes(void* state, uint32_t n, int32_t precision, const std::vector<Block>& blocks, uint32_t w,
                       uint32_t max_cost, bool time_passes, PassStats* stats) {
    const std::vector<WindowPass> passes = plan_window_passes(n, precision, blocks, w, max_cost);
    if (stats) stats->passes = static_cast<uint32_t>(passes.size());
    if (passes.empty()) return;
    if (precision == MH_C128)
        launch_window_passes<double>(state, n, blocks, passes, time_passes, stats);
    else
        launch_window_passes<float>(state, n, blocks, passes, time_passes, stats);
}

}  // namespace mh
