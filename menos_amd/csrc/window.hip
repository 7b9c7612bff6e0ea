// The gate-pass kernel: many plan blocks per HBM read and write of the state.
//
// Each wavefront takes one window of W qubits (the other qubits fixed by the
// wave's index) and holds its 2^W amplitudes in registers: lane bit i of the
// wave's 64 lanes and register bit j of a lane's 2^B amplitudes (B = 4, so
// W = 10) each stand for one window qubit. A
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
#include <chrono>
#include <atomic>
#include <cstdint>
#include <cstring>
#include <exception>
#include <mutex>
#include <thread>

#include "mh_internal.h"
#include "batch_program.h"
#include "window.h"

namespace mh {
namespace {

constexpr uint32_t kLaneBits = 6;
constexpr uint32_t kWavesPerGroup = 4;
constexpr uint32_t kNone = 255;
constexpr uint32_t kMaxWindowDense = 2;  // widest dense block a window pass applies

template <typename R> struct Shape;
template <> struct Shape<double> {
    static constexpr uint32_t kRegBits = 4, kLow = 3;
    using V = double2;
};
template <> struct Shape<float> {
    static constexpr uint32_t kRegBits = 4, kLow = 4;
    using V = float2;
};

// A pass is a list of bundles; a bundle is a 64-bit set of stages that run in
// the fixed order of their bit indices (B register bits):
//   0 sign diagonal, 1 diagonal,
//   for each register bit J, at 2 + 12J: lane bit I <-> J exchange (I = 0..5),
//     then a 1-qubit block on J whose entry (0, 0) is real, then any 1-qubit
//     block on J, then a controlled pair on J whose control is a lane bit or
//     a qubit outside the window, then one whose control is register bit K
//     (one slot per K != J),
//   2 + 12B: 2-qubit block on register bits 0, 1; 3 + 12B sign diagonal,
//   4 + 12B diagonal; 5 + 12B + p: swap of register-bit pair p.
// A controlled pair is a 2-qubit block that never mixes the two values of one
// operand (the control): a 2x2 block on the other operand, chosen by the
// control's bit.
// A loop over bundles whose body is this fixed chain of independent stages
// keeps every amplitude in its registers from stage to stage; a switch over
// op kinds instead makes the compiler move the whole window between
// registers on every op.
struct Stages {
    static constexpr uint32_t kSignEarly = 0, kDiagEarly = 1, kGroupSlots = 12;
    static constexpr uint32_t group(uint32_t j) { return 2 + kGroupSlots * j; }
    static constexpr uint32_t lane_swap(uint32_t i, uint32_t j) { return group(j) + i; }
    static constexpr uint32_t dense1_re(uint32_t j) { return group(j) + 6; }
    static constexpr uint32_t dense1(uint32_t j) { return group(j) + 7; }
    static constexpr uint32_t pair_var(uint32_t j) { return group(j) + 8; }
    static constexpr uint32_t pair_reg(uint32_t j, uint32_t k) { return group(j) + 9 + (k < j ? k : k - 1); }
    static constexpr uint32_t dense2(uint32_t rb) { return 2 + kGroupSlots * rb; }
    static constexpr uint32_t sign_late(uint32_t rb) { return 3 + kGroupSlots * rb; }
    static constexpr uint32_t diag_late(uint32_t rb) { return 4 + kGroupSlots * rb; }
    static constexpr uint32_t pair_index(uint32_t j1, uint32_t j2, uint32_t rb) {
        uint32_t p = 0;
        for (uint32_t a = 0; a < j1; ++a) p += rb - 1 - a;
        return p + (j2 - j1 - 1);
    }
    static constexpr uint32_t reg_swap(uint32_t j1, uint32_t j2, uint32_t rb) { return 5 + kGroupSlots * rb + pair_index(j1, j2, rb); }
};
static_assert(Stages::reg_swap(2, 3, 4) < 64, "stages of one bundle fit 64 bits");

// Every controlled-pair stage bit of a bundle (rb register bits).
constexpr uint64_t pair_stage_bits(uint32_t rb) {
    uint64_t m = 0;
    for (uint32_t j = 0; j < rb; ++j)
        for (uint32_t slot = 8; slot < Stages::kGroupSlots; ++slot) m |= 1ull << (Stages::group(j) + slot);
    return m;
}

bool uses_pairs(const uint64_t* bundles, std::size_t count, uint32_t rb) {
    const uint64_t m = pair_stage_bits(rb);
    for (std::size_t b = 0; b < count; ++b)
        if (bundles[b] & m) return true;
    return false;
}
static_assert(Stages::pair_reg(0, 3) < Stages::group(1) && Stages::pair_reg(3, 2) < Stages::dense2(4),
              "a register bit's stages fit its group");

}  // namespace

// The parameters of a stage that has them (blocks, not exchanges), in stage
// order within a bundle. A sign stage applies a run of consecutive sign
// diagonals at once: sign flips commute and compose exactly, so the run is
// one XOR mask over the lane's amplitudes.
struct WOp {
    uint32_t mat;            // first entry in the matrix table; sign stage: first term
    uint32_t num_outer;      // diagonal: operands outside the window; sign stage: terms;
                             // controlled pair on a register control: 1 for a controlled NOT
    uint64_t neg_mask;       // sign diagonal: bit e set when entry e is -1
    uint8_t lane_bit[8];     // diagonal: operand bit that lane bit i supplies, or kNone;
                             // controlled pair: [0] the control's lane bit, or kNone
    uint8_t outer_q[8];      // diagonal: global qubit of the m-th operand outside the window;
                             // controlled pair: [0] the control when it is outside the window
    uint8_t outer_b[8];      // ... and its operand bit
    uint8_t rpart[32];       // diagonal: operand bits that register index r supplies
};

// One -1 entry of a sign diagonal in a sign stage, resolved on the host
// against where each operand lives: the amplitudes it negates are those whose
// outside-window qubits, lane bits and register bits all match the entry.
struct SignTerm {
    uint64_t omask, oval;    // outside the window: (wave base & omask) == oval
    uint32_t rpat;           // register indices r that match (bit r)
    uint8_t lmask, lval;     // lane bits: (lane & lmask) == lval
    uint8_t pad[2];
};
static_assert(sizeof(SignTerm) == 24, "sign term layout");

// Where a pass's window qubits live at the load and the store, and which
// outside qubits the wave index selects.
struct WLayout {
    uint32_t num_bundles;
    uint32_t normalize;       // add +0 to every component before the store
    uint32_t num_outer;       // qubits outside the window
    uint32_t num_waves;       // windows of the state: 2^(n - W)
    uint8_t outer_q[32];      // wave index bit b -> global qubit outer_q[b]
    uint8_t load_lane_q[8], load_reg_q[8];    // global qubit of each lane / register bit at the load
    uint8_t store_lane_q[8], store_reg_q[8];  // ... and at the store
};

struct WPassArgs {
    void* state;
    const uint64_t* bundles;
    const WOp* ops;
    const SignTerm* terms;
    const void* mats;
    WLayout lay;
};

// One state's pass in a batched launch: offsets from the launch's base pointers.
// generate: the pass starts from the initial state phase * |0...0> (the phase
// at mats[mat_off + phase_at]) instead of loading the state.
struct WBatchArgs {
    uint64_t state_off;       // amplitudes
    uint32_t bundle_off, op_off, term_off, mat_off;
    uint32_t generate, phase_at;
    WLayout lay;
};

namespace {

// The qubits outside the window, taken from the wave index (a bit deposit).
__device__ __forceinline__ uint64_t outer_base(uint32_t wave, const WLayout& lay) {
    uint64_t base = 0;
    for (uint32_t b = 0; b < lay.num_outer; ++b) base |= uint64_t((wave >> b) & 1u) << lay.outer_q[b];
    return base;
}

// The state streams through once per pass: non-temporal loads and stores
// keep it from displacing the op tables and the other waves' lines in the caches.
typedef double VD2 __attribute__((ext_vector_type(2)));
typedef float VF2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ double2 load_amp(const double2* p) {
    const VD2 x = __builtin_nontemporal_load(reinterpret_cast<const VD2*>(p));
    return make_double2(x[0], x[1]);
}

__device__ __forceinline__ float2 load_amp(const float2* p) {
    const VF2 x = __builtin_nontemporal_load(reinterpret_cast<const VF2*>(p));
    return make_float2(x[0], x[1]);
}

__device__ __forceinline__ void store_amp(double2* p, double2 v) {
    VD2 x;
    x[0] = v.x;
    x[1] = v.y;
    __builtin_nontemporal_store(x, reinterpret_cast<VD2*>(p));
}

__device__ __forceinline__ void store_amp(float2* p, float2 v) {
    VF2 x;
    x[0] = v.x;
    x[1] = v.y;
    __builtin_nontemporal_store(x, reinterpret_cast<VF2*>(p));
}

// Ordinary (cached) accesses, for states small enough to stay in the caches
// from one pass to the next.
template <bool kTemporal, typename V>
__device__ __forceinline__ V load_state(const V* p) {
    if constexpr (kTemporal)
        return *p;
    else
        return load_amp(p);
}

template <bool kTemporal, typename V>
__device__ __forceinline__ void store_state(V* p, V v) {
    if constexpr (kTemporal)
        *p = v;
    else
        store_amp(p, v);
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
    } else if constexpr (I >= 2) {
        // Lane bits 2 and 3 select banks of four lanes within a row of 16, so
        // the DPP bank mask picks the lanes that take a partner's value: those
        // with bit I set take y from lane l - 2^I (row_shr), the others x from
        // lane l + 2^I (row_shl); a disabled lane keeps its own value.
        constexpr int shr = I == 2 ? 0x114 : 0x118, shl = I == 2 ? 0x104 : 0x108;
        constexpr int upper_banks = I == 2 ? 0xA : 0xC, lower_banks = I == 2 ? 0x5 : 0x3;
        const uint32_t nx = __builtin_amdgcn_update_dpp(x, y, shr, 0xF, upper_banks, false);
        const uint32_t ny = __builtin_amdgcn_update_dpp(y, x, shl, 0xF, lower_banks, false);
        (void)upper;
        x = nx;
        y = ny;
    } else {
        // Lanes with bit I set take y from lane l - 2^I, the others x from
        // lane l + 2^I: quad_perm lane ^ 1 (0xB1) or lane ^ 2 (0x4E).
        constexpr int from_lower = I == 0 ? 0xB1 : 0x4E;
        constexpr int from_upper = I == 0 ? 0xB1 : 0x4E;
        const uint32_t fy = __builtin_amdgcn_mov_dpp(y, from_lower, 0xF, 0xF, true);
        const uint32_t fx = __builtin_amdgcn_mov_dpp(x, from_upper, 0xF, 0xF, true);
        const uint32_t nx = upper ? fy : x;
        const uint32_t ny = upper ? y : fx;
        x = nx;
        y = ny;
    }
}

template <uint32_t I>
__device__ __forceinline__ void xchg(double& x, double& y, bool upper) {
    uint32_t xl = static_cast<uint32_t>(__double2loint(x)), xh = static_cast<uint32_t>(__double2hiint(x));
    uint32_t yl = static_cast<uint32_t>(__double2loint(y)), yh = static_cast<uint32_t>(__double2hiint(y));
    xchg<I>(xl, yl, upper);
    xchg<I>(xh, yh, upper);
    x = __hiloint2double(static_cast<int>(xh), static_cast<int>(xl));
    y = __hiloint2double(static_cast<int>(yh), static_cast<int>(yl));
}

template <uint32_t I>
__device__ __forceinline__ void xchg(float& x, float& y, bool upper) {
    uint32_t xb = __float_as_uint(x), yb = __float_as_uint(y);
    xchg<I>(xb, yb, upper);
    x = __uint_as_float(xb);
    y = __uint_as_float(yb);
}

// Lane bit I <-> register bit J.
template <typename V, uint32_t E, uint32_t I, uint32_t J>
__device__ __forceinline__ void lane_swap(V* v, uint32_t lane) {
    const bool upper = (lane >> I) & 1u;
#pragma unroll
    for (uint32_t r = 0; r < E; ++r) {
        if (r & (1u << J)) continue;
        xchg<I>(v[r].x, v[r | (1u << J)].x, upper);
        xchg<I>(v[r].y, v[r | (1u << J)].y, upper);
    }
}

// Register bit J1 <-> register bit J2 (a renaming of the amplitudes).
template <typename V, uint32_t E, uint32_t J1, uint32_t J2>
__device__ __forceinline__ void reg_swap(V* v) {
#pragma unroll
    for (uint32_t r = 0; r < E; ++r) {
        if (!((r >> J1) & 1u) || ((r >> J2) & 1u)) continue;
        const uint32_t s = r ^ (1u << J1) ^ (1u << J2);
        const V t = v[r];
        v[r] = v[s];
        v[s] = t;
    }
}

// One pair (x0: the operand's bit clear, x1: set) under a 2x2 block. Every
// product is formed before any sum, so the outputs are defined after the last
// read of the inputs and take over their registers.
template <bool kRe>
__device__ __forceinline__ void pair1(double2& x0, double2& x1, double2 m00, double2 m01, double2 m10, double2 m11) {
    const double2 a0 = x0, a1 = x1;
    const double p0 = m00.x * a0.x, p2 = m00.x * a0.y;
    const double p1 = kRe ? 0.0 : m00.y * a0.y, p3 = kRe ? 0.0 : m00.y * a0.x;
    const double p4 = m01.x * a1.x, p5 = m01.y * a1.y, p6 = m01.x * a1.y, p7 = m01.y * a1.x;
    const double p8 = m10.x * a0.x, p9 = m10.y * a0.y, pa = m10.x * a0.y, pb = m10.y * a0.x;
    const double pc = m11.x * a1.x, pd = m11.y * a1.y, pe = m11.x * a1.y, pf = m11.y * a1.x;
    const double s0r = (kRe ? p0 : p0 - p1) + (p4 - p5);
    const double s0i = (kRe ? p2 : p2 + p3) + (p6 + p7);
    const double s1r = (p8 - p9) + (pc - pd);
    const double s1i = (pa + pb) + (pe + pf);
    x0 = make_double2(s0r, s0i);
    x1 = make_double2(s1r, s1i);
}

// (sr, si) += m * a in FP32 with fused multiply-adds (the complex64 form is
// held to a tolerance, not to the reference's bits).
__device__ __forceinline__ void cmac(float& sr, float& si, float2 m, float2 a) {
    sr = __builtin_fmaf(m.x, a.x, sr);
    sr = __builtin_fmaf(-m.y, a.y, sr);
    si = __builtin_fmaf(m.x, a.y, si);
    si = __builtin_fmaf(m.y, a.x, si);
}

template <bool kRe>
__device__ __forceinline__ void pair1(float2& x0, float2& x1, float2 m00, float2 m01, float2 m10, float2 m11) {
    const float2 a0 = x0, a1 = x1;
    float s0r = m00.x * a0.x, s0i = m00.x * a0.y;
    if constexpr (!kRe) {
        s0r = __builtin_fmaf(-m00.y, a0.y, s0r);
        s0i = __builtin_fmaf(m00.y, a0.x, s0i);
    }
    cmac(s0r, s0i, m01, a1);
    float s1r = 0.0f, s1i = 0.0f;
    cmac(s1r, s1i, m10, a0);
    cmac(s1r, s1i, m11, a1);
    x0 = make_float2(s0r, s0i);
    x1 = make_float2(s1r, s1i);
}

// One 1-qubit block on register bit J.
template <uint32_t E, uint32_t J, bool kRe, typename V>
__device__ __forceinline__ void dense1(V* v, const V* m) {
    const V m00 = m[0], m01 = m[1], m10 = m[2], m11 = m[3];
#pragma unroll
    for (uint32_t r = 0; r < E; ++r) {
        if (r & (1u << J)) continue;
        pair1<kRe>(v[r], v[r | (1u << J)], m00, m01, m10, m11);
    }
}

// A controlled pair on register bit J whose control is register bit K:
// m[0..3] is the block where the control is clear, m[4..7] where it is set.
// A controlled NOT (identity, then the exchange) only exchanges amplitudes:
// a product by an exact 1 or 0 and a sum with a signed zero change no value
// (up to the sign of a zero).
template <uint32_t E, uint32_t J, uint32_t K, typename V>
__device__ __forceinline__ void pair_reg(V* v, const V* m, bool cnot) {
    if (cnot) {
#pragma unroll
        for (uint32_t r = 0; r < E; ++r) {
            if ((r & (1u << J)) || !((r >> K) & 1u)) continue;
            const V t = v[r];
            v[r] = v[r | (1u << J)];
            v[r | (1u << J)] = t;
        }
        return;
    }
    const V a00 = m[0], a01 = m[1], a10 = m[2], a11 = m[3];
    const V b00 = m[4], b01 = m[5], b10 = m[6], b11 = m[7];
#pragma unroll
    for (uint32_t r = 0; r < E; ++r) {
        if (r & (1u << J)) continue;
        if ((r >> K) & 1u)
            pair1<false>(v[r], v[r | (1u << J)], b00, b01, b10, b11);
        else
            pair1<false>(v[r], v[r | (1u << J)], a00, a01, a10, a11);
    }
}

__device__ __forceinline__ double2 pick(bool set, double2 a, double2 b) {
    return make_double2(set ? b.x : a.x, set ? b.y : a.y);
}

__device__ __forceinline__ float2 pick(bool set, float2 a, float2 b) {
    return make_float2(set ? b.x : a.x, set ? b.y : a.y);
}

// A controlled pair on register bit J whose control bit `set` is the same for
// all of the lane's amplitudes (a lane bit, or a qubit outside the window).
template <uint32_t E, uint32_t J, typename V>
__device__ __forceinline__ void pair_var(V* v, const V* m, bool set) {
    const V m00 = pick(set, m[0], m[4]), m01 = pick(set, m[1], m[5]);
    const V m10 = pick(set, m[2], m[6]), m11 = pick(set, m[3], m[7]);
#pragma unroll
    for (uint32_t r = 0; r < E; ++r) {
        if (r & (1u << J)) continue;
        pair1<false>(v[r], v[r | (1u << J)], m00, m01, m10, m11);
    }
}

// A 2-qubit block on register bits 0 (operand 0) and 1 (operand 1).
template <uint32_t E>
__device__ __forceinline__ void dense2(double2* v, const double2* m) {
#pragma unroll
    for (uint32_t hi = 0; hi < (E >> 2); ++hi) {
        double2 a[4];
#pragma unroll
        for (uint32_t in = 0; in < 4; ++in) a[in] = v[in | (hi << 2)];
#pragma unroll
        for (uint32_t o = 0; o < 4; ++o) {
            const double2 e0 = m[o * 4];
            double sr = e0.x * a[0].x - e0.y * a[0].y;
            double si = e0.x * a[0].y + e0.y * a[0].x;
#pragma unroll
            for (uint32_t in = 1; in < 4; ++in) {
                const double2 e = m[o * 4 + in];
                sr = sr + (e.x * a[in].x - e.y * a[in].y);
                si = si + (e.x * a[in].y + e.y * a[in].x);
            }
            v[o | (hi << 2)] = make_double2(sr, si);
        }
    }
}

template <uint32_t E>
__device__ __forceinline__ void dense2(float2* v, const float2* m) {
#pragma unroll
    for (uint32_t hi = 0; hi < (E >> 2); ++hi) {
        float2 a[4];
#pragma unroll
        for (uint32_t in = 0; in < 4; ++in) a[in] = v[in | (hi << 2)];
#pragma unroll
        for (uint32_t o = 0; o < 4; ++o) {
            float sr = 0.0f, si = 0.0f;
#pragma unroll
            for (uint32_t in = 0; in < 4; ++in) cmac(sr, si, m[o * 4 + in], a[in]);
            v[o | (hi << 2)] = make_float2(sr, si);
        }
    }
}

__device__ __forceinline__ double2 diag_mul(double2 d, double2 a) {
    const double pr = d.x * a.x, qr = d.y * a.y, pi = d.x * a.y, qi = d.y * a.x;
    return make_double2(pr - qr, pi + qi);
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

// A run of sign diagonals: bit r of m tells whether amplitude r changes sign.
// Each -1 entry of each diagonal adds (XOR) its register pattern where its
// outside-window qubits (the same for the whole wave) and lane bits match:
// entries that name no lane bit are one scalar test each.
template <typename V, uint32_t E>
__device__ __forceinline__ void apply_signs(V* v, const WOp& op, const SignTerm* terms, uint64_t obase,
                                            uint32_t lane) {
    constexpr uint32_t full = E == 32 ? 0xFFFFFFFFu : (1u << E) - 1u;
    uint32_t mu = 0, ml = 0;
    for (uint32_t t = 0; t < op.num_outer; ++t) {
        const SignTerm& st = terms[op.mat + t];
        if ((obase & st.omask) != st.oval) continue;
        if (st.lmask == 0)
            mu ^= st.rpat;
        else
            ml ^= (lane & st.lmask) == st.lval ? st.rpat : 0u;
    }
    const uint32_t m = (mu ^ ml) & full;
#pragma unroll
    for (uint32_t r = 0; r < E; ++r) flip_sign(v[r], (m >> r) & 1u);
}

template <typename V, uint32_t E>
__device__ __forceinline__ void apply_diag(V* v, const WOp& op, const V* mats, uint64_t obase, uint32_t lane) {
    const V* d = mats + op.mat;
    const uint32_t t = diag_base(op, obase, lane);
    // Entries gathered four at a time (bounds the registers they hold).
#pragma unroll
    for (uint32_t r = 0; r < E; ++r) {
        v[r] = diag_mul(d[t | op.rpart[r]], v[r]);
        if ((r & 3u) == 3u) __builtin_amdgcn_sched_barrier(0);
    }
}

template <typename V, uint32_t E, uint32_t RB, uint32_t J, uint32_t K>
__device__ __forceinline__ void pair_reg_stage(uint64_t c, V* v, const WOp*& op, const V* mats) {
    if constexpr (K != J && K < RB)
        if (c & (1ull << Stages::pair_reg(J, K))) {
            const WOp& o = *op++;
            pair_reg<E, J, K>(v, mats + o.mat, o.num_outer != 0);
        }
}

// The stages of register bit J (exchanges into J, then its 1-qubit blocks and
// controlled pairs).
template <typename V, uint32_t E, uint32_t RB, uint32_t J, bool kPairs>
__device__ __forceinline__ void group_stages(uint64_t c, V* v, const WOp*& op, const V* mats, uint64_t obase,
                                             uint32_t lane) {
    if constexpr (J < RB) {
        if (!(c & ((kPairs ? 0xFFFull : 0xFFull) << Stages::group(J)))) return;
        if (c & (1ull << Stages::lane_swap(0, J))) lane_swap<V, E, 0, J>(v, lane);
        if (c & (1ull << Stages::lane_swap(1, J))) lane_swap<V, E, 1, J>(v, lane);
        if (c & (1ull << Stages::lane_swap(2, J))) lane_swap<V, E, 2, J>(v, lane);
        if (c & (1ull << Stages::lane_swap(3, J))) lane_swap<V, E, 3, J>(v, lane);
        if (c & (1ull << Stages::lane_swap(4, J))) lane_swap<V, E, 4, J>(v, lane);
        if (c & (1ull << Stages::lane_swap(5, J))) lane_swap<V, E, 5, J>(v, lane);
        if (c & (1ull << Stages::dense1_re(J))) dense1<E, J, true>(v, mats + (op++)->mat);
        if (c & (1ull << Stages::dense1(J))) dense1<E, J, false>(v, mats + (op++)->mat);
        if constexpr (!kPairs) return;
        if (c & (1ull << Stages::pair_var(J))) {
            const WOp& o = *op++;
            const bool set = o.lane_bit[0] != kNone ? ((lane >> o.lane_bit[0]) & 1u) != 0
                                                    : ((obase >> o.outer_q[0]) & 1ull) != 0;
            pair_var<E, J>(v, mats + o.mat, set);
        }
        pair_reg_stage<V, E, RB, J, 0>(c, v, op, mats);
        pair_reg_stage<V, E, RB, J, 1>(c, v, op, mats);
        pair_reg_stage<V, E, RB, J, 2>(c, v, op, mats);
        pair_reg_stage<V, E, RB, J, 3>(c, v, op, mats);
        pair_reg_stage<V, E, RB, J, 4>(c, v, op, mats);
    }
}

template <typename V, uint32_t E, uint32_t RB, uint32_t J1, uint32_t J2>
__device__ __forceinline__ void reg_swap_stage(uint64_t c, V* v) {
    if constexpr (J2 < RB)
        if (c & (1ull << Stages::reg_swap(J1, J2, RB))) reg_swap<V, E, J1, J2>(v);
}

// One wave's window of one pass. With `initial` set the window starts from
// the initial state instead of a load: *initial at index 0 (wave 0, lane 0,
// register 0: every index bit clear) and +0 everywhere else, the values the
// initialisation kernel would have written.
template <typename R, bool kPairs, bool kTemporal = false>
__device__ __forceinline__ void window_body(typename Shape<R>::V* state, const uint64_t* bundles, const WOp* ops,
                                            const SignTerm* terms, const typename Shape<R>::V* mats,
                                            const WLayout& a, uint32_t wave, uint32_t lane,
                                            const typename Shape<R>::V* initial = nullptr) {
    using V = typename Shape<R>::V;
    constexpr uint32_t RB = Shape<R>::kRegBits;
    constexpr uint32_t E = 1u << RB;
    const uint64_t obase = outer_base(wave, a);
    V* const st = state + obase;

    V v[E];
    if (initial != nullptr) {
        const V first = *initial;
        const bool origin = wave == 0 && lane == 0;
#pragma unroll
        for (uint32_t r = 0; r < E; ++r) {
            v[r].x = r == 0 && origin ? first.x : R(0);
            v[r].y = r == 0 && origin ? first.y : R(0);
        }
    } else {
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
            v[r] = load_state<kTemporal>(src + h);
        }
    }

    const WOp* op = ops;
    for (uint32_t b = 0; b < a.num_bundles; ++b) {
        const uint64_t c = bundles[b];
        if (c & (1ull << Stages::kSignEarly)) apply_signs<V, E>(v, *op++, terms, obase, lane);
        if (c & (1ull << Stages::kDiagEarly)) apply_diag<V, E>(v, *op++, mats, obase, lane);
        group_stages<V, E, RB, 0, kPairs>(c, v, op, mats, obase, lane);
        group_stages<V, E, RB, 1, kPairs>(c, v, op, mats, obase, lane);
        group_stages<V, E, RB, 2, kPairs>(c, v, op, mats, obase, lane);
        group_stages<V, E, RB, 3, kPairs>(c, v, op, mats, obase, lane);
        group_stages<V, E, RB, 4, kPairs>(c, v, op, mats, obase, lane);
        if (c & (1ull << Stages::dense2(RB))) dense2<E>(v, mats + (op++)->mat);
        if (c & (1ull << Stages::sign_late(RB))) apply_signs<V, E>(v, *op++, terms, obase, lane);
        if (c & (1ull << Stages::diag_late(RB))) apply_diag<V, E>(v, *op++, mats, obase, lane);
        if (c >> Stages::reg_swap(0, 1, RB)) {
            reg_swap_stage<V, E, RB, 0, 1>(c, v);
            reg_swap_stage<V, E, RB, 0, 2>(c, v);
            reg_swap_stage<V, E, RB, 0, 3>(c, v);
            reg_swap_stage<V, E, RB, 0, 4>(c, v);
            reg_swap_stage<V, E, RB, 1, 2>(c, v);
            reg_swap_stage<V, E, RB, 1, 3>(c, v);
            reg_swap_stage<V, E, RB, 1, 4>(c, v);
            reg_swap_stage<V, E, RB, 2, 3>(c, v);
            reg_swap_stage<V, E, RB, 2, 4>(c, v);
            reg_swap_stage<V, E, RB, 3, 4>(c, v);
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
            store_state<kTemporal>(dst + h, v[r]);
        }
    }
}

// Four waves per SIMD: each wave holds its window in at most 128 VGPRs.
// kPairs: the pass has controlled-pair stages (the form without them is the
// leaner code for passes that need none).
template <typename R, bool kPairs>
__global__ void __launch_bounds__(kWavesPerGroup * 64, 4) window_kernel(WPassArgs a) {
    using V = typename Shape<R>::V;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
    window_body<R, kPairs>(static_cast<V*>(a.state), a.bundles, a.ops, a.terms, static_cast<const V*>(a.mats), a.lay,
                           wave, threadIdx.x & 63u);
}

// One pass of many states in one launch: workgroup b belongs to the state c
// with group_begin[c] <= b < group_begin[c + 1] (each state's windows in
// groups of four waves); a wave past its state's last window returns.
template <typename R, bool kPairs, bool kTemporal>
__global__ void __launch_bounds__(kWavesPerGroup * 64, 4)
    window_batch_kernel(const WBatchArgs* __restrict__ args, const uint32_t* __restrict__ group_begin, uint32_t count,
                        typename Shape<R>::V* state, const uint64_t* __restrict__ bundles,
                        const WOp* __restrict__ ops, const SignTerm* __restrict__ terms,
                        const typename Shape<R>::V* __restrict__ mats) {
    const uint32_t b = blockIdx.x;
    uint32_t lo = 0, hi = count;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (group_begin[mid] <= b) {
            lo = mid;
        } else {
            hi = mid;
        }
    }
    const WBatchArgs& a = args[lo];
    const uint32_t wave =
        __builtin_amdgcn_readfirstlane((b - group_begin[lo]) * kWavesPerGroup + (threadIdx.x >> 6));
    if (wave >= a.lay.num_waves) return;
    const typename Shape<R>::V* m = mats + a.mat_off;
    window_body<R, kPairs, kTemporal>(state + a.state_off, bundles + a.bundle_off, ops + a.op_off, terms + a.term_off,
                                      m, a.lay, wave, threadIdx.x & 63u, a.generate ? m + a.phase_at : nullptr);
}

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

// The control operand (0 or 1) of a 2-qubit block that never mixes that
// operand's two values, i.e. every entry between rows and columns that differ
// in its bit is zero, or -1. Such a block is a 2x2 block on the other operand
// chosen by the control's bit (CX, CZ or a controlled rotation with 1-qubit
// gates on its target fused in). The reference sums all four products of a
// row in column order; the two it skips here are a zero entry times an
// amplitude, a signed zero, and adding a signed zero leaves every other sum
// unchanged up to the sign of a zero result, so the pair form is the same
// arithmetic modulo zero signs, which the pass's final normalisation removes.
int pair_control(const Block& b) {
    if (b.diagonal || b.operands.size() != 2) return -1;
    for (uint32_t u = 0; u < 2; ++u) {
        bool ok = true;
        for (uint32_t out = 0; out < 4 && ok; ++out)
            for (uint32_t in = 0; in < 4 && ok; ++in)
                if (((out ^ in) >> u) & 1u) ok = b.matrix[out * 4 + in] == cplx(0.0, 0.0);
        if (ok) return static_cast<int>(u);
    }
    return -1;
}

// A controlled pair that is a controlled NOT: the identity where the control
// is clear, the exchange where it is set (entries exactly 0 and 1).
bool is_cnot(const std::vector<cplx>& m) {
    const cplx one(1.0, 0.0), zero(0.0, 0.0);
    return m[0] == one && m[1] == zero && m[2] == zero && m[3] == one && m[4] == zero && m[5] == one &&
           m[6] == one && m[7] == zero;
}

// The 2x2 blocks of a controlled pair with control operand u: where the
// control is clear (m[0..3]), then where it is set (m[4..7]), each in the
// target's row / column order.
std::vector<cplx> pair_blocks(const Block& b, uint32_t u) {
    std::vector<cplx> m(8);
    for (uint32_t cb = 0; cb < 2; ++cb)
        for (uint32_t to = 0; to < 2; ++to)
            for (uint32_t ti = 0; ti < 2; ++ti) {
                const uint32_t out = (cb << u) | (to << (1 - u)), in = (cb << u) | (ti << (1 - u));
                m[4 * cb + 2 * to + ti] = b.matrix[out * 4 + in];
            }
    return m;
}

// FP operations per amplitude a block costs in a pass (the pass cap unit),
// with an allowance for the exchanges that bring its operands to registers.
uint32_t block_cost(const Block& b) {
    const uint32_t k = static_cast<uint32_t>(b.operands.size());
    if (b.diagonal) return sign_diagonal(b) ? 2 : 6;
    if (k == 1) return (b.matrix[0].imag() == 0.0 ? 12 : 14) + 2;
    if (const int u = pair_control(b); u >= 0) return is_cnot(pair_blocks(b, static_cast<uint32_t>(u))) ? 8 : 14 + 4;
    return 2 * (3 * (1u << k) - 1) + 4;
}

// The qubits a non-diagonal block needs in registers: its operands, or a
// controlled pair's target alone.
uint64_t register_mask(const Block& b) {
    uint64_t m = 0;
    const int u = pair_control(b);
    for (uint32_t t = 0; t < b.operands.size(); ++t)
        if (static_cast<int>(t) != u) m |= 1ull << b.operands[t];
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

    PassCompiler(std::vector<uint64_t>& bundles, std::vector<WOp>& ops, std::vector<SignTerm>& terms,
                 std::vector<V>& mats)
        : bundles_(bundles), ops_(ops), terms_(terms), mats_(mats) {}

    WPassArgs compile(uint32_t n, const std::vector<Block>& blocks, const WindowPass& pass, bool normalize) {
        WPassArgs a;
        std::memset(&a, 0, sizeof(a));
        cur_ = 0;
        last_ = -1;
        open_sign_ = -1;
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
        // Wave index bits -> outside qubits: the lowest three first (eight
        // consecutive waves cover a contiguous kilobyte of every segment they
        // touch), then the others from the highest down. Measured on MI355X
        // at 33 qubits, this ordering holds a pass whose window qubits are
        // all above qubit 22 to about 60 ms (about 93 ms with plain ascending
        // order) and costs no more than a few percent on other windows.
        {
            uint8_t outside[64];
            uint32_t count_out = 0;
            for (uint32_t q = 0; q < n; ++q)
                if (!((wmask >> q) & 1ull)) outside[count_out++] = static_cast<uint8_t>(q);
            const uint32_t head = std::min<uint32_t>(3, count_out);
            for (uint32_t k = 0; k < head; ++k) a.lay.outer_q[a.lay.num_outer++] = outside[k];
            for (uint32_t k = count_out; k-- > head;) a.lay.outer_q[a.lay.num_outer++] = outside[k];
        }
        blocks_.clear();
        for (uint32_t b : pass.blocks) blocks_.push_back(&blocks[b]);
        initial_layout();
        for (uint32_t i = 0; i < kLaneBits; ++i) a.lay.load_lane_q[i] = wq_[lane_[i]];
        for (uint32_t j = 0; j < RB; ++j) a.lay.load_reg_q[j] = wq_[reg_[j]];

        for (std::size_t i = 0; i < blocks_.size(); ++i) {
            const Block& b = *blocks_[i];
            const uint32_t k = static_cast<uint32_t>(b.operands.size());
            if (b.diagonal) {
                emit_diag(b);
                continue;
            }
            const int u = pair_control(b);
            if (k == 1 || u >= 0) {
                const uint8_t p = pos_of_[b.operands[k == 1 ? 0 : 1 - u]];
                int j = reg_index(p);
                if (j < 0) {
                    j = victim(i, p);
                    lane_swap(static_cast<uint32_t>(lane_index(p)), static_cast<uint32_t>(j));
                }
                if (k == 1)
                    emit_dense1(b, static_cast<uint32_t>(j));
                else
                    emit_pair(b, static_cast<uint32_t>(u), static_cast<uint32_t>(j));
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
                emit_dense2(b);
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
        flush();
        for (uint32_t i = 0; i < kLaneBits; ++i) a.lay.store_lane_q[i] = wq_[lane_[i]];
        for (uint32_t j = 0; j < RB; ++j) a.lay.store_reg_q[j] = wq_[reg_[j]];
        a.lay.normalize = normalize ? 1u : 0u;
        return a;
    }

private:
    bool uses(std::size_t i, uint8_t p) const {
        const Block& b = *blocks_[i];
        if (b.diagonal) return false;
        const int u = pair_control(b);
        for (uint32_t t = 0; t < b.operands.size(); ++t)
            if (static_cast<int>(t) != u && pos_of_[b.operands[t]] == p) return true;
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
            const int u = pair_control(*b);
            for (uint32_t t = 0; t < b->operands.size(); ++t) {
                if (static_cast<int>(t) == u) continue;
                const uint8_t p = pos_of_[b->operands[t]];
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

    // Appends a stage to the open bundle when one of its positions comes after
    // the bundle's last stage, else starts a new bundle.
    void put(uint32_t first, int second, const WOp* rec) {
        uint32_t at = first;
        if (static_cast<int>(first) <= last_) {
            if (second > last_) {
                at = static_cast<uint32_t>(second);
            } else {
                flush();
            }
        }
        cur_ |= 1ull << at;
        last_ = static_cast<int>(at);
        open_sign_ = -1;
        if (rec) ops_.push_back(*rec);
    }

    void flush() {
        if (cur_) bundles_.push_back(cur_);
        cur_ = 0;
        last_ = -1;
    }

    void lane_swap(uint32_t i, uint32_t j) {
        put(Stages::lane_swap(i, j), -1, nullptr);
        std::swap(lane_[i], reg_[j]);
    }

    void reg_swap(uint32_t j1, uint32_t j2) {
        put(Stages::reg_swap(std::min(j1, j2), std::max(j1, j2), RB), -1, nullptr);
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
        op.mat = push_matrix(b.matrix);
        put(b.matrix[0].imag() == 0.0 ? Stages::dense1_re(j) : Stages::dense1(j), -1, &op);
    }

    // A controlled pair on register bit j.
    void emit_pair(const Block& b, uint32_t u, uint32_t j) {
        const std::vector<cplx> m = pair_blocks(b, u);
        WOp op;
        std::memset(&op, 0, sizeof(op));
        std::memset(op.lane_bit, kNone, sizeof(op.lane_bit));
        op.mat = push_matrix(m);
        op.num_outer = is_cnot(m) ? 1u : 0u;
        const uint8_t p = pos_of_[b.operands[u]];
        if (p == kNone) {
            op.outer_q[0] = static_cast<uint8_t>(b.operands[u]);
            put(Stages::pair_var(j), -1, &op);
            return;
        }
        const int li = lane_index(p);
        if (li >= 0) {
            op.lane_bit[0] = static_cast<uint8_t>(li);
            put(Stages::pair_var(j), -1, &op);
            return;
        }
        put(Stages::pair_reg(j, static_cast<uint32_t>(reg_index(p))), -1, &op);
    }

    void emit_dense2(const Block& b) {
        WOp op;
        std::memset(&op, 0, sizeof(op));
        op.mat = push_matrix(b.matrix);
        put(Stages::dense2(RB), -1, &op);
    }

    void emit_diag(const Block& b) {
        if (sign_diagonal(b)) {
            emit_sign(b);
            return;
        }
        const uint32_t k = static_cast<uint32_t>(b.operands.size());
        const std::size_t dim = std::size_t{1} << k;
        WOp op;
        std::memset(&op, 0, sizeof(op));
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
        for (std::size_t e = 0; e < dim; ++e) d[e] = b.matrix[e * dim + e];
        op.mat = push_matrix(d);
        put(Stages::kDiagEarly, static_cast<int>(Stages::diag_late(RB)), &op);
    }

    // A sign diagonal joins the sign stage just emitted, if nothing came after
    // it: one term per -1 entry, resolved against the current layout.
    void emit_sign(const Block& b) {
        const uint32_t k = static_cast<uint32_t>(b.operands.size());
        const std::size_t dim = std::size_t{1} << k;
        const uint32_t full = E == 32 ? 0xFFFFFFFFu : (1u << E) - 1u;
        uint32_t added = 0;
        for (std::size_t e = 0; e < dim; ++e) {
            if (b.matrix[e * dim + e].real() != -1.0) continue;
            SignTerm t;
            std::memset(&t, 0, sizeof(t));
            t.rpat = full;
            for (uint32_t u = 0; u < k; ++u) {
                const uint32_t q = b.operands[u];
                const bool one = (e >> u) & 1u;
                const uint8_t p = pos_of_[q];
                if (p == kNone) {
                    t.omask |= 1ull << q;
                    if (one) t.oval |= 1ull << q;
                } else if (const int li = lane_index(p); li >= 0) {
                    t.lmask |= static_cast<uint8_t>(1u << li);
                    if (one) t.lval |= static_cast<uint8_t>(1u << li);
                } else {
                    const int j = reg_index(p);
                    uint32_t pat = 0;
                    for (uint32_t r = 0; r < E; ++r)
                        if ((r >> j) & 1u) pat |= 1u << r;
                    t.rpat &= one ? pat : ~pat;
                }
            }
            terms_.push_back(t);
            ++added;
        }
        if (open_sign_ >= 0) {
            ops_[static_cast<std::size_t>(open_sign_)].num_outer += added;
            return;
        }
        WOp op;
        std::memset(&op, 0, sizeof(op));
        op.mat = static_cast<uint32_t>(terms_.size()) - added;
        op.num_outer = added;
        put(Stages::kSignEarly, static_cast<int>(Stages::sign_late(RB)), &op);
        open_sign_ = static_cast<int>(ops_.size()) - 1;
    }

    std::vector<uint64_t>& bundles_;
    std::vector<WOp>& ops_;
    std::vector<SignTerm>& terms_;
    std::vector<V>& mats_;
    uint64_t cur_ = 0;
    int last_ = -1;
    int open_sign_ = -1;  // index in ops_ of a sign stage that may take more terms
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
    context();
    std::vector<WPassArgs> args(passes.size());
    std::vector<uint32_t> bundle_begin(passes.size(), 0), op_begin(passes.size(), 0);
    std::vector<uint64_t> bundles;
    std::vector<WOp> ops;
    std::vector<SignTerm> terms;
    std::vector<V> mats;
    PassCompiler<R> compiler(bundles, ops, terms, mats);
    for (std::size_t p = 0; p < passes.size(); ++p) {
        bundle_begin[p] = static_cast<uint32_t>(bundles.size());
        op_begin[p] = static_cast<uint32_t>(ops.size());
        if (passes[p].generic) continue;
        const bool last = p + 1 == passes.size();
        args[p] = compiler.compile(n, blocks, passes[p], last && sizeof(R) == 8);
        args[p].lay.num_bundles = static_cast<uint32_t>(bundles.size()) - bundle_begin[p];
    }
    DeviceBuffer dbundles(std::max<std::size_t>(bundles.size(), 1) * sizeof(uint64_t));
    DeviceBuffer dops(std::max<std::size_t>(ops.size(), 1) * sizeof(WOp));
    DeviceBuffer dmats(std::max<std::size_t>(mats.size(), 1) * sizeof(V));
    DeviceBuffer dterms(std::max<std::size_t>(terms.size(), 1) * sizeof(SignTerm));
    upload(dterms.get(), terms.data(), terms.size() * sizeof(SignTerm));
    upload(dbundles.get(), bundles.data(), bundles.size() * sizeof(uint64_t));
    upload(dops.get(), ops.data(), ops.size() * sizeof(WOp));
    upload(dmats.get(), mats.data(), mats.size() * sizeof(V));

    std::vector<hipEvent_t> events;
    if (time_passes) {
        events.resize(passes.size() + 1);
        for (auto& e : events) MH_HIP(hipEventCreate(&e));
        MH_HIP(hipEventRecord(events[0], streams().main));
    }
    // With a monitor: an event after each pass; the monitor for pass p runs
    // once pass p + 1 is queued, so the device never waits for the host.
    const bool monitored = stats && stats->monitor;
    std::vector<hipEvent_t> done;
    struct EventsGuard {
        std::vector<hipEvent_t>& list;
        ~EventsGuard() {
            for (hipEvent_t e : list) (void)hipEventDestroy(e);
        }
    } done_guard{done}, time_guard{events};
    if (monitored) {
        done.resize(passes.size());
        for (auto& e : done) MH_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    auto boundary = [&](std::size_t p) {
        MH_HIP(hipEventSynchronize(done[p]));
        if (stats->monitor(p + 1)) {
            synchronize();
            fail(MH_ERR_ABANDONED, "the execution monitor abandoned the run after " + std::to_string(p + 1) +
                                       " of " + std::to_string(passes.size()) + " passes");
        }
    };
    for (std::size_t p = 0; p < passes.size(); ++p) {
        if (passes[p].generic) {
            const Block& blk = blocks[passes[p].blocks[0]];
            launch_generic(state, n, sizeof(R) == 8 ? MH_C128 : MH_C64, blk.matrix.data(), blk.operands.data(),
                           static_cast<uint32_t>(blk.operands.size()));
        } else {
            WPassArgs a = args[p];
            a.state = state;
            a.bundles = dbundles.as<uint64_t>() + bundle_begin[p];
            a.ops = dops.as<WOp>() + op_begin[p];
            a.terms = dterms.as<SignTerm>();
            a.mats = dmats.get();
            if (n < W || n - W > 30) fail(MH_ERR_INTERNAL, "window pass shape out of range");
            const uint32_t waves_log = n - W;
            const uint32_t group_log = std::min<uint32_t>(waves_log, 2);
            const dim3 grid(1u << (waves_log - group_log)), block(64u << group_log);
            if (uses_pairs(bundles.data() + bundle_begin[p], args[p].lay.num_bundles, Shape<R>::kRegBits))
                hipLaunchKernelGGL((window_kernel<R, true>), grid, block, 0, streams().main, a);
            else
                hipLaunchKernelGGL((window_kernel<R, false>), grid, block, 0, streams().main, a);
            MH_HIP(hipGetLastError());
        }
        if (time_passes) MH_HIP(hipEventRecord(events[p + 1], streams().main));
        if (monitored) {
            MH_HIP(hipEventRecord(done[p], streams().main));
            if (p > 0) boundary(p - 1);
        }
    }
    if (monitored && !passes.empty()) boundary(passes.size() - 1);
    synchronize();
    if (time_passes) {
        stats->pass_ms.resize(passes.size());
        for (std::size_t p = 0; p < passes.size(); ++p)
            MH_HIP(hipEventElapsedTime(&stats->pass_ms[p], events[p], events[p + 1]));
    }
}

// What a launched group of states needs until the stream has run it: its
// device tables and the host copies they were uploaded from.
struct BatchKeep {
    std::vector<DeviceBuffer> device;
    std::vector<std::shared_ptr<void>> host;
    hipEvent_t copied = nullptr;
    void release() {
        if (copied) (void)hipEventDestroy(copied);
        copied = nullptr;
    }
};


// What compile_batch_program keeps per state: where each pass's program
// starts, or which block a generic pass applies.
struct BatchPassRec {
    WLayout lay;
    uint32_t bundle_begin = 0, op_begin = 0;
    bool generic = false;
    uint32_t generic_block = 0;
};

}  // namespace

struct BatchProgramData {
    int32_t precision = MH_C128;
    std::vector<std::size_t> needs_init;             // states an initialisation kernel starts
    std::vector<unsigned char> bundles, ops, terms, mats, args, groups;
    std::vector<std::vector<BatchPassRec>> passes;   // per state
    std::vector<std::vector<std::size_t>> lgroups;   // launch groups of states
    std::vector<BatchLaunch> launches;
    bool has_generic = false;
    uint32_t max_passes = 0;
};

BatchProgram::BatchProgram() : d_(new BatchProgramData()) {}
BatchProgram::~BatchProgram() { delete d_; }
BatchProgram::BatchProgram(BatchProgram&& o) noexcept : d_(o.d_) { o.d_ = nullptr; }
BatchProgram& BatchProgram::operator=(BatchProgram&& o) noexcept {
    if (this != &o) {
        delete d_;
        d_ = o.d_;
        o.d_ = nullptr;
    }
    return *this;
}
int32_t BatchProgram::precision() const { return d_->precision; }
const std::vector<unsigned char>& BatchProgram::bundles() const { return d_->bundles; }
const std::vector<unsigned char>& BatchProgram::ops() const { return d_->ops; }
const std::vector<unsigned char>& BatchProgram::terms() const { return d_->terms; }
const std::vector<unsigned char>& BatchProgram::mats() const { return d_->mats; }
const std::vector<unsigned char>& BatchProgram::args() const { return d_->args; }
const std::vector<unsigned char>& BatchProgram::groups() const { return d_->groups; }
bool BatchProgram::has_generic() const { return d_->has_generic; }
uint32_t BatchProgram::pass_count(std::size_t item) const { return static_cast<uint32_t>(d_->passes[item].size()); }
uint32_t BatchProgram::launch_group_of(std::size_t item) const {
    for (uint32_t lg = 0; lg < d_->lgroups.size(); ++lg)
        for (std::size_t i : d_->lgroups[lg])
            if (i == item) return lg;
    return 0;
}
std::size_t BatchProgram::item_count() const { return d_->passes.size(); }
bool BatchProgram::needs_initialisation() const { return !d_->needs_init.empty(); }
uint32_t BatchProgram::launch_count() const { return static_cast<uint32_t>(d_->launches.size()); }
BatchProgramData& BatchProgram::data() { return *d_; }
const BatchProgramData& BatchProgram::data() const { return *d_; }

bool BatchProgram::same_shape(const BatchProgram& o) const {
    const BatchProgramData& a = *d_;
    const BatchProgramData& b = *o.d_;
    return a.precision == b.precision && a.bundles == b.bundles && a.ops == b.ops && a.terms == b.terms &&
           a.args == b.args && a.groups == b.groups && a.mats.size() == b.mats.size() &&
           a.has_generic == b.has_generic && a.launches.size() == b.launches.size();
}

std::size_t wop_bytes() { return sizeof(WOp); }
std::size_t sign_term_bytes() { return sizeof(SignTerm); }
std::size_t wbatch_args_bytes() { return sizeof(WBatchArgs); }

namespace {

template <typename T>
void append_bytes(std::vector<unsigned char>& dst, const std::vector<T>& src) {
    const std::size_t at = dst.size();
    dst.resize(at + src.size() * sizeof(T));
    if (!src.empty()) std::memcpy(dst.data() + at, src.data(), src.size() * sizeof(T));
}

template <typename R>
void compile_batch_typed(const BatchItem* items, std::size_t count, uint32_t threads, const BatchSchedule& sched,
                         BatchProgramData& prog) {
    using V = typename Shape<R>::V;
    constexpr uint32_t W = width_of<R>();
    struct Compiled {
        std::vector<uint64_t> bundles;
        std::vector<WOp> ops;
        std::vector<SignTerm> terms;
        std::vector<V> mats;
    };
    prog.precision = sizeof(R) == 8 ? MH_C128 : MH_C64;
    std::vector<Compiled> comp(count);
    prog.passes.assign(count, {});
    constexpr uint32_t kNoPhase = UINT32_MAX;
    std::vector<uint32_t> phase_at(count, kNoPhase);
    std::atomic<std::size_t> next{0};
    std::exception_ptr error;
    std::mutex error_mutex;
    auto worker = [&] {
        try {
            for (std::size_t i; (i = next.fetch_add(1)) < count;) {
                const BatchItem& it = items[i];
                Compiled& c = comp[i];
                const std::vector<WindowPass> passes =
                    plan_window_passes(it.n, prog.precision, *it.blocks, W, it.max_cost ? it.max_cost : kBatchPassCost);
                PassCompiler<R> compiler(c.bundles, c.ops, c.terms, c.mats);
                for (std::size_t p = 0; p < passes.size(); ++p) {
                    BatchPassRec rec;
                    rec.bundle_begin = static_cast<uint32_t>(c.bundles.size());
                    rec.op_begin = static_cast<uint32_t>(c.ops.size());
                    if (passes[p].generic) {
                        rec.generic = true;
                        rec.generic_block = passes[p].blocks[0];
                    } else {
                        rec.lay = compiler.compile(it.n, *it.blocks, passes[p], p + 1 == passes.size() && sizeof(R) == 8).lay;
                        rec.lay.num_bundles = static_cast<uint32_t>(c.bundles.size()) - rec.bundle_begin;
                        rec.lay.num_waves = 1u << (it.n - W);
                    }
                    prog.passes[i].push_back(rec);
                }
                // A state to initialise: its first pass generates it from the
                // phase, kept at the end of its matrix entries.
                if (it.initialise && !prog.passes[i].empty() && !prog.passes[i][0].generic) {
                    phase_at[i] = static_cast<uint32_t>(c.mats.size());
                    V ph;
                    ph.x = static_cast<R>(it.phase_re);
                    ph.y = static_cast<R>(it.phase_im);
                    c.mats.push_back(ph);
                }
            }
        } catch (...) {
            std::lock_guard<std::mutex> lock(error_mutex);
            if (!error) error = std::current_exception();
        }
    };
    {
        std::vector<std::thread> pool;
        const uint32_t t = std::max<uint32_t>(1, std::min<uint32_t>(threads, static_cast<uint32_t>(count)));
        for (uint32_t k = 1; k < t; ++k) pool.emplace_back(worker);
        worker();
        for (auto& th : pool) th.join();
    }
    if (error) std::rethrow_exception(error);
    prog.needs_init.clear();
    for (std::size_t i = 0; i < count; ++i)
        if (items[i].initialise && phase_at[i] == kNoPhase) prog.needs_init.push_back(i);

    // Every state's tables, packed one after another.
    std::vector<std::size_t> bundle_base(count), op_base(count), term_base(count), mat_base(count);
    std::size_t nb = 0, no = 0, nt = 0, nm = 0;
    prog.max_passes = 0;
    for (std::size_t i = 0; i < count; ++i) {
        bundle_base[i] = nb;
        op_base[i] = no;
        term_base[i] = nt;
        mat_base[i] = nm;
        nb += comp[i].bundles.size();
        no += comp[i].ops.size();
        nt += comp[i].terms.size();
        nm += comp[i].mats.size();
        prog.max_passes = std::max(prog.max_passes, static_cast<uint32_t>(prog.passes[i].size()));
    }
    prog.bundles.clear();
    prog.ops.clear();
    prog.terms.clear();
    prog.mats.clear();
    prog.bundles.reserve(std::max<std::size_t>(nb, 1) * sizeof(uint64_t));
    prog.ops.reserve(std::max<std::size_t>(no, 1) * sizeof(WOp));
    prog.terms.reserve(std::max<std::size_t>(nt, 1) * sizeof(SignTerm));
    prog.mats.reserve(std::max<std::size_t>(nm, 1) * sizeof(V));
    for (std::size_t i = 0; i < count; ++i) {
        append_bytes(prog.bundles, comp[i].bundles);
        append_bytes(prog.ops, comp[i].ops);
        append_bytes(prog.terms, comp[i].terms);
        append_bytes(prog.mats, comp[i].mats);
    }
    // Tables are never empty (a zero-length device buffer has no address).
    if (prog.bundles.empty()) prog.bundles.resize(sizeof(uint64_t));
    if (prog.ops.empty()) prog.ops.resize(sizeof(WOp));
    if (prog.terms.empty()) prog.terms.resize(sizeof(SignTerm));
    if (prog.mats.empty()) prog.mats.resize(sizeof(V));

    // Launch groups: every state in one group (pass p of all of them per
    // launch), or cache groups of at most sched.group_bytes of state, largest
    // states first, each running its passes back to back.
    prog.lgroups.clear();
    if (sched.group_bytes == 0) {
        prog.lgroups.emplace_back(count);
        for (std::size_t i = 0; i < count; ++i) prog.lgroups[0][i] = i;
    } else {
        std::vector<std::size_t> order(count);
        for (std::size_t i = 0; i < count; ++i) order[i] = i;
        std::stable_sort(order.begin(), order.end(), [&](std::size_t x, std::size_t y) { return items[x].n > items[y].n; });
        uint64_t held = 0;
        for (std::size_t i : order) {
            const uint64_t bytes = (uint64_t{1} << items[i].n) * sizeof(V);
            if (prog.lgroups.empty() || held + bytes > sched.group_bytes) {
                prog.lgroups.emplace_back();
                held = 0;
            }
            prog.lgroups.back().push_back(i);
            held += bytes;
        }
    }
    // Launch lists: pass p of every state of a group that has one. State
    // offsets count amplitudes from the lowest state address.
    uintptr_t base_addr = UINTPTR_MAX;
    for (std::size_t i = 0; i < count; ++i) base_addr = std::min(base_addr, reinterpret_cast<uintptr_t>(items[i].state));
    std::vector<WBatchArgs> all_args;
    std::vector<uint32_t> all_groups;
    prog.launches.clear();
    prog.has_generic = false;
    for (uint32_t lg = 0; lg < prog.lgroups.size(); ++lg) {
        std::size_t group_passes = 0;
        for (std::size_t i : prog.lgroups[lg]) group_passes = std::max(group_passes, prog.passes[i].size());
        for (std::size_t p = 0; p < group_passes; ++p) {
            BatchLaunch l;
            l.args_at = all_args.size();
            l.groups_at = all_groups.size();
            l.group = lg;
            l.pass = static_cast<uint32_t>(p);
            uint32_t g = 0;
            for (std::size_t i : prog.lgroups[lg]) {
                if (p >= prog.passes[i].size()) continue;
                const BatchPassRec& rec = prog.passes[i][p];
                if (rec.generic) {
                    prog.has_generic = true;
                    continue;
                }
                const uintptr_t addr = reinterpret_cast<uintptr_t>(items[i].state);
                if ((addr - base_addr) % sizeof(V) != 0)
                    fail(MH_ERR_INVALID, "a batched state is not aligned to its amplitudes");
                WBatchArgs a;
                std::memset(&a, 0, sizeof(a));
                a.state_off = (addr - base_addr) / sizeof(V);
                a.bundle_off = static_cast<uint32_t>(bundle_base[i] + rec.bundle_begin);
                a.op_off = static_cast<uint32_t>(op_base[i] + rec.op_begin);
                a.term_off = static_cast<uint32_t>(term_base[i]);
                a.mat_off = static_cast<uint32_t>(mat_base[i]);
                if (p == 0 && phase_at[i] != kNoPhase) {
                    a.generate = 1;
                    a.phase_at = phase_at[i];
                }
                a.lay = rec.lay;
                if (uses_pairs(comp[i].bundles.data() + rec.bundle_begin, rec.lay.num_bundles, Shape<R>::kRegBits))
                    l.pairs = true;
                all_args.push_back(a);
                all_groups.push_back(g);
                g += (a.lay.num_waves + kWavesPerGroup - 1) / kWavesPerGroup;
            }
            l.len = all_args.size() - l.args_at;
            l.total_groups = g;
            prog.launches.push_back(l);
        }
    }
    if (all_args.empty()) all_args.resize(1);
    if (all_groups.empty()) all_groups.resize(1);
    prog.args.clear();
    prog.groups.clear();
    append_bytes(prog.args, all_args);
    append_bytes(prog.groups, all_groups);
}

template <typename R>
uint32_t issue_batch_typed(const BatchProgramData& prog, const BatchItem* items, const BatchDeviceTables& dev,
                           const BatchSchedule& sched, uint32_t first_stream, const AfterLaunch& after) {
    using V = typename Shape<R>::V;
    uintptr_t base_addr = UINTPTR_MAX;
    for (std::size_t i = 0; i < prog.passes.size(); ++i)
        base_addr = std::min(base_addr, reinterpret_cast<uintptr_t>(items[i].state));
    V* const state_base = reinterpret_cast<V*>(base_addr);
    const V* const mats = static_cast<const V*>(dev.mats);
    const uint32_t nstreams =
        prog.has_generic ? 1u : std::max(1u, std::min(sched.kernel_streams, Streams::kKernelStreams));
    uint32_t launches = 0;
    // States that no window pass can generate start from an initialisation
    // kernel on their launch group's stream.
    if (!prog.needs_init.empty()) {
        std::vector<uint32_t> group_of(prog.passes.size(), 0);
        for (uint32_t lg = 0; lg < prog.lgroups.size(); ++lg)
            for (std::size_t i : prog.lgroups[lg]) group_of[i] = lg;
        for (std::size_t i : prog.needs_init) {
            const hipStream_t stream =
                prog.has_generic ? streams().main : streams().kernel((first_stream + group_of[i]) % nstreams);
            launch_init_on(items[i].state, items[i].n, prog.precision, items[i].phase_re, items[i].phase_im, stream);
            ++launches;
        }
    }
    for (const BatchLaunch& l : prog.launches) {
        const hipStream_t stream = prog.has_generic ? streams().main : streams().kernel((first_stream + l.group) % nstreams);
        if (l.len > 0) {
            const dim3 grid(l.total_groups), block(kWavesPerGroup * 64);
            const WBatchArgs* args_p = dev.args + l.args_at;
            const uint32_t* groups_p = dev.groups + l.groups_at;
            const uint32_t len = static_cast<uint32_t>(l.len);
            if (l.pairs && sched.temporal)
                hipLaunchKernelGGL((window_batch_kernel<R, true, true>), grid, block, 0, stream, args_p, groups_p, len,
                                   state_base, dev.bundles, dev.ops, dev.terms, mats);
            else if (l.pairs)
                hipLaunchKernelGGL((window_batch_kernel<R, true, false>), grid, block, 0, stream, args_p, groups_p, len,
                                   state_base, dev.bundles, dev.ops, dev.terms, mats);
            else if (sched.temporal)
                hipLaunchKernelGGL((window_batch_kernel<R, false, true>), grid, block, 0, stream, args_p, groups_p, len,
                                   state_base, dev.bundles, dev.ops, dev.terms, mats);
            else
                hipLaunchKernelGGL((window_batch_kernel<R, false, false>), grid, block, 0, stream, args_p, groups_p, len,
                                   state_base, dev.bundles, dev.ops, dev.terms, mats);
            MH_HIP(hipGetLastError());
            ++launches;
        }
        for (std::size_t i : prog.lgroups[l.group]) {
            if (l.pass >= prog.passes[i].size() || !prog.passes[i][l.pass].generic) continue;
            const Block& blk = (*items[i].blocks)[prog.passes[i][l.pass].generic_block];
            launch_generic(items[i].state, items[i].n, prog.precision, blk.matrix.data(), blk.operands.data(),
                           static_cast<uint32_t>(blk.operands.size()));
            ++launches;
        }
        if (after) after(l, stream);
    }
    return launches;
}

// A device copy of a table kept with the group that reads it.
void keep_table(BatchKeep& keep, const std::vector<unsigned char>& src, hipStream_t copy_stream, const void*& out) {
    DeviceBuffer d(src.size());
    auto held = std::make_shared<std::vector<unsigned char>>(src);
    MH_HIP(hipMemcpyAsync(d.get(), held->data(), held->size(), hipMemcpyHostToDevice, copy_stream));
    keep.host.push_back(held);
    out = d.get();
    keep.device.push_back(std::move(d));
}

// Many states, each with its own block list, compiled into one program whose
// tables go up on the copy stream; its launches then wait for them on the
// kernel streams (see issue_batch_program). Returns without waiting; `keep`
// holds what the launches read.
void launch_window_passes_batch(int32_t precision, const BatchItem* items, std::size_t count, uint32_t threads,
                                BatchStats* stats, BatchKeep& keep, const BatchSchedule& sched, uint32_t first_stream,
                                const BatchHooks* hooks) {
    context();
    const auto c0 = std::chrono::steady_clock::now();
    const BatchProgram prog = compile_batch_program(precision, items, count, threads, sched);
    BatchDeviceTables dev;
    const void* p = nullptr;
    keep_table(keep, prog.bundles(), streams().copy, p);
    dev.bundles = static_cast<const uint64_t*>(p);
    keep_table(keep, prog.ops(), streams().copy, p);
    dev.ops = static_cast<const WOp*>(p);
    keep_table(keep, prog.terms(), streams().copy, p);
    dev.terms = static_cast<const SignTerm*>(p);
    keep_table(keep, prog.mats(), streams().copy, dev.mats);
    keep_table(keep, prog.args(), streams().copy, p);
    dev.args = static_cast<const WBatchArgs*>(p);
    keep_table(keep, prog.groups(), streams().copy, p);
    dev.groups = static_cast<const uint32_t*>(p);
    MH_HIP(hipEventCreateWithFlags(&keep.copied, hipEventDisableTiming));
    MH_HIP(hipEventRecord(keep.copied, streams().copy));
    if (prog.has_generic()) {
        MH_HIP(hipStreamWaitEvent(streams().main, keep.copied, 0));
    } else {
        const uint32_t nstreams = std::max(1u, std::min(sched.kernel_streams, Streams::kKernelStreams));
        for (uint32_t k = 0; k < nstreams; ++k) MH_HIP(hipStreamWaitEvent(streams().kernel(k), keep.copied, 0));
    }
    if (stats) stats->compile_seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - c0).count();
    if (hooks && hooks->compiled) hooks->compiled(prog);
    const uint32_t launches =
        issue_batch_program(prog, items, dev, sched, first_stream, hooks ? hooks->after : AfterLaunch{});
    if (stats) {
        stats->launches += launches;
        stats->max_passes = std::max(stats->max_passes, prog.data().max_passes);
    }
}

}  // namespace

BatchProgram compile_batch_program(int32_t precision, const BatchItem* items, std::size_t count, uint32_t threads,
                                   const BatchSchedule& sched) {
    BatchProgram prog;
    if (precision == MH_C128)
        compile_batch_typed<double>(items, count, threads, sched, prog.data());
    else
        compile_batch_typed<float>(items, count, threads, sched, prog.data());
    return prog;
}

uint32_t issue_batch_program(const BatchProgram& prog, const BatchItem* items, const BatchDeviceTables& dev,
                             const BatchSchedule& sched, uint32_t first_stream, const AfterLaunch& after) {
    if (prog.precision() == MH_C128)
        return issue_batch_typed<double>(prog.data(), items, dev, sched, first_stream, after);
    return issue_batch_typed<float>(prog.data(), items, dev, sched, first_stream, after);
}

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
        const uint64_t dense = blk.diagonal ? 0 : register_mask(blk);
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

struct BatchPassRunner::Impl {
    int32_t precision = MH_C128;
    uint32_t threads = 1;
    BatchSchedule schedule;
    uint32_t added = 0;
    hipEvent_t start = nullptr;  // the work queued before the runner, for the second stream to wait on
    std::vector<BatchKeep> keeps;
};

BatchPassRunner::BatchPassRunner(int32_t precision, uint32_t threads, const BatchSchedule& schedule)
    : impl_(std::make_unique<Impl>()) {
    amp_bytes(precision);
    impl_->precision = precision;
    impl_->threads = threads;
    impl_->schedule = schedule;
    context();
    MH_HIP(hipEventCreateWithFlags(&impl_->start, hipEventDisableTiming));
    MH_HIP(hipEventRecord(impl_->start, streams().main));
    for (uint32_t k = 1; k < Streams::kKernelStreams; ++k) MH_HIP(hipStreamWaitEvent(streams().kernel(k), impl_->start, 0));
    MH_HIP(hipStreamWaitEvent(streams().reduce, impl_->start, 0));
}

BatchPassRunner::~BatchPassRunner() {
    // Whatever was launched or copied reads the kept tables: wait before they go.
    if (!impl_->keeps.empty()) {
        (void)hipStreamSynchronize(streams().copy);
        for (uint32_t k = 0; k < Streams::kKernelStreams; ++k) (void)hipStreamSynchronize(streams().kernel(k));
        (void)hipStreamSynchronize(streams().reduce);
    }
    for (BatchKeep& k : impl_->keeps) k.release();
    if (impl_->start) (void)hipEventDestroy(impl_->start);
}

void BatchPassRunner::add(const BatchItem* items, std::size_t count, BatchStats* stats, const BatchHooks* hooks) {
    if (count == 0) return;
    const uint32_t width = window_width(impl_->precision);
    for (std::size_t i = 0; i < count; ++i)
        if (items[i].n < width) fail(MH_ERR_INVALID, "a batched state has fewer qubits than one window");
    impl_->keeps.emplace_back();
    // Calls alternate between the kernel streams, so the short launches at
    // the end of one call overlap the next call's.
    context();
    const uint32_t first_stream = impl_->added++;
    launch_window_passes_batch(impl_->precision, items, count, impl_->threads, stats, impl_->keeps.back(),
                               impl_->schedule, first_stream, hooks);
}

void BatchPassRunner::finish() {
    MH_HIP(hipStreamSynchronize(streams().copy));
    for (uint32_t k = 1; k < Streams::kKernelStreams; ++k) MH_HIP(hipStreamSynchronize(streams().kernel(k)));
    MH_HIP(hipStreamSynchronize(streams().reduce));
    synchronize();
    for (BatchKeep& k : impl_->keeps) k.release();
    impl_->keeps.clear();
}

void run_window_passes_batch(int32_t precision, const std::vector<BatchItem>& items, uint32_t threads,
                             BatchStats* stats) {
    if (items.empty()) return;
    BatchPassRunner runner(precision, threads);
    runner.add(items.data(), items.size(), stats);
    runner.finish();
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
