// The gate-pass kernel: many plan blocks per HBM read and write of the state.
//
// A pass owns a window of W qubits: the low run 0..L-1 (so global accesses are
// contiguous runs) plus the higher qubits its dense blocks act on. Each
// workgroup takes one 2^W-amplitude window (the other qubits fixed by its
// workgroup index) and holds it in registers: 2^(W-4) threads x 16 amplitudes.
// The 16 amplitudes of a thread differ in 4 window bits, its "register slots";
// the thread index supplies the other W-4 bits. A dense block acts inside each
// thread when its operands are register slots, so the pass is split into
// phases, each with its own slot assignment, and the window is re-distributed
// between phases through LDS (a transpose: write 16 amplitudes, barrier, read
// 16). Diagonal blocks need no slots: every amplitude knows its own index.
//
// Loads and stores use the "home" assignment (slots = the top 4 window bits,
// thread bits = the low ones), so consecutive lanes touch consecutive
// amplitudes. LDS holds one window (64 KiB) and nothing else.
//
// Arithmetic (complex128) is the reference generic kernel's, bit for bit
// (reference src/kernels/cpu/dense_kernels.cpp:465-524): every output starts at
// (+0, +0) and adds M[out][in] * a[in] for in ascending, each product
// (mr*ar - mi*ai, mr*ai + mi*ar), every operation rounded on its own (this file
// is compiled with -ffp-contract=off). A diagonal block computes (+0) + d * a,
// which equals the generic sum (reference dense_kernels.cpp:299-313). A
// diagonal whose entries are all (+-1, +-0) computes (+-a) + 0 per component:
// with d = (s, +-0), s*ar - (+-0)*ai is s*ar whenever s*ar != 0 and a signed
// zero otherwise, and (+0) + x equals x + (+0) for every x, so the result is
// the same bits.
#include <algorithm>
#include <cstring>

#include "mh_internal.h"
#include "window.h"

namespace mh {
namespace {

constexpr uint32_t kRB = kWindowRegisterBits;
constexpr uint32_t kE = 1u << kRB;
constexpr uint32_t kMaxW = 13;
constexpr uint32_t kNone = 255;

enum OpKind : uint32_t { kDense1 = 0, kDenseK = 1, kDiag = 2, kDiagSign = 3 };

struct WOp {
    uint32_t kind;
    uint32_t k;
    uint32_t slot;          // kDense1: register slot of the operand
    uint32_t pad;
    uint64_t mat;           // entries: 4^k (dense) or 2^k (diagonal)
    uint64_t neg_mask;      // kDiagSign: bit i set when entry i is -1
    uint8_t dpos[8];        // diagonal operand j: window position, or kNone
    uint8_t dq[8];          // diagonal operand j: global qubit
};

struct WPhase {
    uint8_t slot[kRB];          // window position of register slot j
    uint8_t tpos[kMaxW];        // window position of thread bit j (w - 4 of them)
    uint32_t op_begin, op_end;
    uint32_t transpose;         // the assignment differs from the one before
    uint32_t pad;
};

struct WPassArgs {
    void* state;
    const WPhase* phases;
    const WOp* ops;
    const void* mats;
    uint64_t window_mask;
    uint32_t n, w, num_phases, final_transpose;
    uint8_t wq[kMaxW];          // global qubit of window position i
    uint8_t home_q[kRB];        // global qubits of the home register slots (wq[w-4 .. w-1])
};

template <typename R> struct Vec;
template <> struct Vec<double> { using type = double2; };
template <> struct Vec<float> { using type = float2; };

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

// Bank-spreading permutation of a window index: the low 3 bits are XORed with
// the XOR of the higher 3-bit groups (a bijection on [0, 2^w)). It is linear
// over XOR, and the thread's and the register's window bits are disjoint, so
// swz(thread | reg) = swz(thread) ^ swz(reg): the thread part is computed once
// per assignment and the register part is uniform (scalar) per element.
__device__ __forceinline__ uint32_t swz(uint32_t l) { return l ^ (((l >> 3) ^ (l >> 6) ^ (l >> 9) ^ (l >> 12)) & 7u); }

template <typename V>
__device__ __forceinline__ V* lds_at(unsigned char* base, uint32_t thread_off, uint32_t l_reg) {
    return reinterpret_cast<V*>(base + (thread_off ^ (swz(l_reg) * static_cast<uint32_t>(sizeof(V)))));
}

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

// Window position of each bit of register index r under slot assignment `slot`.
__device__ __forceinline__ uint32_t slot_bits(const uint8_t* slot, uint32_t r) {
    uint32_t l = 0;
#pragma unroll
    for (uint32_t j = 0; j < kRB; ++j) l |= ((r >> j) & 1u) << slot[j];
    return l;
}

__device__ __forceinline__ uint32_t thread_bits(const uint8_t* tpos, uint32_t tb, uint32_t t) {
    uint32_t l = 0;
#pragma unroll
    for (uint32_t j = 0; j < kMaxW - kRB; ++j)
        if (j < tb) l |= ((t >> j) & 1u) << tpos[j];
    return l;
}

template <typename R, uint32_t J>
__device__ __forceinline__ void apply1(typename Vec<R>::type* v, const typename Vec<R>::type* m) {
    using V = typename Vec<R>::type;
    const V m00 = m[0], m01 = m[1], m10 = m[2], m11 = m[3];
#pragma unroll
    for (uint32_t r = 0; r < kE; ++r) {
        if (r & (1u << J)) continue;
        const V a0 = v[r], a1 = v[r | (1u << J)];
        R s0r = R(0), s0i = R(0), s1r = R(0), s1i = R(0);
        cmac(s0r, s0i, m00.x, m00.y, a0.x, a0.y);
        cmac(s0r, s0i, m01.x, m01.y, a1.x, a1.y);
        cmac(s1r, s1i, m10.x, m10.y, a0.x, a0.y);
        cmac(s1r, s1i, m11.x, m11.y, a1.x, a1.y);
        v[r].x = s0r;
        v[r].y = s0i;
        v[r | (1u << J)].x = s1r;
        v[r | (1u << J)].y = s1i;
        // One pair at a time: interleaving all eight only raises register
        // pressure past two waves per SIMD.
        __builtin_amdgcn_sched_barrier(0);
    }
}

// A k-operand block whose operands are register slots 0..k-1 in operand order.
template <typename R, uint32_t K>
__device__ __forceinline__ void applyk(typename Vec<R>::type* v, const typename Vec<R>::type* m) {
    using V = typename Vec<R>::type;
    constexpr uint32_t dim = 1u << K;
#pragma unroll
    for (uint32_t hi = 0; hi < (kE >> K); ++hi) {
        V a[dim];
#pragma unroll
        for (uint32_t in = 0; in < dim; ++in) a[in] = v[in | (hi << K)];
#pragma unroll
        for (uint32_t o = 0; o < dim; ++o) {
            R sr = R(0), si = R(0);
#pragma unroll
            for (uint32_t in = 0; in < dim; ++in) {
                const V e = m[o * dim + in];
                cmac(sr, si, e.x, e.y, a[in].x, a[in].y);
            }
            v[o | (hi << K)].x = sr;
            v[o | (hi << K)].y = si;
        }
    }
}

// KMAX: the widest dense block the instance applies (2 or 4), so passes of 1-
// and 2-qubit blocks run with the registers those need.
template <typename R, uint32_t KMAX>
__global__ void __launch_bounds__(sizeof(R) == 8 ? 256 : 512, sizeof(R) == 8 ? 2 : 4) window_kernel(WPassArgs a) {
    using V = typename Vec<R>::type;
    extern __shared__ unsigned char lds_raw[];
    const uint32_t t = threadIdx.x;
    const uint32_t tb = a.w - kRB;
    const uint64_t obase = outer_base(blockIdx.x, a.window_mask, a.n);
    V* st = static_cast<V*>(a.state);
    const V* mats = static_cast<const V*>(a.mats);

    // Home assignment: thread bits are window positions 0..tb-1, slots tb..tb+3.
    V v[kE];
    {
        uint64_t tglob = 0;
#pragma unroll
        for (uint32_t j = 0; j < kMaxW - kRB; ++j)
            if (j < tb) tglob |= uint64_t((t >> j) & 1u) << a.wq[j];
        const V* src = st + obase + tglob;
#pragma unroll
        for (uint32_t r = 0; r < kE; ++r) {
            uint64_t h = 0;
#pragma unroll
            for (uint32_t j = 0; j < kRB; ++j) h |= uint64_t((r >> j) & 1u) << a.home_q[j];
            v[r] = src[h];
        }
    }

    uint8_t cur_slot[kRB];
#pragma unroll
    for (uint32_t j = 0; j < kRB; ++j) cur_slot[j] = static_cast<uint8_t>(tb + j);
    uint32_t cur_t = t;  // window bits the thread index supplies under the current assignment
    bool lds_used = false;

    for (uint32_t p = 0; p < a.num_phases; ++p) {
        const WPhase ph = a.phases[p];
        if (ph.transpose) {
            if (lds_used) __syncthreads();
            {
                const uint32_t toff = swz(cur_t) * static_cast<uint32_t>(sizeof(V));
#pragma unroll
                for (uint32_t r = 0; r < kE; ++r) *lds_at<V>(lds_raw, toff, slot_bits(cur_slot, r)) = v[r];
            }
            __syncthreads();
            cur_t = thread_bits(ph.tpos, tb, t);
#pragma unroll
            for (uint32_t j = 0; j < kRB; ++j) cur_slot[j] = ph.slot[j];
            {
                const uint32_t toff = swz(cur_t) * static_cast<uint32_t>(sizeof(V));
#pragma unroll
                for (uint32_t r = 0; r < kE; ++r) v[r] = *lds_at<V>(lds_raw, toff, slot_bits(cur_slot, r));
            }
            lds_used = true;
        }
        for (uint32_t i = ph.op_begin; i < ph.op_end; ++i) {
            const WOp op = a.ops[i];
            const V* m = mats + op.mat;
            if (op.kind == kDense1) {
                switch (op.slot) {
                    case 0: apply1<R, 0>(v, m); break;
                    case 1: apply1<R, 1>(v, m); break;
                    case 2: apply1<R, 2>(v, m); break;
                    default: apply1<R, 3>(v, m); break;
                }
            } else if (op.kind == kDenseK) {
                if (op.k == 2) {
                    applyk<R, 2>(v, m);
                } else if constexpr (KMAX >= 4) {
                    if (op.k == 3)
                        applyk<R, 3>(v, m);
                    else
                        applyk<R, 4>(v, m);
                }
            } else {
                // Index bits from the window positions the thread supplies and
                // from non-window qubits: once per op. Bits from register slots:
                // uniform per element.
                uint32_t tpart = 0;
#pragma unroll
                for (uint32_t j = 0; j < 8; ++j) {
                    if (j >= op.k) continue;
                    if (op.dpos[j] == kNone)
                        tpart |= uint32_t((obase >> op.dq[j]) & 1ull) << j;
                    else
                        tpart |= ((cur_t >> op.dpos[j]) & 1u) << j;
                }
#pragma unroll
                for (uint32_t r = 0; r < kE; ++r) {
                    const uint32_t lreg = slot_bits(cur_slot, r);
                    uint32_t rpart = 0;
#pragma unroll
                    for (uint32_t j = 0; j < 8; ++j)
                        if (j < op.k && op.dpos[j] != kNone) rpart |= ((lreg >> op.dpos[j]) & 1u) << j;
                    const uint32_t idx = tpart | rpart;
                    if (op.kind == kDiagSign) {
                        const bool neg = (op.neg_mask >> idx) & 1ull;
                        v[r].x = (neg ? -v[r].x : v[r].x) + R(0);
                        v[r].y = (neg ? -v[r].y : v[r].y) + R(0);
                    } else {
                        const V d = m[idx];
                        R sr = R(0), si = R(0);
                        cmac(sr, si, d.x, d.y, v[r].x, v[r].y);
                        v[r].x = sr;
                        v[r].y = si;
                    }
                }
            }
        }
    }

    if (a.final_transpose) {
        if (lds_used) __syncthreads();
        {
            const uint32_t toff = swz(cur_t) * static_cast<uint32_t>(sizeof(V));
#pragma unroll
            for (uint32_t r = 0; r < kE; ++r) *lds_at<V>(lds_raw, toff, slot_bits(cur_slot, r)) = v[r];
        }
        __syncthreads();
        const uint32_t toff = swz(t) * static_cast<uint32_t>(sizeof(V));
#pragma unroll
        for (uint32_t r = 0; r < kE; ++r) v[r] = *lds_at<V>(lds_raw, toff, r << tb);
    }
    // The store addresses are recomputed from an opaque copy of the thread
    // index rather than kept live (as 16 address pairs) across the whole pass.
    uint32_t t2 = t;
    asm volatile("" : "+v"(t2));
    uint64_t tglob = 0;
#pragma unroll
    for (uint32_t j = 0; j < kMaxW - kRB; ++j)
        if (j < tb) tglob |= uint64_t((t2 >> j) & 1u) << a.wq[j];
    V* dst = st + obase + tglob;
#pragma unroll
    for (uint32_t r = 0; r < kE; ++r) {
        uint64_t h = 0;
#pragma unroll
        for (uint32_t j = 0; j < kRB; ++j) h |= uint64_t((r >> j) & 1u) << a.home_q[j];
        dst[h] = v[r];
    }
}

// ---- host: passes and phases

// FP64 (or FP32) operations per amplitude a block costs (the pass cap unit).
uint32_t block_cost(const Block& b) {
    const uint32_t k = static_cast<uint32_t>(b.operands.size());
    if (b.diagonal) return 8;
    return 8u << k;
}

bool sign_diagonal(const Block& b) {
    const std::size_t dim = std::size_t{1} << b.operands.size();
    for (std::size_t r = 0; r < dim; ++r) {
        const cplx e = b.matrix[r * dim + r];
        if (!((e.real() == 1.0 || e.real() == -1.0) && e.imag() == 0.0)) return false;
    }
    return true;
}

uint64_t mask_of(const std::vector<uint32_t>& q) {
    uint64_t m = 0;
    for (uint32_t x : q) m |= 1ull << x;
    return m;
}

struct PhasePlan {
    std::vector<uint8_t> slots;  // up to kRB, then filled
    std::vector<WOp> ops;
    bool has_dense = false;
};

}  // namespace

std::vector<WindowPass> plan_window_passes(uint32_t n, const std::vector<Block>& blocks, uint32_t w,
                                           uint32_t max_cost) {
    const uint32_t low = std::min(std::min(n, w), 5u);
    const uint64_t low_mask = (1ull << low) - 1ull;
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
        const uint32_t c = block_cost(blk);
        if (!blk.diagonal && (k > kRB || n < kRB)) {
            // Wider than the register slots (or a state smaller than one
            // thread's amplitudes): a pass of its own on the generic kernel.
            close();
            WindowPass g;
            g.generic = true;
            g.blocks.push_back(b);
            passes.push_back(g);
            continue;
        }
        if (blk.diagonal && n < kRB) {
            close();
            WindowPass g;
            g.generic = true;
            g.blocks.push_back(b);
            passes.push_back(g);
            continue;
        }
        const uint64_t dense = blk.diagonal ? 0 : mask_of(blk.operands);
        const uint64_t want = cur.dense_mask | dense;
        const bool fits = __builtin_popcountll(want | low_mask) <= static_cast<int>(w);
        if (open && fits && (cost + c <= max_cost || cur.blocks.empty())) {
            cur.dense_mask = want;
            if (!blk.diagonal) cur.kmax = std::max(cur.kmax, k);
            cur.blocks.push_back(b);
            cost += c;
            continue;
        }
        close();
        if (__builtin_popcountll(dense | low_mask) > static_cast<int>(w))
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

namespace {

template <typename R>
void launch_window_passes(void* state, uint32_t n, const std::vector<Block>& blocks,
                          const std::vector<WindowPass>& passes, uint32_t w, bool time_passes, PassStats* stats) {
    using V = typename Vec<R>::type;
    Context& ctx = context();
    std::vector<WPassArgs> args;
    std::vector<WPhase> phases;
    std::vector<WOp> ops;
    std::vector<V> mats;
    std::vector<uint32_t> phase_begin;
    std::vector<int> generic_block(passes.size(), -1);
    const uint32_t low = std::min(std::min(n, w), 5u);
    for (std::size_t p = 0; p < passes.size(); ++p) {
        const WindowPass& pass = passes[p];
        WPassArgs a;
        std::memset(&a, 0, sizeof(a));
        phase_begin.push_back(static_cast<uint32_t>(phases.size()));
        if (pass.generic) {
            generic_block[p] = static_cast<int>(pass.blocks[0]);
            args.push_back(a);
            continue;
        }
        uint64_t wmask = pass.dense_mask | ((1ull << low) - 1ull);
        for (uint32_t q = 0; q < n && static_cast<uint32_t>(__builtin_popcountll(wmask)) < w; ++q) wmask |= 1ull << q;
        uint8_t pos_of[64];
        std::memset(pos_of, kNone, sizeof(pos_of));
        uint32_t wcount = 0;
        for (uint32_t q = 0; q < n; ++q)
            if ((wmask >> q) & 1ull) {
                a.wq[wcount] = static_cast<uint8_t>(q);
                pos_of[q] = static_cast<uint8_t>(wcount++);
            }
        a.n = n;
        a.w = wcount;
        a.window_mask = wmask;
        const uint32_t tb = wcount - kRB;
        for (uint32_t j = 0; j < kRB; ++j) a.home_q[j] = a.wq[tb + j];

        std::vector<PhasePlan> plan(1);
        for (uint32_t b : pass.blocks) {
            const Block& blk = blocks[b];
            const uint32_t k = static_cast<uint32_t>(blk.operands.size());
            const std::size_t dim = std::size_t{1} << k;
            WOp op;
            std::memset(&op, 0, sizeof(op));
            op.k = k;
            op.mat = mats.size();
            if (blk.diagonal) {
                op.kind = sign_diagonal(blk) ? kDiagSign : kDiag;
                for (uint32_t j = 0; j < k; ++j) {
                    op.dpos[j] = pos_of[blk.operands[j]];
                    op.dq[j] = static_cast<uint8_t>(blk.operands[j]);
                }
                for (std::size_t r = 0; r < dim; ++r) {
                    const cplx e = blk.matrix[r * dim + r];
                    if (e.real() == -1.0) op.neg_mask |= 1ull << r;
                    V x;
                    x.x = static_cast<R>(e.real());
                    x.y = static_cast<R>(e.imag());
                    mats.push_back(x);
                }
                plan.back().ops.push_back(op);
                continue;
            }
            for (std::size_t e = 0; e < dim * dim; ++e) {
                V x;
                x.x = static_cast<R>(blk.matrix[e].real());
                x.y = static_cast<R>(blk.matrix[e].imag());
                mats.push_back(x);
            }
            std::vector<uint8_t> want;
            for (uint32_t q : blk.operands) want.push_back(pos_of[q]);
            PhasePlan* ph = &plan.back();
            if (k == 1) {
                op.kind = kDense1;
                auto it = std::find(ph->slots.begin(), ph->slots.end(), want[0]);
                if (it == ph->slots.end() && ph->slots.size() >= kRB) {
                    plan.emplace_back();
                    ph = &plan.back();
                    it = ph->slots.end();
                }
                if (it == ph->slots.end()) {
                    ph->slots.push_back(want[0]);
                    it = ph->slots.end() - 1;
                }
                op.slot = static_cast<uint32_t>(it - ph->slots.begin());
            } else {
                op.kind = kDenseK;
                const bool prefix = ph->slots.size() >= k && std::equal(want.begin(), want.end(), ph->slots.begin());
                if (!prefix) {
                    if (ph->has_dense) {
                        plan.emplace_back();
                        ph = &plan.back();
                    }
                    ph->slots = want;
                }
            }
            ph->has_dense = true;
            ph->ops.push_back(op);
        }
        // Fill every phase's slots to kRB; thread bits are the rest, ascending.
        std::vector<uint8_t> prev_slot(kRB), prev_t(tb);
        for (uint32_t j = 0; j < kRB; ++j) prev_slot[j] = static_cast<uint8_t>(tb + j);
        for (uint32_t j = 0; j < tb; ++j) prev_t[j] = static_cast<uint8_t>(j);
        for (PhasePlan& pp : plan) {
            for (int cand = static_cast<int>(wcount) - 1; pp.slots.size() < kRB && cand >= 0; --cand)
                if (std::find(pp.slots.begin(), pp.slots.end(), cand) == pp.slots.end())
                    pp.slots.push_back(static_cast<uint8_t>(cand));
            WPhase wp;
            std::memset(&wp, 0, sizeof(wp));
            std::vector<uint8_t> tpos;
            for (uint32_t x = 0; x < wcount; ++x)
                if (std::find(pp.slots.begin(), pp.slots.end(), x) == pp.slots.end()) tpos.push_back(static_cast<uint8_t>(x));
            for (uint32_t j = 0; j < kRB; ++j) wp.slot[j] = pp.slots[j];
            for (uint32_t j = 0; j < tb; ++j) wp.tpos[j] = tpos[j];
            wp.transpose = (pp.slots != prev_slot || tpos != prev_t) ? 1u : 0u;
            wp.op_begin = static_cast<uint32_t>(ops.size());
            ops.insert(ops.end(), pp.ops.begin(), pp.ops.end());
            wp.op_end = static_cast<uint32_t>(ops.size());
            phases.push_back(wp);
            prev_slot = pp.slots;
            prev_t = tpos;
        }
        a.num_phases = static_cast<uint32_t>(plan.size());
        bool home = true;
        for (uint32_t j = 0; j < kRB; ++j) home = home && prev_slot[j] == tb + j;
        for (uint32_t j = 0; j < tb; ++j) home = home && prev_t[j] == j;
        a.final_transpose = home ? 0u : 1u;
        args.push_back(a);
    }
    DeviceBuffer dph(std::max<std::size_t>(phases.size(), 1) * sizeof(WPhase));
    DeviceBuffer dops(std::max<std::size_t>(ops.size(), 1) * sizeof(WOp));
    DeviceBuffer dmats(std::max<std::size_t>(mats.size(), 1) * sizeof(V));
    upload(dph.get(), phases.data(), phases.size() * sizeof(WPhase));
    upload(dops.get(), ops.data(), ops.size() * sizeof(WOp));
    upload(dmats.get(), mats.data(), mats.size() * sizeof(V));

    std::vector<hipEvent_t> events;
    if (time_passes) {
        events.resize(passes.size() + 1);
        for (auto& e : events) MH_HIP(hipEventCreate(&e));
        MH_HIP(hipEventRecord(events[0], ctx.stream));
    }
    for (std::size_t p = 0; p < passes.size(); ++p) {
        if (generic_block[p] >= 0) {
            const Block& blk = blocks[static_cast<std::size_t>(generic_block[p])];
            launch_generic(state, n, sizeof(V) == 16 ? MH_C128 : MH_C64, blk.matrix.data(), blk.operands.data(),
                           static_cast<uint32_t>(blk.operands.size()));
        } else {
            WPassArgs a = args[p];
            a.state = state;
            a.phases = dph.as<WPhase>() + phase_begin[p];
            a.ops = dops.as<WOp>();
            a.mats = dmats.get();
            const uint32_t threads = 1u << (a.w - kRB);
            const std::size_t lds = sizeof(V) << a.w;
            if (a.w > kMaxW || threads > (sizeof(V) == 16 ? 256u : 512u) || lds > 65536)
                fail(MH_ERR_INTERNAL, "window pass shape out of range");
            if ((n - a.w) > 31) fail(MH_ERR_INTERNAL, "too many windows for one launch");
            if (passes[p].kmax <= 2)
                hipLaunchKernelGGL((window_kernel<R, 2>), dim3(1u << (n - a.w)), dim3(threads), lds, ctx.stream, a);
            else
                hipLaunchKernelGGL((window_kernel<R, 4>), dim3(1u << (n - a.w)), dim3(threads), lds, ctx.stream, a);
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

void run_window_passes(void* state, uint32_t n, int32_t precision, const std::vector<Block>& blocks, uint32_t w,
                       uint32_t max_cost, bool time_passes, PassStats* stats) {
    const std::vector<WindowPass> passes = plan_window_passes(n, blocks, w, max_cost);
    if (stats) stats->passes = static_cast<uint32_t>(passes.size());
    if (passes.empty()) return;
    if (precision == MH_C128)
        launch_window_passes<double>(state, n, blocks, passes, w, time_passes, stats);
    else
        launch_window_passes<float>(state, n, blocks, passes, w, time_passes, stats);
}

}  // namespace mh
