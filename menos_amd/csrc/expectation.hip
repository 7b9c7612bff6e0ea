// Pauli-string expectation reductions (FP64 for both precisions).
//
// The raw term of P = (x, z) is (reference src/kernels/dense_kernels.h:145-194,
// dense_kernels.cpp:544-569)
//
//     i^popcount(x & z) * sum_m conj(psi[m]) * psi[m ^ x] * (-1)^popcount(z & (m ^ x)).
//
// Terms sharing an X mask share their products. X masks whose bits fit an LDS
// tile (with the lowest qubits) are reduced from tiles: one read of the state
// serves a whole set of X masks (expectation_tile_kernel). A wider X mask reads
// the state once for up to kTermsPerLaunch terms (expectation_group). For x != 0
// the sum is taken over
// pairs (m, m' = m ^ x) with bit h = msb(x) of m clear: with p = conj(psi[m]) *
// psi[m'], the pair contributes p * s(m') + conj(p) * s(m), s(k) =
// (-1)^popcount(z & k); conj(p) is bitwise conj(psi[m']) * psi[m], so each
// product is the reference's, read once. Each thread sums a fixed set of items
// in order, then the workgroup and the partials are added in a fixed order:
// the value is reproducible run to run, the same for one state alone and in a
// batch (both take this path), and differs from the
// reference's sequential index-order sum by rounding only (the reference's
// own CUDA contract, dense_kernels.h:237-246).
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <map>
#include <mutex>

#include "mh_internal.h"

namespace mh {
namespace {

constexpr int kThreads = 256;
constexpr uint32_t kTermsPerLaunch = 16;
constexpr uint32_t kMaxBlocks = 2048;

template <typename V>
__device__ __forceinline__ void load_amp(const V* s, uint64_t i, double& re, double& im) {
    const V v = s[i];
    re = static_cast<double>(v.x);
    im = static_cast<double>(v.y);
}

// Workgroup `block` of `blocks` for one group of terms sharing an X mask:
// writes partials[(t * blocks + block) * 2 + {0, 1}] for t < terms.
template <typename V, uint32_t T>
__device__ __forceinline__ void expectation_group(const V* s, double* partials, uint64_t pairs, uint64_t x, uint32_t h,
                                                  uint32_t terms, const uint64_t* z, uint32_t block, uint32_t blocks) {
    __shared__ double red[2][kThreads];
    // A fixed grid-stride assignment (coalesced; the same for every run).
    const uint64_t total_threads = uint64_t(blocks) * kThreads;
    const uint64_t tid = uint64_t(block) * kThreads + threadIdx.x;
    double accr[T], acci[T];
#pragma unroll
    for (uint32_t t = 0; t < T; ++t) accr[t] = acci[t] = 0.0;

    // Items are taken kUnroll at a time: every load of a batch is issued
    // before any is used (the loop is otherwise bound by memory latency), and
    // the sums still run over the items in order.
    constexpr uint32_t kUnroll = 4;
    if (x == 0) {
        for (uint64_t m0 = tid; m0 < pairs; m0 += kUnroll * total_threads) {
            double ar[kUnroll], ai[kUnroll];
#pragma unroll
            for (uint32_t k = 0; k < kUnroll; ++k) {
                const uint64_t m = m0 + k * total_threads;
                if (m < pairs) load_amp(s, m, ar[k], ai[k]);
            }
#pragma unroll
            for (uint32_t k = 0; k < kUnroll; ++k) {
                const uint64_t m = m0 + k * total_threads;
                if (m >= pairs) break;
                // conj(a) * a: real ar*ar + ai*ai, imaginary ar*ai - ai*ar (exactly 0).
                const double pr = ar[k] * ar[k] + ai[k] * ai[k];
                const double pi = ar[k] * ai[k] - ai[k] * ar[k];
#pragma unroll
                for (uint32_t t = 0; t < T; ++t) {
                    if (t >= terms) break;
                    const bool neg = (__popcll(z[t] & m) & 1) != 0;
                    accr[t] = neg ? accr[t] - pr : accr[t] + pr;
                    acci[t] = neg ? acci[t] - pi : acci[t] + pi;
                }
            }
        }
    } else {
        const uint64_t lowm = (1ull << h) - 1ull;
        for (uint64_t r0 = tid; r0 < pairs; r0 += kUnroll * total_threads) {
            double ar[kUnroll], ai[kUnroll], br[kUnroll], bi[kUnroll];
#pragma unroll
            for (uint32_t k = 0; k < kUnroll; ++k) {
                const uint64_t r = r0 + k * total_threads;
                if (r >= pairs) continue;
                const uint64_t m = ((r & ~lowm) << 1) | (r & lowm);
                load_amp(s, m, ar[k], ai[k]);
                load_amp(s, m ^ x, br[k], bi[k]);
            }
#pragma unroll
            for (uint32_t k = 0; k < kUnroll; ++k) {
                const uint64_t r = r0 + k * total_threads;
                if (r >= pairs) break;
                const uint64_t m = ((r & ~lowm) << 1) | (r & lowm);
                const uint64_t mp = m ^ x;
                // p = conj(a) * b
                const double pr = ar[k] * br[k] + ai[k] * bi[k];
                const double pi = ar[k] * bi[k] - ai[k] * br[k];
#pragma unroll
                for (uint32_t t = 0; t < T; ++t) {
                    if (t >= terms) break;
                    const bool neg_mp = (__popcll(z[t] & mp) & 1) != 0;  // s(m') for p
                    const bool neg_m = (__popcll(z[t] & m) & 1) != 0;    // s(m) for conj(p)
                    accr[t] = neg_mp ? accr[t] - pr : accr[t] + pr;
                    acci[t] = neg_mp ? acci[t] - pi : acci[t] + pi;
                    accr[t] = neg_m ? accr[t] - pr : accr[t] + pr;
                    acci[t] = neg_m ? acci[t] + pi : acci[t] - pi;
                }
            }
        }
    }
    // Fixed-order tree over the workgroup, one term at a time.
    for (uint32_t t = 0; t < terms && t < T; ++t) {
        double vr = 0.0, vi = 0.0;
#pragma unroll
        for (uint32_t u = 0; u < T; ++u)
            if (u == t) { vr = accr[u]; vi = acci[u]; }
        red[0][threadIdx.x] = vr;
        red[1][threadIdx.x] = vi;
        __syncthreads();
        for (int stride = kThreads / 2; stride > 0; stride >>= 1) {
            if (threadIdx.x < stride) {
                red[0][threadIdx.x] = red[0][threadIdx.x] + red[0][threadIdx.x + stride];
                red[1][threadIdx.x] = red[1][threadIdx.x] + red[1][threadIdx.x + stride];
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            partials[(uint64_t(t) * blocks + block) * 2] = red[0][0];
            partials[(uint64_t(t) * blocks + block) * 2 + 1] = red[1][0];
        }
        __syncthreads();
    }
}

// One group of terms of one state in a batched launch; offsets count from the
// launch's base pointers (amplitudes, doubles).
struct ExpJob {
    uint64_t state_off;
    uint64_t pairs;
    uint64_t x;
    uint32_t h, terms, blocks, pad;
    uint64_t partial_off;
    uint64_t z[kTermsPerLaunch];
};

// Workgroup b belongs to job j with block_begin[j] <= b < block_begin[j + 1].
template <typename V>
__global__ void __launch_bounds__(kThreads)
    expectation_batch_kernel(const ExpJob* __restrict__ jobs, const uint32_t* __restrict__ block_begin, uint32_t count,
                             const V* base, double* partials) {
    const uint32_t b = blockIdx.x;
    uint32_t lo = 0, hi = count;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (block_begin[mid] <= b) {
            lo = mid;
        } else {
            hi = mid;
        }
    }
    const ExpJob& j = jobs[lo];
    expectation_group<V, kTermsPerLaunch>(base + j.state_off, partials + j.partial_off, j.pairs, j.x, j.h, j.terms,
                                          j.z, b - block_begin[lo], j.blocks);
}

// Tiles: a workgroup loads 2^k amplitudes of one state into LDS at a time
// (the tile bits T: every X bit of a set of up to kTermsPerLaunch terms, the
// lowest qubits for whole 128-byte runs, then the next lowest), and reduces
// every term of the set from there, so the state is read from HBM once per set
// rather than once per X mask. Every (m, m ^ x) pair of a term lies in one
// tile because x's bits are in T. A workgroup takes tiles b, b + blocks, ...
// in order and keeps one sum per term across them.
constexpr uint32_t kTileBits = 11;

struct TileJob {
    uint64_t state_off;      // amplitudes from the launch base
    uint64_t outer_mask;     // index bits outside T: the tile index deposits here, ascending
    uint32_t k;              // tile bits (2^k amplitudes per tile)
    uint32_t term_begin, term_count;
    uint32_t tiles, blocks;
    uint32_t pad;
    uint8_t tpos[16];        // tile bit i -> qubit, ascending
};

// What a tile term sums: one real number per term (tile_term).
enum TileKind : uint32_t { kDiagonalTerm = 0, kRealTerm = 1, kImagTerm = 2 };

struct TileTerm {
    uint64_t x, z;
    uint32_t xl;             // x on the tile bits
    uint32_t hl;             // msb of xl (x != 0)
    uint32_t zl;             // z on the tile bits
    uint32_t kind;           // TileKind
    uint64_t partial_off;    // partials[(block) * 2 + {0, 1}] from here
};

// d with its sign flipped when `neg`: a + flip(p, neg) is bitwise a - p or a + p.
__device__ __forceinline__ double flip(double d, bool neg) {
    return __longlong_as_double(__double_as_longlong(d) ^ (static_cast<long long>(neg) << 63));
}

// One term over one tile in LDS, for thread u. Thread u handles tile entries
// u, u + kThreads, ... (pairs r = u, u + kThreads, ... for an X term). Entry
// and pair indices split into the thread's bits and a step that is the same
// across the workgroup, and both the index map and the sign parity are linear
// over that split: the per-step parts are uniform, so each entry costs two
// index XORs and one sign XOR. s(m ^ x) = s(m) ^ popcount(z & x). kFull: the
// tile has 2^kTileBits entries, so no entry is out of range.
//
// Each term needs one real sum. x = 0 (kDiagonalTerm): s(m) |a_m|^2. x != 0:
// the pair's s(m') p + s(m) conj(p), with s(m') = s(m) (-1)^sx and sx =
// popcount(z & x) mod 2, is 2 s(m) Re p for sx = 0 (kRealTerm) and
// 2i s(m') Im p for sx = 1 (kImagTerm); the factor 2 is applied to the
// workgroup's sum (exact), and the other part is zero.
template <bool kFull, typename V>
__device__ __forceinline__ void tile_term(const V* tile, const TileTerm& tt, uint32_t pz, uint32_t u, uint32_t size,
                                          double& acc) {
    if (tt.kind == kDiagonalTerm) {
        const uint32_t su = (static_cast<uint32_t>(__popc(tt.zl & u)) ^ pz) & 1u;
#pragma unroll
        for (uint32_t step = 0; step < (1u << kTileBits); step += kThreads) {
            const uint32_t j = step | u;
            if (!kFull && j >= size) continue;
            const double ar = static_cast<double>(tile[j].x), ai = static_cast<double>(tile[j].y);
            const double pr = ar * ar + ai * ai;
            const bool neg = ((su ^ static_cast<uint32_t>(__popc(tt.zl & step))) & 1u) != 0;
            acc = acc + flip(pr, neg);
        }
        return;
    }
    // Pair r -> m: r's bits with a zero inserted at bit hl (the top bit of xl).
    const uint32_t lowm = (1u << tt.hl) - 1u;
    const uint32_t ju = ((u & ~lowm) << 1) | (u & lowm);
    const uint32_t jpu = ju ^ tt.xl;
    // The sign taken: s(m) for a real term, s(m') = s(m) ^ 1 for an imaginary one.
    const uint32_t su = (static_cast<uint32_t>(__popc(tt.zl & ju)) ^ pz ^ (tt.kind == kImagTerm ? 1u : 0u)) & 1u;
    if (tt.kind == kRealTerm) {
#pragma unroll
        for (uint32_t step = 0; step < (1u << kTileBits) / 2; step += kThreads) {
            if (!kFull && (step | u) >= size / 2) continue;
            const uint32_t js = ((step & ~lowm) << 1) | (step & lowm);
            const uint32_t j = ju | js, jp = jpu ^ js;
            const double ar = static_cast<double>(tile[j].x), ai = static_cast<double>(tile[j].y);
            const double br = static_cast<double>(tile[jp].x), bi = static_cast<double>(tile[jp].y);
            const double pr = ar * br + ai * bi;  // Re conj(a) b
            const bool neg = ((su ^ static_cast<uint32_t>(__popc(tt.zl & js))) & 1u) != 0;
            acc = acc + flip(pr, neg);
        }
    } else {
#pragma unroll
        for (uint32_t step = 0; step < (1u << kTileBits) / 2; step += kThreads) {
            if (!kFull && (step | u) >= size / 2) continue;
            const uint32_t js = ((step & ~lowm) << 1) | (step & lowm);
            const uint32_t j = ju | js, jp = jpu ^ js;
            const double ar = static_cast<double>(tile[j].x), ai = static_cast<double>(tile[j].y);
            const double br = static_cast<double>(tile[jp].x), bi = static_cast<double>(tile[jp].y);
            const double pi = ar * bi - ai * br;  // Im conj(a) b
            const bool neg = ((su ^ static_cast<uint32_t>(__popc(tt.zl & js))) & 1u) != 0;
            acc = acc + flip(pi, neg);
        }
    }
}

// Five waves per SIMD: the kernel waits on LDS and memory latency, and the
// sixteen per-term sums in double fit beside the tile indices without spills.
// complex128 reaches four (its 32 KB tile caps the CU at four workgroups) but
// keeps the same register budget; the compiler's note that five is out of
// reach there is silenced.
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Wpass-failed"
template <typename V>
__global__ void __launch_bounds__(kThreads, 5)
    expectation_tile_kernel(const TileJob* __restrict__ jobs, const uint32_t* __restrict__ block_begin,
                            uint32_t count, const TileTerm* __restrict__ terms, const V* base, double* partials) {
    __shared__ V tile[1u << kTileBits];
    __shared__ uint64_t dep_lo[64], dep_hi[32];
    __shared__ double red[kThreads / 64][kTermsPerLaunch];
    const uint32_t b = blockIdx.x;
    uint32_t lo = 0, hi = count;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (block_begin[mid] <= b) {
            lo = mid;
        } else {
            hi = mid;
        }
    }
    const TileJob& job = jobs[lo];
    const uint32_t blk = b - block_begin[lo];
    const uint32_t k = job.k, size = 1u << k;
    // Tile-local index -> global index bits, as two lookup tables (bits 0-5, 6-10).
    if (threadIdx.x < 96) {
        const uint32_t part = threadIdx.x < 64 ? threadIdx.x : (threadIdx.x - 64) << 6;
        uint64_t g = 0;
        for (uint32_t i = 0; i < k; ++i)
            if ((part >> i) & 1u) g |= 1ull << job.tpos[i];
        if (threadIdx.x < 64)
            dep_lo[threadIdx.x] = g;
        else
            dep_hi[threadIdx.x - 64] = g;
    }
    const V* s = base + job.state_off;
    const uint32_t nt = job.term_count;
    double acc[kTermsPerLaunch];
#pragma unroll
    for (uint32_t t = 0; t < kTermsPerLaunch; ++t) acc[t] = 0.0;

    for (uint32_t tl = blk; tl < job.tiles; tl += job.blocks) {
        uint64_t tbase = 0;
        {
            uint64_t om = job.outer_mask;
            for (uint32_t t = tl; om != 0; t >>= 1) {
                const uint64_t low = om & (~om + 1);
                if (t & 1u) tbase |= low;
                om ^= low;
            }
        }
        __syncthreads();  // the tables are written; the previous tile is no longer read
        for (uint32_t j = threadIdx.x; j < size; j += kThreads) tile[j] = s[tbase | dep_lo[j & 63u] | dep_hi[j >> 6]];
        __syncthreads();
        // s(m) = (-1)^popcount(z & m) splits into the tile's outside bits (one
        // parity per tile) and the tile bits (z and x taken onto them).
        // Whole tiles take the form without range checks, whose LDS reads
        // the compiler can issue together; both forms add in the same order.
        const uint32_t u = threadIdx.x;
#pragma unroll
        for (uint32_t t = 0; t < kTermsPerLaunch; ++t) {
            if (t >= nt) continue;
            const TileTerm& tt = terms[job.term_begin + t];
            const uint32_t pz = static_cast<uint32_t>(__popcll(tt.z & tbase)) & 1u;
            if (size == (1u << kTileBits))
                tile_term<true>(tile, tt, pz, u, size, acc[t]);
            else
                tile_term<false>(tile, tt, pz, u, size, acc[t]);
        }
    }
    // Fixed order: a butterfly within each wave, then the waves in order.
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
#pragma unroll
    for (uint32_t t = 0; t < kTermsPerLaunch; ++t) {
        if (t >= nt) continue;
        double v = acc[t];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v = v + __shfl_xor(v, off, 64);
        if (lane == 0) red[wave][t] = v;
    }
    __syncthreads();
    if (threadIdx.x < nt) {
        const uint32_t t = threadIdx.x;
        double v = red[0][t];
#pragma unroll
        for (uint32_t w = 1; w < kThreads / 64; ++w) v = v + red[w][t];
        const TileTerm& tt = terms[job.term_begin + t];
        const double sum = tt.kind == kDiagonalTerm ? v : 2.0 * v;
        double* pp = partials + tt.partial_off + uint64_t(blk) * 2;
        pp[0] = tt.kind == kImagTerm ? 0.0 : sum;
        pp[1] = tt.kind == kImagTerm ? sum : 0.0;
    }
}

#pragma clang diagnostic pop

// One term slot of a batched launch.
struct ExpSlot {
    uint64_t partial_off;    // the term's first partial (doubles)
    uint32_t blocks;
    int32_t phase_k;         // popcount(x & z)
    uint32_t out;            // output term index
    uint32_t pad;
};

// The block partials of one term in a fixed tree order, then the phase
// i^popcount(x & z) as a complex product (reference dense_kernels.cpp:568).
__device__ __forceinline__ void finish_term(const double* partials, uint32_t blocks, int32_t phase_k, double* out) {
    __shared__ double red[2][kThreads];
    double vr = 0.0, vi = 0.0;
    for (uint32_t b = threadIdx.x; b < blocks; b += kThreads) {
        vr = vr + partials[uint64_t(b) * 2];
        vi = vi + partials[uint64_t(b) * 2 + 1];
    }
    red[0][threadIdx.x] = vr;
    red[1][threadIdx.x] = vi;
    __syncthreads();
    for (int stride = kThreads / 2; stride > 0; stride >>= 1) {
        if (threadIdx.x < stride) {
            red[0][threadIdx.x] = red[0][threadIdx.x] + red[0][threadIdx.x + stride];
            red[1][threadIdx.x] = red[1][threadIdx.x] + red[1][threadIdx.x + stride];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const int k = phase_k & 3;
        const double fr = k == 0 ? 1.0 : (k == 2 ? -1.0 : 0.0);
        const double fi = k == 1 ? 1.0 : (k == 3 ? -1.0 : 0.0);
        const double sr = red[0][0], si = red[1][0];
        out[0] = fr * sr - fi * si;
        out[1] = fr * si + fi * sr;
    }
}

// One workgroup per term.
__global__ void __launch_bounds__(kThreads) expectation_finish_batch_kernel(const ExpSlot* __restrict__ slots,
                                                                            const double* partials, double* out) {
    const ExpSlot& s = slots[blockIdx.x];
    finish_term(partials + s.partial_off, s.blocks, s.phase_k, out + 2 * uint64_t(s.out));
}

uint32_t blocks_for(uint64_t pairs) {
    return static_cast<uint32_t>(std::min<uint64_t>(kMaxBlocks, std::max<uint64_t>(1, pairs / (kThreads * 16))));
}

}  // namespace

void expectation_terms(const void* state, uint32_t n, int32_t precision, const uint64_t* x, const uint64_t* z,
                       uint32_t terms, cplx* out) {
    ExpectationItem it;
    it.state = state;
    it.n = n;
    it.terms = terms;
    it.x = x;
    it.z = z;
    it.out_index = 0;
    expectation_batch(precision, {it}, reinterpret_cast<double*>(out));
}

// The launches of one expectation call: the descriptors of every tile set,
// per-mask job and term slot, and where each lands in one device workspace
// (descriptors, block partials, output terms; 256-byte aligned regions).
struct ExpPlan {
    uintptr_t base_addr = UINTPTR_MAX;  // the lowest state address: kernels address states from it
    std::vector<ExpJob> jobs;
    std::vector<uint32_t> block_begin;
    std::vector<TileJob> tjobs;
    std::vector<TileTerm> tterms;
    std::vector<uint32_t> tblock_begin;
    std::vector<ExpSlot> slots;
    uint64_t partial_doubles = 0;
    uint32_t total_blocks = 0, tile_blocks = 0;
    uint32_t out_terms = 0;
    // Workspace offsets in bytes.
    uint64_t at_slots = 0, at_tjobs = 0, at_tterms = 0, at_tbegin = 0, at_jobs = 0, at_begin = 0;
    uint64_t at_partials = 0, at_out = 0, bytes = 0;
};

// Plans every item's terms: X masks whose bits fit a tile (with the lowest
// qubits) reduce from LDS tiles, a set of them per state read; any other X
// mask (more than kTileBits - 3 bits) takes the per-mask kernel; then one
// finishing launch runs over every term.
template <typename V>
ExpPlan plan_expectation(const std::vector<ExpectationItem>& items) {
    ExpPlan P;
    uintptr_t& base_addr = P.base_addr;
    for (const ExpectationItem& it : items) base_addr = std::min(base_addr, reinterpret_cast<uintptr_t>(it.state));
    std::vector<ExpJob>& jobs = P.jobs;
    std::vector<uint32_t>& block_begin = P.block_begin;
    std::vector<TileJob>& tjobs = P.tjobs;
    std::vector<TileTerm>& tterms = P.tterms;
    std::vector<uint32_t>& tblock_begin = P.tblock_begin;
    std::vector<ExpSlot>& slots = P.slots;
    uint64_t& partial_doubles = P.partial_doubles;
    uint32_t& total_blocks = P.total_blocks;
    uint32_t& tile_blocks = P.tile_blocks;
    auto add_slots = [&](uint64_t xm, const uint64_t* zs, const std::vector<uint32_t>& members, std::size_t at,
                         std::size_t cnt, uint32_t blocks, uint32_t out_index) {
        for (std::size_t k = 0; k < cnt; ++k) {
            const uint32_t t = members[at + k];
            ExpSlot sl;
            std::memset(&sl, 0, sizeof(sl));
            sl.partial_off = partial_doubles + uint64_t(k) * blocks * 2;
            sl.blocks = blocks;
            sl.phase_k = __builtin_popcountll(xm & zs[t]);
            sl.out = out_index + t;
            slots.push_back(sl);
        }
        partial_doubles += uint64_t(cnt) * blocks * 2;
    };
    for (const ExpectationItem& it : items) {
        if (it.n == 0 || it.n > 40) fail(MH_ERR_INVALID, "expectation: num_qubits must be 1..40");
        const uint64_t limit = (uint64_t{1} << it.n) - 1;
        std::map<uint64_t, std::vector<uint32_t>> groups;
        for (uint32_t t = 0; t < it.terms; ++t) {
            if ((it.x[t] & ~limit) || (it.z[t] & ~limit)) fail(MH_ERR_INVALID, "a Pauli mask names a qubit beyond the state");
            groups[it.x[t]].push_back(t);
        }
        const uintptr_t addr = reinterpret_cast<uintptr_t>(it.state);
        const uint64_t state_off = (addr - base_addr) / sizeof(V);
        const uint32_t k = std::min<uint32_t>(kTileBits, it.n);
        const uint64_t low = (uint64_t{1} << std::min<uint32_t>(3, k)) - 1;
        // Sets of up to kTermsPerLaunch terms whose X bits, with the low run,
        // fit k tile bits (terms taken in X-mask order).
        std::vector<std::vector<uint32_t>> sets;
        std::vector<uint64_t> wide;
        uint64_t uni = 0;
        for (const auto& [xm, members] : groups) {
            if (static_cast<uint32_t>(__builtin_popcountll(xm | low)) > k) {
                wide.push_back(xm);
                continue;
            }
            for (uint32_t t : members) {
                if (sets.empty() || sets.back().size() >= kTermsPerLaunch ||
                    static_cast<uint32_t>(__builtin_popcountll(uni | xm | low)) > k) {
                    sets.emplace_back();
                    uni = low;
                }
                uni |= xm;
                sets.back().push_back(t);
            }
        }
        const uint32_t tiles = static_cast<uint32_t>(uint64_t{1} << (it.n - k));
        // About eight tiles per workgroup for large states, one for small ones.
        const uint32_t tile_blocks_each = std::max<uint32_t>(1, std::min<uint32_t>(tiles, std::max<uint32_t>(tiles / 8, 1)));
        for (const std::vector<uint32_t>& set : sets) {
            uint64_t tmask = low;
            for (uint32_t t : set) tmask |= it.x[t];
            for (uint32_t q = 0; q < it.n && static_cast<uint32_t>(__builtin_popcountll(tmask)) < k; ++q)
                tmask |= uint64_t{1} << q;
            TileJob tj;
            std::memset(&tj, 0, sizeof(tj));
            tj.state_off = state_off;
            tj.outer_mask = limit & ~tmask;
            tj.k = k;
            tj.tiles = tiles;
            tj.blocks = tile_blocks_each;
            uint32_t at_bit = 0;
            for (uint32_t q = 0; q < it.n; ++q)
                if ((tmask >> q) & 1ull) tj.tpos[at_bit++] = static_cast<uint8_t>(q);
            tj.term_begin = static_cast<uint32_t>(tterms.size());
            tj.term_count = static_cast<uint32_t>(set.size());
            for (uint32_t t : set) {
                TileTerm tt;
                std::memset(&tt, 0, sizeof(tt));
                tt.x = it.x[t];
                tt.z = it.z[t];
                for (uint32_t i = 0; i < k; ++i) {
                    if ((tt.x >> tj.tpos[i]) & 1ull) tt.xl |= 1u << i;
                    if ((tt.z >> tj.tpos[i]) & 1ull) tt.zl |= 1u << i;
                }
                tt.hl = tt.xl ? 31u - static_cast<uint32_t>(__builtin_clz(tt.xl)) : 0u;
                tt.kind = tt.x == 0 ? kDiagonalTerm
                                    : ((__builtin_popcountll(tt.x & tt.z) & 1) ? kImagTerm : kRealTerm);
                tt.partial_off = partial_doubles;
                ExpSlot sl;
                std::memset(&sl, 0, sizeof(sl));
                sl.partial_off = partial_doubles;
                sl.blocks = tj.blocks;
                sl.phase_k = __builtin_popcountll(tt.x & tt.z);
                sl.out = it.out_index + t;
                slots.push_back(sl);
                partial_doubles += uint64_t(tj.blocks) * 2;
                tterms.push_back(tt);
            }
            tblock_begin.push_back(tile_blocks);
            tile_blocks += tj.blocks;
            tjobs.push_back(tj);
        }
        for (uint64_t xm : wide) {
            const std::vector<uint32_t>& members = groups[xm];
            for (std::size_t at = 0; at < members.size(); at += kTermsPerLaunch) {
                const std::size_t cnt = std::min<std::size_t>(kTermsPerLaunch, members.size() - at);
                ExpJob j;
                std::memset(&j, 0, sizeof(j));
                j.state_off = state_off;
                j.x = xm;
                j.h = 63u - static_cast<uint32_t>(__builtin_clzll(xm));
                j.pairs = uint64_t{1} << (it.n - 1);
                j.terms = static_cast<uint32_t>(cnt);
                j.blocks = blocks_for(j.pairs);
                j.partial_off = partial_doubles;
                for (std::size_t q = 0; q < cnt; ++q) j.z[q] = it.z[members[at + q]];
                add_slots(xm, it.z, members, at, cnt, j.blocks, it.out_index);
                block_begin.push_back(total_blocks);
                total_blocks += j.blocks;
                jobs.push_back(j);
            }
        }
    }
    for (const ExpSlot& sl : slots) P.out_terms = std::max(P.out_terms, sl.out + 1);
    uint64_t at = 0;
    auto region = [&at](uint64_t bytes) {
        const uint64_t here = at;
        at += (bytes + 255) / 256 * 256;
        return here;
    };
    P.at_slots = region(slots.size() * sizeof(ExpSlot));
    P.at_tjobs = region(tjobs.size() * sizeof(TileJob));
    P.at_tterms = region(tterms.size() * sizeof(TileTerm));
    P.at_tbegin = region(tblock_begin.size() * sizeof(uint32_t));
    P.at_jobs = region(jobs.size() * sizeof(ExpJob));
    P.at_begin = region(block_begin.size() * sizeof(uint32_t));
    P.at_partials = region(partial_doubles * sizeof(double));
    P.at_out = region(uint64_t(P.out_terms) * 2 * sizeof(double));
    P.bytes = at;
    return P;
}

// Copies a plan's descriptors into its workspace on `stream` (not waited for;
// the plan's host vectors must outlive the copy).
void upload_expectation(const ExpPlan& P, void* ws, hipStream_t stream) {
    char* w = static_cast<char*>(ws);
    auto put = [&](std::size_t at, const void* src, std::size_t bytes) {
        if (bytes) MH_HIP(hipMemcpyAsync(w + at, src, bytes, hipMemcpyHostToDevice, stream));
    };
    put(P.at_slots, P.slots.data(), P.slots.size() * sizeof(ExpSlot));
    if (!P.tjobs.empty()) {
        put(P.at_tjobs, P.tjobs.data(), P.tjobs.size() * sizeof(TileJob));
        put(P.at_tterms, P.tterms.data(), P.tterms.size() * sizeof(TileTerm));
        put(P.at_tbegin, P.tblock_begin.data(), P.tblock_begin.size() * sizeof(uint32_t));
    }
    if (!P.jobs.empty()) {
        put(P.at_jobs, P.jobs.data(), P.jobs.size() * sizeof(ExpJob));
        put(P.at_begin, P.block_begin.data(), P.block_begin.size() * sizeof(uint32_t));
    }
}

// The reduction kernels of a plan whose descriptors are in the workspace
// (on `stream`; no copies). The terms land in the workspace's output region,
// or at out[2 * term] when `out` is given.
template <typename V>
void launch_expectation(const ExpPlan& P, void* ws, hipStream_t stream, double* out = nullptr) {
    char* w = static_cast<char*>(ws);
    double* partials = reinterpret_cast<double*>(w + P.at_partials);
    double* dout = out ? out : reinterpret_cast<double*>(w + P.at_out);
    if (!P.tjobs.empty()) {
        hipLaunchKernelGGL(expectation_tile_kernel<V>, dim3(P.tile_blocks), dim3(kThreads), 0, stream,
                           reinterpret_cast<const TileJob*>(w + P.at_tjobs),
                           reinterpret_cast<const uint32_t*>(w + P.at_tbegin), static_cast<uint32_t>(P.tjobs.size()),
                           reinterpret_cast<const TileTerm*>(w + P.at_tterms), reinterpret_cast<const V*>(P.base_addr),
                           partials);
        MH_HIP(hipGetLastError());
    }
    if (!P.jobs.empty()) {
        hipLaunchKernelGGL(expectation_batch_kernel<V>, dim3(P.total_blocks), dim3(kThreads), 0, stream,
                           reinterpret_cast<const ExpJob*>(w + P.at_jobs),
                           reinterpret_cast<const uint32_t*>(w + P.at_begin), static_cast<uint32_t>(P.jobs.size()),
                           reinterpret_cast<const V*>(P.base_addr), partials);
        MH_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(expectation_finish_batch_kernel, dim3(static_cast<uint32_t>(P.slots.size())), dim3(kThreads), 0,
                       stream, reinterpret_cast<const ExpSlot*>(w + P.at_slots),
                       static_cast<const double*>(partials), dout);
    MH_HIP(hipGetLastError());
}

// Runs a plan with its workspace at `ws` (P.bytes bytes of device memory,
// nothing in it need be zero: every partial and output is written before it is
// read), then copies the terms to out[2 * term]. Uses the calling thread's main
// stream and waits for it.
template <typename V>
void run_expectation(const ExpPlan& P, void* ws, double* out) {
    if (P.slots.empty()) return;
    upload_expectation(P, ws, streams().main);
    launch_expectation<V>(P, ws, streams().main);
    const double* dout = reinterpret_cast<const double*>(static_cast<const char*>(ws) + P.at_out);
    std::vector<double> host(uint64_t(P.out_terms) * 2);
    download(host.data(), dout, host.size() * sizeof(double));
    for (const ExpSlot& sl : P.slots) {
        out[2 * uint64_t(sl.out)] = host[2 * uint64_t(sl.out)];
        out[2 * uint64_t(sl.out) + 1] = host[2 * uint64_t(sl.out) + 1];
    }
}

// Every item's terms, with the workspace from the pool.
template <typename V>
void expectation_batch_typed(const std::vector<ExpectationItem>& items, double* out) {
    context();
    const ExpPlan P = plan_expectation<V>(items);
    if (P.slots.empty()) return;
    DeviceBuffer ws(P.bytes);
    run_expectation<V>(P, ws.get(), out);
}

struct ExpectationProgram::Impl {
    int32_t precision = MH_C128;
    ExpPlan plan;
};

ExpectationProgram::ExpectationProgram(int32_t precision, const std::vector<ExpectationItem>& items)
    : impl_(std::make_unique<Impl>()) {
    amp_bytes(precision);
    impl_->precision = precision;
    impl_->plan = precision == MH_C128 ? plan_expectation<double2>(items) : plan_expectation<float2>(items);
}

ExpectationProgram::~ExpectationProgram() = default;

uint64_t ExpectationProgram::workspace_bytes() const { return std::max<uint64_t>(impl_->plan.bytes, 256); }

void ExpectationProgram::upload_on(void* workspace, hipStream_t stream) const {
    if (impl_->plan.slots.empty()) return;
    upload_expectation(impl_->plan, workspace, stream);
}

void ExpectationProgram::launch_on(void* workspace, hipStream_t stream, double* out) const {
    if (impl_->plan.slots.empty()) return;
    if (impl_->precision == MH_C128)
        launch_expectation<double2>(impl_->plan, workspace, stream, out);
    else
        launch_expectation<float2>(impl_->plan, workspace, stream, out);
}

void expectation_batch(int32_t precision, const std::vector<ExpectationItem>& items, double* out) {
    amp_bytes(precision);
    if (precision == MH_C128)
        expectation_batch_typed<double2>(items, out);
    else
        expectation_batch_typed<float2>(items, out);
}

namespace {

// Largest |a_i - b_i| (complex modulus, FP64) over two states of any precisions:
// one partial per workgroup.
template <typename VA, typename VB>
__global__ void __launch_bounds__(kThreads) max_diff_kernel(const VA* a, const VB* b, uint64_t size, double* partials) {
    __shared__ double red[kThreads];
    double m = 0.0;
    const uint64_t stride = uint64_t(gridDim.x) * kThreads;
    for (uint64_t i = uint64_t(blockIdx.x) * kThreads + threadIdx.x; i < size; i += stride) {
        const double dr = static_cast<double>(a[i].x) - static_cast<double>(b[i].x);
        const double di = static_cast<double>(a[i].y) - static_cast<double>(b[i].y);
        m = fmax(m, sqrt(dr * dr + di * di));
    }
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = kThreads / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = red[0];
}

template <typename VA, typename VB>
double max_diff(const void* a, const void* b, uint64_t size) {
    context();
    const uint32_t grid = static_cast<uint32_t>(std::min<uint64_t>(4096, std::max<uint64_t>(1, size / kThreads)));
    DeviceBuffer partials(grid * sizeof(double));
    hipLaunchKernelGGL((max_diff_kernel<VA, VB>), dim3(grid), dim3(kThreads), 0, streams().main, static_cast<const VA*>(a),
                       static_cast<const VB*>(b), size, partials.as<double>());
    MH_HIP(hipGetLastError());
    std::vector<double> host(grid);
    download(host.data(), partials.get(), grid * sizeof(double));
    return *std::max_element(host.begin(), host.end());
}

}  // namespace

}  // namespace mh

namespace mh {
namespace {

// The workspace of mh_expectation_noalloc: one array per device, held for the
// process's lifetime and shared by every call on that device (each call
// overwrites it; none needs it zeroed). Host code never passes the array's
// name to a kernel or a copy (there it is a host-side stand-in): it uses the
// device address hipGetSymbolAddress gives for the current device.
constexpr uint64_t kStaticWorkspaceBytes = uint64_t{32} << 20;
static __device__ __align__(256) unsigned char scratch[kStaticWorkspaceBytes];

// The device address of `scratch` on the current device, looked up on the
// first call there (with that device current) and kept per device.
void* static_workspace() {
    context();
    int device = 0;
    MH_HIP(hipGetDevice(&device));
    static std::mutex mu;
    static std::map<int, void*> per_device;
    std::lock_guard<std::mutex> lock(mu);
    const auto it = per_device.find(device);
    if (it != per_device.end()) return it->second;
    void* ptr = nullptr;
    MH_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (ptr == nullptr) fail(MH_ERR_INTERNAL, "hipGetSymbolAddress gave no address for the static workspace");
    per_device.emplace(device, ptr);
    return ptr;
}

template <typename V>
void expectation_static(const ExpectationItem& item, double* out) {
    const ExpPlan P = plan_expectation<V>({item});
    if (P.slots.empty()) return;
    if (P.bytes > kStaticWorkspaceBytes)
        fail(MH_ERR_INVALID, "mh_expectation_noalloc: these terms need " + std::to_string(P.bytes) +
                                 " bytes of workspace; the static workspace holds " +
                                 std::to_string(kStaticWorkspaceBytes) + " (use mh_expectation)");
    run_expectation<V>(P, static_workspace(), out);
}

}  // namespace
}  // namespace mh

using namespace mh;

extern "C" int mh_expectation(const void* state, uint32_t num_qubits, int32_t precision, const uint64_t* x_masks,
                              const uint64_t* z_masks, uint32_t num_terms, double* out) {
    return guarded([&] {
        expectation_terms(state, num_qubits, precision, x_masks, z_masks, num_terms, reinterpret_cast<cplx*>(out));
    });
}

extern "C" int mh_state_max_abs_diff(const void* a, int32_t precision_a, const void* b, int32_t precision_b,
                                     uint32_t num_qubits, double* out) {
    return guarded([&] {
        amp_bytes(precision_a);
        amp_bytes(precision_b);
        if (num_qubits > 40) fail(MH_ERR_INVALID, "num_qubits above 40");
        if (out == nullptr) fail(MH_ERR_INVALID, "null output");
        const uint64_t size = uint64_t{1} << num_qubits;
        if (precision_a == MH_C128 && precision_b == MH_C128)
            *out = max_diff<double2, double2>(a, b, size);
        else if (precision_a == MH_C128)
            *out = max_diff<double2, float2>(a, b, size);
        else if (precision_b == MH_C128)
            *out = max_diff<float2, double2>(a, b, size);
        else
            *out = max_diff<float2, float2>(a, b, size);
    });
}

extern "C" int mh_expectation_noalloc(const void* state, uint32_t num_qubits, int32_t precision,
                                      const uint64_t* x_masks, const uint64_t* z_masks, uint32_t num_terms,
                                      double* out) {
    return guarded([&] {
        amp_bytes(precision);
        if (num_terms > 0 && (x_masks == nullptr || z_masks == nullptr || out == nullptr))
            fail(MH_ERR_INVALID, "mh_expectation_noalloc: null term masks or output");
        ExpectationItem it;
        it.state = state;
        it.n = num_qubits;
        it.terms = num_terms;
        it.x = x_masks;
        it.z = z_masks;
        if (precision == MH_C128)
            expectation_static<double2>(it, out);
        else
            expectation_static<float2>(it, out);
    });
}
