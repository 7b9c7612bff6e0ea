// The per-group gate kernel (menos_amd/csrc/gates.hip, generic_gate_kernel)
// replayed on the CPU: every lane of every workgroup of the launch the host
// makes (group_grid), the grid-stride loop, the index arithmetic of
// gate_index.h and the kernel's complex128 arithmetic order, with every
// access bounds-checked and every amplitude required to be read and written
// exactly once. The result must equal the oracle's generic kernel bit for bit.
// Covers k = 1..6 on n = 10..16 qubits, including operands on the lowest and
// highest qubits, in ascending, descending and shuffled operand order.
#define __host__
#define __device__
#include "gate_index.h"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "oracle.h"

namespace {

int failures = 0;

void fail_case(const char* what, uint32_t n, uint32_t k) {
    std::fprintf(stderr, "n=%u k=%u: %s\n", n, k, what);
    ++failures;
}

// One launch of generic_gate_kernel<double, k> replayed lane by lane.
void replay(std::vector<double>& state, uint32_t n, const std::vector<double>& m, const std::vector<uint32_t>& ops) {
    const uint32_t k = static_cast<uint32_t>(ops.size());
    uint32_t sorted[6];
    std::copy(ops.begin(), ops.end(), sorted);
    std::sort(sorted, sorted + k);
    const uint64_t size = uint64_t{1} << n;
    const uint64_t groups = size >> k;
    const uint32_t grid = mh::group_grid(groups);
    const uint64_t stride = uint64_t(grid) * mh::kGroupThreads;
    std::vector<uint8_t> reads(size, 0), writes(size, 0);
    const uint32_t dim = 1u << k;
    for (uint32_t block = 0; block < grid; ++block)
        for (uint32_t t = 0; t < mh::kGroupThreads; ++t)
            for (uint64_t g = uint64_t(block) * mh::kGroupThreads + t; g < groups; g += stride) {
                const uint64_t base = mh::group_base(g, sorted, k);
                double v[64][2];
                for (uint32_t l = 0; l < dim; ++l) {
                    const uint64_t idx = mh::group_member(base, l, ops.data(), k);
                    if (idx >= size) {
                        fail_case("read out of range", n, k);
                        return;
                    }
                    ++reads[idx];
                    v[l][0] = state[2 * idx];
                    v[l][1] = state[2 * idx + 1];
                }
                for (uint32_t o = 0; o < dim; ++o) {
                    double sr = 0.0, si = 0.0;
                    for (uint32_t in = 0; in < dim; ++in) {
                        const double mr = m[2 * (o * dim + in)], mi = m[2 * (o * dim + in) + 1];
                        const double pr = mr * v[in][0] - mi * v[in][1];
                        const double pi = mr * v[in][1] + mi * v[in][0];
                        sr = sr + pr;
                        si = si + pi;
                    }
                    const uint64_t idx = mh::group_member(base, o, ops.data(), k);
                    if (idx >= size) {
                        fail_case("write out of range", n, k);
                        return;
                    }
                    ++writes[idx];
                    state[2 * idx] = sr;
                    state[2 * idx + 1] = si;
                }
            }
    for (uint64_t i = 0; i < size; ++i)
        if (reads[i] != 1 || writes[i] != 1) {
            fail_case("an amplitude not read and written exactly once", n, k);
            return;
        }
}

}  // namespace

int main() {
    std::mt19937_64 g(20261203);
    std::uniform_real_distribution<double> u(-1.0, 1.0);
    uint32_t cases = 0;
    for (uint32_t n = 10; n <= 16; n += 2)
        for (uint32_t k = 1; k <= 6; ++k)
            for (int variant = 0; variant < 4; ++variant) {
                std::vector<uint32_t> q(n);
                for (uint32_t i = 0; i < n; ++i) q[i] = i;
                std::vector<uint32_t> ops;
                if (variant == 0) {
                    ops.assign(q.begin(), q.begin() + k);                       // the lowest qubits, ascending
                } else if (variant == 1) {
                    ops.assign(q.end() - k, q.end());                            // the highest, descending
                    std::reverse(ops.begin(), ops.end());
                } else {
                    std::shuffle(q.begin(), q.end(), g);                         // anywhere, any order
                    ops.assign(q.begin(), q.begin() + k);
                }
                const uint32_t dim = 1u << k;
                std::vector<double> m(2 * dim * dim), state(2 * (std::size_t{1} << n));
                for (double& x : m) x = u(g);
                if (variant == 3) m[2] = -0.0;                                   // signed-zero entries
                for (double& x : state) x = u(g);
                std::vector<double> want = state;
                orc_apply_generic(want.data(), n, m.data(), ops.data(), k, 1);
                replay(state, n, m, ops);
                if (std::memcmp(state.data(), want.data(), state.size() * sizeof(double)) != 0)
                    fail_case("result differs from the oracle", n, k);
                ++cases;
            }
    // A launch over more groups than one grid of kGroupMaxBlocks workgroups
    // covers: the stride loop's index stays in range (checked arithmetically).
    {
        const uint64_t groups = (uint64_t{1} << 33) >> 1;
        const uint32_t grid = mh::group_grid(groups);
        const uint64_t stride = uint64_t(grid) * mh::kGroupThreads;
        if (grid != mh::kGroupMaxBlocks || groups % stride != 0) ++failures;
    }
    if (failures) {
        std::fprintf(stderr, "%d failures\n", failures);
        return 1;
    }
    std::printf("generic kernel replay: %u cases bit-exact, every access in range\n", cases);
    return 0;
}
