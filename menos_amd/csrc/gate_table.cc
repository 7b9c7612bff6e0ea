// Named-gate matrices and the initial phase, from host libm.
//
// This file is compiled with g++, not with hipcc's clang: the matrix bits are
// part of the bit-exact contract (reference src/ir/gate_table.cpp:69-98 builds
// them with std::complex and libm), and the two compilers lower
// std::exp(std::complex<double>) differently -- g++ calls glibc's cexp, clang
// expands it into separate exp/cos/sin calls -- which differ in the last bit
// for some angles. The reference is built with g++, so this is too.
#include <cmath>
#include <complex>
#include <numbers>

#include "gate_table.h"

namespace mh {

namespace {
using C = std::complex<double>;
using namespace std::complex_literals;
constexpr double kPi = std::numbers::pi_v<double>;

enum : int32_t { kX, kY, kZ, kH, kS, kSdg, kT, kTdg, kSX, kSXdg, kRX, kRY, kRZ, kP, kU, kCX, kCZ, kSWAP, kCP, kRZZ, kCCX };

void set4(C* o, C a, C b, C c, C d) {
    o[0] = a;
    o[1] = b;
    o[2] = c;
    o[3] = d;
}

void diag(C* o, std::size_t dim, std::size_t at, C v) { o[at * dim + at] = v; }
}  // namespace

// reference src/ir/gate_table.cpp:19-206: the same std::complex expressions.
void host_gate_matrix(int32_t op, const double* p, C* o) {
    switch (op) {
        case kX: set4(o, 0.0, 1.0, 1.0, 0.0); return;
        case kY: set4(o, 0.0, -1.0i, 1.0i, 0.0); return;
        case kZ: set4(o, 1.0, 0.0, 0.0, -1.0); return;
        case kH: {
            const double s = 1.0 / std::sqrt(2.0);
            set4(o, s, s, s, -s);
            return;
        }
        case kS: set4(o, 1.0, 0.0, 0.0, 1.0i); return;
        case kSdg: set4(o, 1.0, 0.0, 0.0, -1.0i); return;
        case kT: set4(o, 1.0, 0.0, 0.0, std::exp(1.0i * (kPi / 4.0))); return;
        case kTdg: set4(o, 1.0, 0.0, 0.0, std::exp(-1.0i * (kPi / 4.0))); return;
        case kSX: set4(o, 0.5 * (1.0 + 1.0i), 0.5 * (1.0 - 1.0i), 0.5 * (1.0 - 1.0i), 0.5 * (1.0 + 1.0i)); return;
        case kSXdg: set4(o, 0.5 * (1.0 - 1.0i), 0.5 * (1.0 + 1.0i), 0.5 * (1.0 + 1.0i), 0.5 * (1.0 - 1.0i)); return;
        case kRX: {
            const double c = std::cos(p[0] / 2.0);
            const double s = std::sin(p[0] / 2.0);
            set4(o, c, -1.0i * s, -1.0i * s, c);
            return;
        }
        case kRY: {
            const double c = std::cos(p[0] / 2.0);
            const double s = std::sin(p[0] / 2.0);
            set4(o, c, -s, s, c);
            return;
        }
        case kRZ: set4(o, std::exp(-1.0i * (p[0] / 2.0)), 0.0, 0.0, std::exp(1.0i * (p[0] / 2.0))); return;
        case kP: set4(o, 1.0, 0.0, 0.0, std::exp(1.0i * p[0])); return;
        case kU: {
            const double c = std::cos(p[0] / 2.0);
            const double s = std::sin(p[0] / 2.0);
            set4(o, c, -std::exp(1.0i * p[2]) * s, std::exp(1.0i * p[1]) * s, std::exp(1.0i * (p[1] + p[2])) * c);
            return;
        }
        default: break;
    }
    const std::size_t dim = op == kCCX ? 8 : 4;
    for (std::size_t e = 0; e < dim * dim; ++e) o[e] = C{0.0, 0.0};
    switch (op) {
        case kCX:
            o[0] = 1.0; o[7] = 1.0; o[10] = 1.0; o[13] = 1.0;
            return;
        case kCZ:
            diag(o, 4, 0, 1.0); diag(o, 4, 1, 1.0); diag(o, 4, 2, 1.0); diag(o, 4, 3, -1.0);
            return;
        case kSWAP:
            o[0] = 1.0; o[6] = 1.0; o[9] = 1.0; o[15] = 1.0;
            return;
        case kCP:
            diag(o, 4, 0, 1.0); diag(o, 4, 1, 1.0); diag(o, 4, 2, 1.0); diag(o, 4, 3, std::exp(1.0i * p[0]));
            return;
        case kRZZ: {
            const C same = std::exp(-0.5i * p[0]);
            const C differ = std::exp(0.5i * p[0]);
            diag(o, 4, 0, same); diag(o, 4, 1, differ); diag(o, 4, 2, differ); diag(o, 4, 3, same);
            return;
        }
        default:  // kCCX
            for (std::size_t i = 0; i < 8; ++i) diag(o, 8, i, 1.0);
            o[3 * 8 + 3] = 0.0; o[7 * 8 + 7] = 0.0; o[3 * 8 + 7] = 1.0; o[7 * 8 + 3] = 1.0;
            return;
    }
}

C host_phase(double global_phase) { return std::exp(C{0.0, global_phase}); }

}  // namespace mh
