// Host libm work of the execution plan, kept in one translation unit that
// build.py compiles with g++ (gate_table.cc).
#pragma once

#include <complex>
#include <cstdint>

namespace mh {

// Writes the row-major 2^k x 2^k matrix of a named gate (reference opcode
// numbering, include/menos/circuit.h) into out; params as the gate takes them.
void host_gate_matrix(int32_t opcode, const double* params, std::complex<double>* out);

// exp(i * global_phase), the initial amplitude (reference
// src/kernels/cpu/dense_kernels.cpp:21-27).
std::complex<double> host_phase(double global_phase);

}  // namespace mh
