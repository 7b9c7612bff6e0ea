// Device calibration for the time model (replaces the reference's
// measure_cuda_calibration, src/plan/cuda_calibration.cpp:42-153, and the
// transfer and launch probes of src/kernels/cuda/device_benchmarks.cpp:27-74).
//
// Measured once per call on this device: the HBM copy bandwidth (a streaming
// copy, bytes read + written), a window pass at 26 qubits (complex128, eight
// 1-qubit gates), the synchronised launch cost of an empty kernel, pinned
// host<->device bandwidth and the separate (non-fused) FP64 multiply/add rate.
// Each figure is the best of a few repetitions after a warm-up.
#include <algorithm>
#include <chrono>
#include <cmath>

#include "mh_internal.h"
#include "window.h"

namespace mh {

void set_calibration(const mh_calibration& c);

namespace {

constexpr int kThreads = 256;

__global__ void __launch_bounds__(kThreads) copy_kernel(const double2* __restrict__ src, double2* __restrict__ dst,
                                                        uint64_t n) {
    const uint64_t stride = uint64_t(gridDim.x) * kThreads;
    for (uint64_t i = uint64_t(blockIdx.x) * kThreads + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

__global__ void empty_kernel() {}

// 4 independent chains of alternating multiply and add per thread.
__global__ void __launch_bounds__(kThreads) fp64_kernel(double* out, uint32_t iters, double a, double b) {
    double x0 = threadIdx.x, x1 = x0 + 1.0, x2 = x0 + 2.0, x3 = x0 + 3.0;
    for (uint32_t i = 0; i < iters; ++i) {
        x0 = x0 * a; x1 = x1 * a; x2 = x2 * a; x3 = x3 * a;
        x0 = x0 + b; x1 = x1 + b; x2 = x2 + b; x3 = x3 + b;
        x0 = x0 * a; x1 = x1 * a; x2 = x2 * a; x3 = x3 * a;
        x0 = x0 + b; x1 = x1 + b; x2 = x2 + b; x3 = x3 + b;
    }
    if (x0 + x1 + x2 + x3 == 12345.678) out[0] = x0;
}

template <typename F>
double best_seconds(int reps, F&& f) {
    context();
    hipEvent_t e0, e1;
    MH_HIP(hipEventCreate(&e0));
    MH_HIP(hipEventCreate(&e1));
    f();  // warm-up
    double best = 1e30;
    for (int i = 0; i < reps; ++i) {
        MH_HIP(hipEventRecord(e0, streams().main));
        f();
        MH_HIP(hipEventRecord(e1, streams().main));
        MH_HIP(hipEventSynchronize(e1));
        float ms = 0;
        MH_HIP(hipEventElapsedTime(&ms, e0, e1));
        best = std::min(best, ms * 1e-3);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return best;
}

}  // namespace

// A streaming copy of `bytes` between two device buffers (bytes read + bytes written).
void launch_copy(const void* src, void* dst, uint64_t bytes) {
    context();
    const uint64_t n = bytes / sizeof(double2);
    const uint32_t grid = static_cast<uint32_t>(std::min<uint64_t>((n + kThreads - 1) / kThreads, 1u << 20));
    hipLaunchKernelGGL(copy_kernel, dim3(grid), dim3(kThreads), 0, streams().main, static_cast<const double2*>(src),
                       static_cast<double2*>(dst), n);
    MH_HIP(hipGetLastError());
}

}  // namespace mh

using namespace mh;

extern "C" int mh_calibrate(mh_calibration* out) {
    return guarded([&] {
        if (out == nullptr) fail(MH_ERR_INVALID, "null output");
        Context& ctx = context();
        mh_calibration c{};
        {
            const uint64_t bytes = uint64_t{1} << 31;  // 2 GiB each way
            DeviceBuffer a(bytes), b(bytes);
            MH_HIP(hipMemsetAsync(a.get(), 0, bytes, streams().main));
            const double s = best_seconds(5, [&] { launch_copy(a.get(), b.get(), bytes); });
            c.hbm_bytes_per_second = 2.0 * static_cast<double>(bytes) / s;
        }
        {
            const uint32_t n = 26;
            DeviceBuffer st((uint64_t{1} << n) * 16);
            launch_init(st.get(), n, MH_C128, 1.0, 0.0);
            std::vector<Block> blocks;
            for (uint32_t q = 0; q < 8; ++q) {
                Block b;
                b.operands = {q * 3};
                const double t = 0.1 + 0.2 * q;
                b.matrix = {std::cos(t), -std::sin(t), std::sin(t), std::cos(t)};
                blocks.push_back(b);
            }
            PassStats stats;
            const double s = best_seconds(3, [&] { apply_blocks(st.get(), n, MH_C128, blocks, 0, false, &stats, 0); });
            c.pass_bytes_per_second = 2.0 * 16.0 * static_cast<double>(uint64_t{1} << n) * stats.passes / s;
        }
        c.launch_seconds = best_seconds(20, [&] {
            hipLaunchKernelGGL(empty_kernel, dim3(1), dim3(64), 0, streams().main);
            synchronize();
        });
        {
            const uint64_t bytes = uint64_t{256} << 20;
            void* host = nullptr;
            MH_HIP(hipHostMalloc(&host, bytes, hipHostMallocDefault));
            DeviceBuffer d(bytes);
            try {
                const double h2d = best_seconds(3, [&] {
                    MH_HIP(hipMemcpyAsync(d.get(), host, bytes, hipMemcpyHostToDevice, streams().main));
                });
                const double d2h = best_seconds(3, [&] {
                    MH_HIP(hipMemcpyAsync(host, d.get(), bytes, hipMemcpyDeviceToHost, streams().main));
                });
                c.h2d_bytes_per_second = static_cast<double>(bytes) / h2d;
                c.d2h_bytes_per_second = static_cast<double>(bytes) / d2h;
            } catch (...) {
                (void)hipHostFree(host);
                throw;
            }
            MH_HIP(hipHostFree(host));
        }
        {
            DeviceBuffer sink(64);
            const uint32_t iters = 4096;
            const uint32_t grid = static_cast<uint32_t>(ctx.props.multiProcessorCount) * 16;
            const double s = best_seconds(3, [&] {
                hipLaunchKernelGGL(fp64_kernel, dim3(grid), dim3(kThreads), 0, streams().main, sink.as<double>(), iters,
                                   1.0000001, 1e-9);
            });
            c.fp64_ops_per_second = static_cast<double>(grid) * kThreads * iters * 16.0 / s;
        }
        {
            // A complex64 6-qubit block on the matrix cores at 26 qubits
            // (arithmetic-bound): 8 * 64 flops per amplitude.
            const uint32_t n = 26, k = 6;
            DeviceBuffer st((uint64_t{1} << n) * 8);
            launch_init(st.get(), n, MH_C64, 1.0, 0.0);
            std::vector<cplx> m(std::size_t{1} << (2 * k));
            for (std::size_t e = 0; e < m.size(); ++e) m[e] = cplx{e % 65 == 0 ? 1.0 : 0.0, 0.0};
            const uint32_t ops[] = {3, 7, 11, 15, 19, 23};
            const double s = best_seconds(3, [&] { launch_generic(st.get(), n, MH_C64, m.data(), ops, k); });
            c.mfma_flops_per_second = 8.0 * 64.0 * static_cast<double>(uint64_t{1} << n) / s;
        }
        synchronize();
        set_calibration(c);
        *out = c;
    });
}
