// Batched execution of many independent circuits (mh_batch_execute).
//
// The reference runs a batch as one full `run` per circuit
// (src/api/batch.cpp:10-27); its CUDA engine pays, per circuit, a plan, an
// allocation, a launch and a synchronisation per block, and a copy back. Here
// the per-circuit cost is host planning (fusion plan and block matrices, on
// host threads) and a share of a few launches:
//
//   1. every circuit is validated and planned (host threads);
//   2. circuits are taken in order into rounds whose states fit round_bytes;
//      a round's states share one pooled device arena;
//   3. per round: one clear of the arena plus one kernel that writes each
//      state's first amplitude (the global phase), then pass p of every
//      circuit in one launch (run_window_passes_batch), then every circuit's
//      expectation terms in two launches (expectation_batch), one copy back.
//
// The passes and reductions are the single-circuit engine's kernels, so
// results equal mh_engine_execute's bit for bit.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>
#include <exception>
#include <mutex>
#include <thread>

#include "gate_table.h"
#include "mh_internal.h"
#include "window.h"

namespace mh {
namespace {

constexpr int kThreads = 256;

template <typename V>
__global__ void __launch_bounds__(kThreads) first_amplitude_kernel(V* base, const uint64_t* offsets,
                                                                   const double* phases, uint32_t count) {
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= count) return;
    V v;
    v.x = static_cast<decltype(v.x)>(phases[2 * i]);
    v.y = static_cast<decltype(v.y)>(phases[2 * i + 1]);
    base[offsets[i]] = v;
}

struct Planned {
    uint32_t n = 0;
    std::vector<Block> blocks;
};

template <typename F>
void parallel_for(std::size_t count, uint32_t threads, F&& body) {
    std::atomic<std::size_t> next{0};
    std::exception_ptr error;
    std::mutex m;
    auto worker = [&] {
        try {
            for (std::size_t i; (i = next.fetch_add(1)) < count;) body(i);
        } catch (...) {
            std::lock_guard<std::mutex> lock(m);
            if (!error) error = std::current_exception();
        }
    };
    std::vector<std::thread> pool;
    const std::size_t t = std::max<std::size_t>(1, std::min<std::size_t>(threads, count));
    for (std::size_t k = 1; k < t; ++k) pool.emplace_back(worker);
    worker();
    for (auto& th : pool) th.join();
    if (error) std::rethrow_exception(error);
}

}  // namespace
}  // namespace mh

using namespace mh;

extern "C" int mh_batch_execute(const mh_batch_request* req, mh_batch_result* res) {
    return guarded([&] {
        using clock = std::chrono::steady_clock;
        const auto t0 = clock::now();
        if (req == nullptr || res == nullptr) fail(MH_ERR_INVALID, "null request or result");
        const std::size_t amp = amp_bytes(req->precision);
        const uint32_t count = req->num_circuits;
        if (count == 0) return;
        if (req->circuits == nullptr || req->term_offsets == nullptr) fail(MH_ERR_INVALID, "null circuits or term offsets");
        if (req->term_offsets[0] != 0) fail(MH_ERR_INVALID, "term_offsets must start at 0");
        const uint32_t total_terms = req->term_offsets[count];
        if (total_terms > 0 && (req->x_masks == nullptr || req->z_masks == nullptr || res->expectation_terms == nullptr))
            fail(MH_ERR_INVALID, "terms without masks or an output");
        if (req->want_statevectors && res->statevectors == nullptr) fail(MH_ERR_INVALID, "no statevector output");
        const uint32_t width = window_width(req->precision);
        uint32_t threads = req->host_threads;
        if (threads == 0) threads = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));

        // 1. Plans.
        std::vector<Planned> plans(count);
        parallel_for(count, threads, [&](std::size_t i) {
            const mh_circuit* c = &req->circuits[i];
            validate_circuit(c);
            if (c->num_qubits < width || c->num_qubits > 30)
                fail(MH_ERR_INVALID, "a batched circuit needs " + std::to_string(width) + " to 30 qubits");
            if (req->term_offsets[i + 1] < req->term_offsets[i]) fail(MH_ERR_INVALID, "term_offsets must not decrease");
            const CircuitView cv{c};
            const FusionPlan fp = plan_fusion(cv, req->fusion, req->fusion_width);
            plans[i].n = c->num_qubits;
            plans[i].blocks = build_blocks(cv, fp);
        });
        res->plan_seconds = std::chrono::duration<double>(clock::now() - t0).count();

        // 2. Rounds.
        uint64_t round_bytes = req->round_bytes;
        if (round_bytes == 0) {
            size_t free_b = 0, total_b = 0;
            MH_HIP(hipMemGetInfo(&free_b, &total_b));
            round_bytes = (static_cast<uint64_t>(free_b) + pool_reusable_bytes()) / 2;
        }
        constexpr uint64_t kAlign = 256;
        Context& ctx = context();
        uint64_t sv_at = 0;
        uint32_t rounds = 0, launches = 0;
        double device_seconds = 0;
        for (uint32_t first = 0; first < count;) {
            uint32_t last = first;
            uint64_t bytes = 0;
            while (last < count) {
                const uint64_t b = ((uint64_t{1} << plans[last].n) * amp + kAlign - 1) / kAlign * kAlign;
                if (last > first && bytes + b > round_bytes) break;
                bytes += b;
                ++last;
            }
            const auto r0 = clock::now();
            DeviceBuffer arena(bytes);
            std::vector<uint64_t> offsets(last - first);  // amplitudes from the arena start
            std::vector<double> phases(2 * (last - first));
            std::vector<BatchItem> items(last - first);
            std::vector<ExpectationItem> exp_items;
            uint64_t at = 0;
            for (uint32_t i = first; i < last; ++i) {
                const uint32_t k = i - first;
                offsets[k] = at / amp;
                items[k].state = static_cast<unsigned char*>(arena.get()) + at;
                items[k].n = plans[i].n;
                items[k].blocks = &plans[i].blocks;
                const std::complex<double> ph = host_phase(req->global_phases ? req->global_phases[i] : 0.0);
                phases[2 * k] = ph.real();
                phases[2 * k + 1] = ph.imag();
                const uint32_t t_begin = req->term_offsets[i], t_end = req->term_offsets[i + 1];
                if (t_end > t_begin) {
                    ExpectationItem e;
                    e.state = items[k].state;
                    e.n = plans[i].n;
                    e.terms = t_end - t_begin;
                    e.x = req->x_masks + t_begin;
                    e.z = req->z_masks + t_begin;
                    e.out_index = t_begin;
                    exp_items.push_back(e);
                }
                at += ((uint64_t{1} << plans[i].n) * amp + kAlign - 1) / kAlign * kAlign;
            }
            // 3. Initialise, apply, reduce.
            MH_HIP(hipMemsetAsync(arena.get(), 0, bytes, ctx.stream));
            DeviceBuffer doffsets(offsets.size() * sizeof(uint64_t)), dphases(phases.size() * sizeof(double));
            upload(doffsets.get(), offsets.data(), offsets.size() * sizeof(uint64_t));
            upload(dphases.get(), phases.data(), phases.size() * sizeof(double));
            const uint32_t n_here = last - first;
            const dim3 grid((n_here + kThreads - 1) / kThreads), block(kThreads);
            if (req->precision == MH_C128)
                hipLaunchKernelGGL(first_amplitude_kernel<double2>, grid, block, 0, ctx.stream, arena.as<double2>(),
                                   static_cast<const uint64_t*>(doffsets.as<uint64_t>()),
                                   static_cast<const double*>(dphases.as<double>()), n_here);
            else
                hipLaunchKernelGGL(first_amplitude_kernel<float2>, grid, block, 0, ctx.stream, arena.as<float2>(),
                                   static_cast<const uint64_t*>(doffsets.as<uint64_t>()),
                                   static_cast<const double*>(dphases.as<double>()), n_here);
            MH_HIP(hipGetLastError());
            BatchStats stats;
            run_window_passes_batch(req->precision, items, threads, &stats);
            launches += 2 + stats.launches;
            if (!exp_items.empty()) {
                expectation_batch(req->precision, exp_items, res->expectation_terms);
                launches += 2;
            }
            if (req->want_statevectors) {
                for (uint32_t i = first; i < last; ++i) {
                    const uint64_t sb = (uint64_t{1} << plans[i].n) * amp;
                    download(static_cast<unsigned char*>(res->statevectors) + sv_at, items[i - first].state, sb);
                    sv_at += sb;
                }
            }
            synchronize();
            device_seconds += std::chrono::duration<double>(clock::now() - r0).count();
            const double done = std::chrono::duration<double>(clock::now() - t0).count();
            if (res->completion_seconds)
                for (uint32_t i = first; i < last; ++i) res->completion_seconds[i] = done;
            ++rounds;
            first = last;
        }
        res->rounds = rounds;
        res->launches = launches;
        res->device_seconds = device_seconds;
        res->total_seconds = std::chrono::duration<double>(clock::now() - t0).count();
    });
}
