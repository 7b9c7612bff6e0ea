// A resident batch replayed from a captured HIP graph (mh_batch_graph_*).
//
// The variational estimator runs the same circuits again and again with new
// parameters. The reference reruns each one from scratch (run_batch_with_options,
// src/api/batch.cpp:10-27; plan reuse by plan_job_cached /
// reparameterize_direct_plan, src/plan/plan_cache.cpp:233-276). Here the whole
// round is made resident once: fusion plans, pass programs, the state arena,
// every device table and the term-reduction descriptors; its device work is
// captured once into a graph:
//
//   matrix table <- pinned staging (copy node; it also holds each circuit's
//                   initial phase)
//   every pass of every circuit (the batch program's launches, forked over the
//   kernel streams and joined back; each circuit's first pass generates its
//   initial state, so the arena needs no clearing)
//   term reductions (kernels)
//   terms -> pinned staging (copy node)
//
// A parameter update rebuilds the block matrices and pass programs on host
// threads and, when every circuit's program keeps its shape (same passes, ops,
// sign terms and launch lists: only matrix entries differ), writes the new
// matrix table into the pinned staging the graph copies from. A run is one graph
// launch. Results are the same kernels' as mh_batch_execute's, so terms equal it
// bit for bit.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "batch.h"
#include "batch_program.h"
#include "engine.hpp"
#include "gate_table.h"
#include "mh_internal.h"
#include "window.h"

namespace mh {
namespace {

// Pinned host memory, freed on destruction.
class Pinned {
public:
    Pinned() = default;
    explicit Pinned(std::size_t bytes) : bytes_(bytes) {
        MH_HIP(hipHostMalloc(&ptr_, std::max<std::size_t>(bytes, 16), hipHostMallocDefault));
    }
    ~Pinned() {
        if (ptr_) (void)hipHostFree(ptr_);
    }
    Pinned(const Pinned&) = delete;
    Pinned& operator=(const Pinned&) = delete;
    Pinned(Pinned&& o) noexcept : ptr_(o.ptr_), bytes_(o.bytes_) {
        o.ptr_ = nullptr;
        o.bytes_ = 0;
    }
    Pinned& operator=(Pinned&& o) noexcept {
        if (this != &o) {
            if (ptr_) (void)hipHostFree(ptr_);
            ptr_ = o.ptr_;
            bytes_ = o.bytes_;
            o.ptr_ = nullptr;
            o.bytes_ = 0;
        }
        return *this;
    }
    void* get() const { return ptr_; }
    std::size_t bytes() const { return bytes_; }

private:
    void* ptr_ = nullptr;
    std::size_t bytes_ = 0;
};

// A circuit's structure: what its fusion plan and pass programs depend on
// besides matrix values (opcodes, qubits, parameter counts).
bool same_structure(const mh_circuit& a, const mh_circuit& b) {
    if (a.num_qubits != b.num_qubits || a.num_ops != b.num_ops) return false;
    for (uint32_t i = 0; i < a.num_ops; ++i) {
        if (a.opcodes[i] != b.opcodes[i]) return false;
        if (a.qubit_offsets[i + 1] - a.qubit_offsets[i] != b.qubit_offsets[i + 1] - b.qubit_offsets[i]) return false;
        if (a.param_offsets[i + 1] - a.param_offsets[i] != b.param_offsets[i + 1] - b.param_offsets[i]) return false;
        for (uint32_t q = a.qubit_offsets[i]; q < a.qubit_offsets[i + 1]; ++q)
            if (a.qubits[q] != b.qubits[q - a.qubit_offsets[i] + b.qubit_offsets[i]]) return false;
    }
    return true;
}

// Deep copy of a circuit's structure arrays (parameters are not kept).
struct Structure {
    std::vector<int32_t> opcodes;
    std::vector<uint32_t> qubit_offsets, qubits, param_offsets;
    mh_circuit view{};

    void copy_from(const mh_circuit& c) {
        opcodes.assign(c.opcodes, c.opcodes + c.num_ops);
        qubit_offsets.assign(c.qubit_offsets, c.qubit_offsets + c.num_ops + 1);
        qubits.assign(c.qubits, c.qubits + qubit_offsets.back());
        param_offsets.assign(c.param_offsets, c.param_offsets + c.num_ops + 1);
        view = c;
        view.opcodes = opcodes.data();
        view.qubit_offsets = qubit_offsets.data();
        view.qubits = qubits.data();
        view.param_offsets = param_offsets.data();
        view.params = nullptr;
    }
};

}  // namespace
}  // namespace mh

struct mh_batch_graph {
    int32_t precision = MH_C128;
    int32_t fusion = MH_FUSE_CONTAINMENT;
    uint32_t fusion_width = 0;
    uint32_t max_pass_cost = 0;
    uint32_t threads = 1;
    mh::BatchSchedule schedule;
    uint32_t count = 0;
    uint32_t total_terms = 0;
    std::vector<mh::Structure> structure;
    std::vector<std::vector<mh::Block>> blocks;
    std::vector<mh::BatchItem> items;
    mh::BatchProgram program;
    std::unique_ptr<mh::EarlyReductions> reductions;
    std::vector<uint32_t> term_offsets;
    // Device residency.
    mh::DeviceBuffer arena, bundles, ops, terms, mats, args, groups, dterms;
    uint64_t arena_bytes = 0;
    mh::Pinned pinned_mats, pinned_terms;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint32_t launches = 0;
    double plan_seconds = 0, compile_seconds = 0;

    ~mh_batch_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
    }
};

namespace mh {
namespace {

using clock_type = std::chrono::steady_clock;

double seconds_since(clock_type::time_point t0) {
    return std::chrono::duration<double>(clock_type::now() - t0).count();
}

// Fusion plans and block matrices of every circuit (host threads).
void plan_all(mh_batch_graph& g, const mh_circuit* circuits) {
    const auto t0 = clock_type::now();
    parallel_for(g.count, g.threads, [&](std::size_t i) {
        validate_circuit(&circuits[i]);
        const CircuitView cv{&circuits[i]};
        g.blocks[i] = build_blocks(cv, plan_fusion(cv, g.fusion, g.fusion_width));
    });
    g.plan_seconds = seconds_since(t0);
}

// Each item's initial phase exp(i * global_phase), from host libm.
void set_phases(std::vector<BatchItem>& items, const double* global_phases) {
    for (std::size_t i = 0; i < items.size(); ++i) {
        const std::complex<double> ph = host_phase(global_phases ? global_phases[i] : 0.0);
        items[i].initialise = true;
        items[i].phase_re = ph.real();
        items[i].phase_im = ph.imag();
    }
}

// Everything on the main stream. A graph captured over several streams (the
// kernel streams forked off the main stream and joined back) kept the device
// memory its kernels used referenced after the graph and its executable were
// destroyed (measured on ROCm 7.2: hipFree returned success and the memory did
// not come back), so resident batches were never released; captured on one
// stream, the memory returns. One launch group at a time on one stream costs
// the default schedule nothing (it has one launch group), and the reduction
// buckets follow the launch that finishes their states.
void capture(mh_batch_graph& g) {
    context();
    const hipStream_t main = streams().main;
    BatchSchedule one = g.schedule;
    one.kernel_streams = 1;
    BatchDeviceTables dev;
    dev.bundles = g.bundles.as<uint64_t>();
    dev.ops = static_cast<const WOp*>(g.ops.get());
    dev.terms = static_cast<const SignTerm*>(g.terms.get());
    dev.mats = g.mats.get();
    dev.args = static_cast<const WBatchArgs*>(g.args.get());
    dev.groups = g.groups.as<uint32_t>();
    MH_HIP(hipStreamBeginCapture(main, hipStreamCaptureModeThreadLocal));
    try {
        MH_HIP(hipMemcpyAsync(g.mats.get(), g.pinned_mats.get(), g.program.mats().size(), hipMemcpyHostToDevice, main));
        AfterLaunch after;
        if (g.reductions) after = [&](const BatchLaunch& l, hipStream_t s) { g.reductions->after(l, s); };
        g.launches = issue_batch_program(g.program, g.items.data(), dev, one, 0, after);
        if (g.reductions) {
            g.reductions->flush({main});
            MH_HIP(hipMemcpyAsync(g.pinned_terms.get(), g.dterms.get(), g.dterms.bytes(), hipMemcpyDeviceToHost, main));
            g.launches += g.reductions->launches();
        }
    } catch (...) {
        hipGraph_t partial = nullptr;
        (void)hipStreamEndCapture(main, &partial);
        if (partial) (void)hipGraphDestroy(partial);
        throw;
    }
    MH_HIP(hipStreamEndCapture(main, &g.graph));
    MH_HIP(hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0));
}

}  // namespace
}  // namespace mh

namespace mh {
namespace {
// New parameters (circuits of the captured structure) into the staged matrix
// table; refuses a change of any pass program.
void update_blocks(mh_batch_graph& g, const mh_circuit* circuits, const double* global_phases);
}  // namespace
}  // namespace mh

using namespace mh;

extern "C" {

int mh_batch_graph_create(const mh_batch_request* req, mh_batch_graph** out) {
    return guarded([&] {
        if (req == nullptr || out == nullptr) fail(MH_ERR_INVALID, "null request or output");
        *out = nullptr;
        amp_bytes(req->precision);
        if (req->num_circuits == 0 || req->circuits == nullptr || req->term_offsets == nullptr)
            fail(MH_ERR_INVALID, "a captured batch needs circuits and term offsets");
        if (req->want_statevectors) fail(MH_ERR_INVALID, "a captured batch returns terms only");
        if (req->term_offsets[0] != 0) fail(MH_ERR_INVALID, "term_offsets must start at 0");
        auto g = std::make_unique<mh_batch_graph>();
        g->precision = req->precision;
        g->fusion = req->fusion;
        g->fusion_width = req->fusion_width;
        g->max_pass_cost = req->max_pass_cost;
        g->threads = req->host_threads ? req->host_threads : std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        g->schedule.group_bytes = req->group_bytes;
        if (req->kernel_streams) g->schedule.kernel_streams = req->kernel_streams;
        g->schedule.temporal = req->cached_states != 0;
        g->count = req->num_circuits;
        g->total_terms = req->term_offsets[g->count];
        if (g->total_terms > 0 && (req->x_masks == nullptr || req->z_masks == nullptr))
            fail(MH_ERR_INVALID, "terms without masks");
        const uint32_t width = window_width(req->precision);
        g->structure.resize(g->count);
        g->blocks.resize(g->count);
        g->items.resize(g->count);
        g->term_offsets.assign(req->term_offsets, req->term_offsets + g->count + 1);
        uint64_t arena_bytes = 0;
        for (uint32_t i = 0; i < g->count; ++i) {
            const mh_circuit& c = req->circuits[i];
            if (c.num_qubits < width || c.num_qubits > 30)
                fail(MH_ERR_INVALID, "a batched circuit needs " + std::to_string(width) + " to 30 qubits");
            if (req->term_offsets[i + 1] < req->term_offsets[i]) fail(MH_ERR_INVALID, "term_offsets must not decrease");
            g->structure[i].copy_from(c);
            arena_bytes += round_state_bytes(c.num_qubits, req->precision);
        }
        context();
        uint64_t available = 0;
        {
            size_t free_b = 0, total_b = 0;
            MH_HIP(hipMemGetInfo(&free_b, &total_b));
            available = static_cast<uint64_t>(free_b) + pool_reusable_bytes();
        }
        if (arena_bytes > available)
            fail(MH_ERR_RESOURCE, "a captured batch keeps every state resident: " + std::to_string(arena_bytes) +
                                      " bytes, " + std::to_string(available) + " available");
        plan_all(*g, req->circuits);
        g->arena_bytes = arena_bytes;
        g->arena = DeviceBuffer(arena_bytes);
        unsigned char* base = static_cast<unsigned char*>(g->arena.get());
        std::vector<ExpectationItem> exp_items(g->count);
        bool any_terms = false;
        uint64_t at = 0;
        for (uint32_t i = 0; i < g->count; ++i) {
            BatchItem& it = g->items[i];
            it.state = base + at;
            it.n = req->circuits[i].num_qubits;
            it.blocks = &g->blocks[i];
            it.max_cost = req->max_pass_cost;
            const uint32_t nt = req->term_offsets[i + 1] - req->term_offsets[i];
            if (nt > 0) {
                ExpectationItem& e = exp_items[i];
                e.state = it.state;
                e.n = it.n;
                e.terms = nt;
                e.x = req->x_masks + req->term_offsets[i];
                e.z = req->z_masks + req->term_offsets[i];
                e.out_index = req->term_offsets[i];
                any_terms = true;
            }
            at += round_state_bytes(it.n, req->precision);
        }
        set_phases(g->items, req->global_phases);
        const auto c0 = std::chrono::steady_clock::now();
        g->program = compile_batch_program(req->precision, g->items.data(), g->count, g->threads, g->schedule);
        g->compile_seconds = seconds_since(c0);
        if (g->program.has_generic() || g->program.needs_initialisation())
            fail(MH_ERR_INVALID, "a captured batch takes circuits with at least one gate and blocks of at most 2 "
                                 "qubits (dense) or diagonal blocks");
        auto table = [](DeviceBuffer& d, const std::vector<unsigned char>& src) {
            d = DeviceBuffer(src.size());
            upload(d.get(), src.data(), src.size());
        };
        table(g->bundles, g->program.bundles());
        table(g->ops, g->program.ops());
        table(g->terms, g->program.terms());
        table(g->args, g->program.args());
        table(g->groups, g->program.groups());
        g->mats = DeviceBuffer(g->program.mats().size());
        g->pinned_mats = Pinned(g->program.mats().size());
        std::memcpy(g->pinned_mats.get(), g->program.mats().data(), g->program.mats().size());
        if (any_terms) {
            g->dterms = DeviceBuffer(uint64_t(g->total_terms) * 2 * sizeof(double));
            g->reductions = std::make_unique<EarlyReductions>(req->precision, g->program, exp_items,
                                                              g->dterms.as<double>(), streams().main);
            g->reductions->upload(streams().copy, false);
            MH_HIP(hipStreamSynchronize(streams().copy));
            g->pinned_terms = Pinned(g->dterms.bytes());
        }
        synchronize();
        capture(*g);
        *out = g.release();
    });
}

int mh_batch_graph_update(mh_batch_graph* g, const mh_circuit* circuits, const double* global_phases) {
    return guarded([&] {
        if (g == nullptr || circuits == nullptr) fail(MH_ERR_INVALID, "null graph or circuits");
        for (uint32_t i = 0; i < g->count; ++i)
            if (!same_structure(circuits[i], g->structure[i].view))
                fail(MH_ERR_INVALID, "circuit " + std::to_string(i) +
                                         " differs in structure from the captured one (only parameters may change)");
        update_blocks(*g, circuits, global_phases);
    });
}

int mh_batch_graph_set_params(mh_batch_graph* g, const double* params, const double* global_phases) {
    return guarded([&] {
        if (g == nullptr) fail(MH_ERR_INVALID, "null graph");
        std::vector<mh_circuit> views(g->count);
        uint64_t at = 0;
        for (uint32_t i = 0; i < g->count; ++i) {
            views[i] = g->structure[i].view;
            const uint64_t np = views[i].param_offsets[views[i].num_ops];
            if (np > 0 && params == nullptr) fail(MH_ERR_INVALID, "null parameters");
            views[i].params = np > 0 ? params + at : nullptr;
            at += np;
        }
        update_blocks(*g, views.data(), global_phases);
    });
}

int mh_batch_graph_param_count(const mh_batch_graph* g, uint64_t* out) {
    return guarded([&] {
        if (g == nullptr || out == nullptr) fail(MH_ERR_INVALID, "null graph or output");
        uint64_t total = 0;
        for (const Structure& st : g->structure) total += st.view.param_offsets[st.view.num_ops];
        *out = total;
    });
}

}  // extern "C"

namespace mh {
namespace {

void update_blocks(mh_batch_graph& gr, const mh_circuit* circuits, const double* global_phases) {
    mh_batch_graph* g = &gr;
    {
        std::vector<std::vector<Block>> fresh(g->count);
        const auto t0 = clock_type::now();
        parallel_for(g->count, g->threads, [&](std::size_t i) {
            validate_circuit(&circuits[i]);
            const CircuitView cv{&circuits[i]};
            fresh[i] = build_blocks(cv, plan_fusion(cv, g->fusion, g->fusion_width));
        });
        g->plan_seconds = seconds_since(t0);
        std::vector<BatchItem> items = g->items;
        for (uint32_t i = 0; i < g->count; ++i) items[i].blocks = &fresh[i];
        set_phases(items, global_phases);
        const auto c0 = clock_type::now();
        BatchProgram prog = compile_batch_program(g->precision, items.data(), g->count, g->threads, g->schedule);
        g->compile_seconds = seconds_since(c0);
        if (!prog.same_shape(g->program))
            fail(MH_ERR_INVALID, "the new parameters change a pass program (a matrix entry became or stopped being "
                                 "exactly zero or +-1); create a new captured batch for them");
        std::memcpy(g->pinned_mats.get(), prog.mats().data(), prog.mats().size());
        g->blocks = std::move(fresh);
        g->items = std::move(items);
        for (uint32_t i = 0; i < g->count; ++i) g->items[i].blocks = &g->blocks[i];
        g->program = std::move(prog);
    }
}

}  // namespace
}  // namespace mh

extern "C" {

int mh_batch_graph_run(mh_batch_graph* g, mh_batch_result* res) {
    return guarded([&] {
        if (g == nullptr || res == nullptr) fail(MH_ERR_INVALID, "null graph or result");
        if (g->total_terms > 0 && res->expectation_terms == nullptr) fail(MH_ERR_INVALID, "no term output");
        const auto t0 = clock_type::now();
        context();
        MH_HIP(hipGraphLaunch(g->exec, streams().main));
        synchronize();
        if (g->reductions)
            std::memcpy(res->expectation_terms, g->pinned_terms.get(), uint64_t(g->total_terms) * 2 * sizeof(double));
        const double secs = seconds_since(t0);
        if (res->completion_seconds)
            for (uint32_t i = 0; i < g->count; ++i) res->completion_seconds[i] = secs;
        res->rounds = 1;
        res->launches = g->launches;
        res->plan_seconds = g->plan_seconds;
        res->compile_seconds = g->compile_seconds;
        res->device_seconds = secs;
        res->total_seconds = secs;
    });
}

int mh_batch_graph_destroy(mh_batch_graph* g) {
    return guarded([&] {
        if (g == nullptr) return;
        (void)hipStreamSynchronize(streams().main);
        delete g;
    });
}

}  // extern "C"
