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
//   3. per round, in groups of circuits (the host plans a group while the
//      device runs the one before): pass p of every circuit of the group in
//      one launch (BatchPassRunner; the first pass writes each state, global
//      phase included, so the arena is never cleared), then the group's
//      expectation terms in two launches behind its passes; one copy back
//      per group at the end.
//
// The passes and reductions are the single-circuit engine's kernels, so
// results equal mh_engine_execute's bit for bit.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <exception>
#include <mutex>
#include <thread>

#include "batch.h"
#include "engine.hpp"
#include "gate_table.h"
#include "mh_internal.h"
#include "window.h"

namespace mh {
namespace {


struct Planned {
    uint32_t n = 0;
    std::vector<Block> blocks;
};

constexpr uint64_t kAlign = 256;

}  // namespace

void parallel_for(std::size_t count, uint32_t threads, const std::function<void(std::size_t)>& body) {
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

uint64_t round_state_bytes(uint32_t n, int32_t precision) {
    return ((uint64_t{1} << n) * amp_bytes(precision) + kAlign - 1) / kAlign * kAlign;
}

EarlyReductions::EarlyReductions(int32_t precision, const BatchProgram& prog,
                                 const std::vector<ExpectationItem>& terms, double* out, hipStream_t reduce_stream)
    : out_(out), reduce_(reduce_stream) {
    static constexpr double kCuts[] = {0.5, 0.85, 1.0};
    static constexpr std::size_t kMinSplit = 48;  // states with terms below which a group keeps one bucket
    std::map<uint32_t, std::vector<std::size_t>> by_group;
    for (std::size_t i = 0; i < terms.size(); ++i)
        if (terms[i].terms > 0) by_group[prog.launch_group_of(i)].push_back(i);
    for (auto& [group, members] : by_group) {
        std::stable_sort(members.begin(), members.end(),
                         [&](std::size_t a, std::size_t b) { return prog.pass_count(a) < prog.pass_count(b); });
        double total = 0;
        for (std::size_t i : members) total += std::ldexp(double(terms[i].terms), static_cast<int>(terms[i].n));
        const std::size_t cuts = members.size() < kMinSplit ? 1 : sizeof(kCuts) / sizeof(kCuts[0]);
        double done = 0;
        std::size_t at = 0;
        for (std::size_t c = 0; c < cuts && at < members.size(); ++c) {
            const double goal = cuts == 1 ? total : kCuts[c] * total;
            std::vector<ExpectationItem> bucket;
            // Whole pass counts at a time: a bucket ends where a pass count ends.
            while (at < members.size() && (bucket.empty() || done < goal || c + 1 == cuts)) {
                const uint32_t p = prog.pass_count(members[at]);
                while (at < members.size() && prog.pass_count(members[at]) == p) {
                    const ExpectationItem& e = terms[members[at]];
                    done += std::ldexp(double(e.terms), static_cast<int>(e.n));
                    bucket.push_back(e);
                    ++at;
                }
            }
            Bucket b;
            b.group = group;
            const uint32_t last = prog.pass_count(members[at - 1]);
            b.pass = last > 0 ? last - 1 : 0;
            b.queued = false;
            b.program = std::make_unique<ExpectationProgram>(precision, bucket);
            b.workspace = DeviceBuffer(b.program->workspace_bytes());
            // A bucket of states without passes is queued by flush().
            if (last == 0) b.pass = UINT32_MAX;
            buckets_.push_back(std::move(b));
        }
    }
}

EarlyReductions::~EarlyReductions() {
    for (hipEvent_t e : events_) (void)hipEventDestroy(e);
    if (uploaded_) (void)hipEventDestroy(uploaded_);
}

hipEvent_t EarlyReductions::event() {
    hipEvent_t e = nullptr;
    MH_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    events_.push_back(e);
    return e;
}

void EarlyReductions::upload(hipStream_t copy, bool ordered) {
    for (Bucket& b : buckets_) b.program->upload_on(b.workspace.get(), copy);
    wait_upload_ = ordered;
    if (ordered) {
        MH_HIP(hipEventCreateWithFlags(&uploaded_, hipEventDisableTiming));
        MH_HIP(hipEventRecord(uploaded_, copy));
    }
}

void EarlyReductions::queue(Bucket& b) {
    if (wait_upload_) {
        MH_HIP(hipStreamWaitEvent(reduce_, uploaded_, 0));
        wait_upload_ = false;
    }
    b.program->launch_on(b.workspace.get(), reduce_, out_);
    b.queued = true;
    launches_ += 2;
}

void EarlyReductions::after(const BatchLaunch& l, hipStream_t stream) {
    hipEvent_t done = nullptr;
    for (Bucket& b : buckets_) {
        if (b.queued || b.group != l.group || b.pass != l.pass) continue;
        // On the launch's own stream the order is already the stream's.
        if (done == nullptr && stream != reduce_) {
            done = event();
            MH_HIP(hipEventRecord(done, stream));
            MH_HIP(hipStreamWaitEvent(reduce_, done, 0));
        }
        queue(b);
    }
}

void EarlyReductions::flush(const std::vector<hipStream_t>& streams) {
    bool waited = false;
    for (Bucket& b : buckets_) {
        if (b.queued) continue;
        if (!waited) {
            for (hipStream_t st : streams) {
                if (st == reduce_) continue;
                hipEvent_t e = event();
                MH_HIP(hipEventRecord(e, st));
                MH_HIP(hipStreamWaitEvent(reduce_, e, 0));
            }
            waited = true;
        }
        queue(b);
    }
}

void execute_round(int32_t precision, std::vector<RoundItem>& items, uint32_t threads, double* terms_out,
                   RoundStats* stats, void* arena_in, const std::function<void(std::size_t, std::size_t)>& prepare,
                   const BatchSchedule& schedule) {
    const std::size_t amp = amp_bytes(precision);
    const uint32_t width = window_width(precision);
    if (items.empty()) return;
    uint64_t bytes = 0;
    for (const RoundItem& it : items) {
        if (it.n < width || it.n > 30)
            fail(MH_ERR_INVALID, "a batched circuit needs " + std::to_string(width) + " to 30 qubits");
        if (it.num_terms > 0 && terms_out == nullptr) fail(MH_ERR_INVALID, "terms without an output");
        bytes += round_state_bytes(it.n, precision);
    }
    context();
    DeviceBuffer own(arena_in ? 0 : bytes);
    unsigned char* const arena = static_cast<unsigned char*>(arena_in ? arena_in : own.get());
    const std::size_t count = items.size();
    std::vector<BatchItem> passes(count);
    uint64_t at = 0;
    for (std::size_t k = 0; k < count; ++k) {
        const RoundItem& it = items[k];
        passes[k].state = arena + at;
        passes[k].n = it.n;
        passes[k].max_cost = it.max_pass_cost;
        const std::complex<double> ph = host_phase(it.global_phase);
        passes[k].initialise = true;
        passes[k].phase_re = ph.real();
        passes[k].phase_im = ph.imag();
        at += round_state_bytes(it.n, precision);
    }
    // On an error past this point, wait for the queued work before the
    // buffers it uses go back to the pool.
    struct WaitOnUnwind {
        hipStream_t stream;
        int pending = std::uncaught_exceptions();
        ~WaitOnUnwind() {
            if (std::uncaught_exceptions() > pending) (void)hipStreamSynchronize(stream);
        }
    } wait_on_unwind{streams().main};
    // No initialisation pass: each state's first window pass generates
    // phase * |0...0> instead of loading it (window_body), and the rare state
    // whose first pass is not a window pass gets an initialisation kernel.
    // Apply, reduce, sample, copy back.
    using clock = std::chrono::steady_clock;
    const auto p0 = clock::now();
    BatchStats bs;
    double prepare_seconds = 0;
    double reduce_seconds = 0;
    uint32_t launches = 0;
    // The round's terms on the device, from its lowest term index; each
    // planning group's reductions write their own terms there.
    uint32_t t_lo = UINT32_MAX, t_hi = 0;
    for (const RoundItem& it : items)
        if (it.num_terms > 0) {
            t_lo = std::min(t_lo, it.term_out);
            t_hi = std::max(t_hi, it.term_out + it.num_terms);
        }
    DeviceBuffer dterms(t_lo < t_hi ? uint64_t(t_hi - t_lo) * 2 * sizeof(double) : 0);
    std::vector<std::unique_ptr<EarlyReductions>> reductions;
    {
        // Groups of items: the host plans and compiles group g + 1 while the
        // device runs group g. Group sizes grow (1, 1, 2, 4, 8 sixteenths) so
        // that only a sixteenth of the host work is exposed before the device
        // starts, and every later group's host work is shorter than the device
        // work queued before it. A group's terms are reduced in buckets as its
        // states become final (EarlyReductions), on a stream of their own.
        BatchPassRunner runner(precision, threads, schedule);
        const hipStream_t reduce_stream = streams().reduce;
        std::vector<hipStream_t> all_streams{streams().main};
        for (uint32_t k = 0; k < Streams::kKernelStreams; ++k) all_streams.push_back(streams().kernel(k));
        static constexpr std::size_t kEnds[] = {1, 2, 4, 8, 16};
        const std::size_t groups = count >= 64 ? 5 : 1;
        for (std::size_t g = 0; g < groups; ++g) {
            const std::size_t lo = groups == 1 ? 0 : count * (g ? kEnds[g - 1] : 0) / 16;
            const std::size_t hi = groups == 1 ? count : count * kEnds[g] / 16;
            if (prepare) {
                const auto q0 = clock::now();
                prepare(lo, hi);
                prepare_seconds += std::chrono::duration<double>(clock::now() - q0).count();
            }
            for (std::size_t k = lo; k < hi; ++k) {
                if (items[k].blocks == nullptr) fail(MH_ERR_INTERNAL, "a round item without blocks");
                passes[k].blocks = items[k].blocks;
            }
            std::vector<ExpectationItem> group_terms(hi - lo);
            bool any_terms = false;
            for (std::size_t k = lo; k < hi; ++k) {
                const RoundItem& it = items[k];
                if (it.num_terms == 0) continue;
                ExpectationItem& e = group_terms[k - lo];
                e.state = passes[k].state;
                e.n = it.n;
                e.terms = it.num_terms;
                e.x = it.x;
                e.z = it.z;
                e.out_index = it.term_out - t_lo;
                any_terms = true;
            }
            EarlyReductions* red = nullptr;
            BatchHooks hooks;
            if (any_terms) {
                hooks.compiled = [&](const BatchProgram& prog) {
                    const auto r0 = clock::now();
                    reductions.push_back(std::make_unique<EarlyReductions>(precision, prog, group_terms,
                                                                           dterms.as<double>(), reduce_stream));
                    red = reductions.back().get();
                    red->upload(streams().copy, true);
                    reduce_seconds += std::chrono::duration<double>(clock::now() - r0).count();
                };
                hooks.after = [&](const BatchLaunch& l, hipStream_t stream) { red->after(l, stream); };
            }
            runner.add(passes.data() + lo, hi - lo, &bs, any_terms ? &hooks : nullptr);
            if (red) {
                red->flush(all_streams);
                launches += red->launches();
            }
        }
        runner.finish();
    }
    const auto p1 = clock::now();
    launches += bs.launches;
    if (t_lo < t_hi) download(terms_out + 2 * uint64_t(t_lo), dterms.get(), dterms.bytes());
    const auto p2 = clock::now();
    for (std::size_t k = 0; k < count; ++k) {
        const RoundItem& it = items[k];
        if (it.shots > 0) {
            static const std::vector<std::pair<uint32_t, uint32_t>> kNone;
            *it.counts = sample_counts(passes[k].state, it.n, precision, it.shots, it.seed,
                                       it.targets ? *it.targets : kNone);
            launches += 7;
        }
        if (it.statevector) download(it.statevector, passes[k].state, (uint64_t{1} << it.n) * amp);
    }
    synchronize();
    if (stats) {
        const auto p3 = clock::now();
        stats->launches += launches;
        stats->compile_seconds += bs.compile_seconds;
        stats->prepare_seconds += prepare_seconds;
        stats->pass_seconds +=
            std::chrono::duration<double>(p1 - p0).count() - bs.compile_seconds - prepare_seconds - reduce_seconds;
        stats->expectation_seconds += std::chrono::duration<double>(p2 - p1).count() + reduce_seconds;
        stats->sample_seconds += std::chrono::duration<double>(p3 - p2).count();
    }
}

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

        // 1. Checks that need no plan; the plans are made round by round, a
        // group of circuits at a time, while the device runs the group before.
        for (uint32_t i = 0; i < count; ++i) {
            const mh_circuit* c = &req->circuits[i];
            if (c->num_qubits < width || c->num_qubits > 30)
                fail(MH_ERR_INVALID, "a batched circuit needs " + std::to_string(width) + " to 30 qubits");
            if (req->term_offsets[i + 1] < req->term_offsets[i]) fail(MH_ERR_INVALID, "term_offsets must not decrease");
        }
        std::vector<Planned> plans(count);
        for (uint32_t i = 0; i < count; ++i) plans[i].n = req->circuits[i].num_qubits;
        double plan_seconds = 0;

        // 2. Rounds of states that fit round_bytes; 3. each round executed.
        uint64_t round_bytes = req->round_bytes;
        if (round_bytes == 0) {
            size_t free_b = 0, total_b = 0;
            MH_HIP(hipMemGetInfo(&free_b, &total_b));
            round_bytes = (static_cast<uint64_t>(free_b) + pool_reusable_bytes()) / 2;
        }
        uint64_t sv_at = 0;
        uint32_t rounds = 0;
        RoundStats rs;
        double device_seconds = 0;
        for (uint32_t first = 0; first < count;) {
            uint32_t last = first;
            uint64_t bytes = 0;
            while (last < count) {
                const uint64_t b = round_state_bytes(plans[last].n, req->precision);
                if (last > first && bytes + b > round_bytes) break;
                bytes += b;
                ++last;
            }
            const auto r0 = clock::now();
            std::vector<RoundItem> items(last - first);
            for (uint32_t i = first; i < last; ++i) {
                RoundItem& it = items[i - first];
                it.n = plans[i].n;
                it.max_pass_cost = req->max_pass_cost;
                it.global_phase = req->global_phases ? req->global_phases[i] : 0.0;
                it.num_terms = req->term_offsets[i + 1] - req->term_offsets[i];
                it.x = req->x_masks + req->term_offsets[i];
                it.z = req->z_masks + req->term_offsets[i];
                it.term_out = req->term_offsets[i];
                if (req->want_statevectors) {
                    it.statevector = static_cast<unsigned char*>(res->statevectors) + sv_at;
                    sv_at += (uint64_t{1} << plans[i].n) * amp;
                }
            }
            const uint32_t round_first = first;
            auto plan_group = [&](std::size_t lo, std::size_t hi) {
                const auto q0 = clock::now();
                parallel_for(hi - lo, threads, [&](std::size_t k) {
                    const uint32_t i = round_first + static_cast<uint32_t>(lo + k);
                    const mh_circuit* c = &req->circuits[i];
                    validate_circuit(c);
                    const CircuitView cv{c};
                    plans[i].blocks = build_blocks(cv, plan_fusion(cv, req->fusion, req->fusion_width));
                    items[lo + k].blocks = &plans[i].blocks;
                });
                plan_seconds += std::chrono::duration<double>(clock::now() - q0).count();
            };
            BatchSchedule schedule;
            schedule.group_bytes = req->group_bytes;
            if (req->kernel_streams != 0) schedule.kernel_streams = req->kernel_streams;
            schedule.temporal = req->cached_states != 0;
            execute_round(req->precision, items, threads, res->expectation_terms, &rs, nullptr, plan_group, schedule);
            device_seconds += std::chrono::duration<double>(clock::now() - r0).count();
            const double done = std::chrono::duration<double>(clock::now() - t0).count();
            if (res->completion_seconds)
                for (uint32_t i = first; i < last; ++i) res->completion_seconds[i] = done;
            ++rounds;
            first = last;
        }
        const uint32_t launches = rs.launches;
        res->rounds = rounds;
        res->launches = launches;
        res->plan_seconds = plan_seconds;
        res->device_seconds = device_seconds - plan_seconds;
        res->compile_seconds = rs.compile_seconds;
        res->total_seconds = std::chrono::duration<double>(clock::now() - t0).count();
    });
}
