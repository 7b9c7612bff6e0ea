// The serving loop (mh_server_*): the job server's device backend and its C ABI.
//
// The reference serves a stream of estimator and sampler requests from its
// Python primitives, one full `run` per observable or pub
// (python/menos/qiskit/estimator.py:91-146, sampler.py:93-129), each run
// planning (with a structure-keyed plan cache, src/plan/plan_cache.cpp:233-276)
// and executing alone. Here the host core (server_core.cpp) queues, admits and
// completes jobs on several lanes; this backend gives each lane its own
// streams, plans a job (the fusion plan from the cache, block matrices from
// the job's own parameters) and runs a lane's jobs of one precision as one
// execute_round in the lane's share of the server's arena: pass p of every
// circuit in one launch, all terms in two launches, each sampler's draw.
#include <algorithm>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "batch.h"
#include "engine.hpp"
#include "mh_internal.h"
#include "server_core.hpp"
#include "window.h"

namespace mh {
namespace {

constexpr uint32_t kMaxJobQubits = 30;

class HipServerBackend final : public serve::ServerBackend {
public:
    uint64_t capacity_bytes() override {
        size_t free_b = 0, total_b = 0;
        MH_HIP(hipMemGetInfo(&free_b, &total_b));
        return static_cast<uint64_t>(free_b) + pool_reusable_bytes();
    }

    void validate_circuit(const mh_circuit* c) override { mh::validate_circuit(c); }

    uint32_t min_qubits(int32_t precision) override { return window_width(precision); }
    uint32_t max_qubits() override { return kMaxJobQubits; }

    uint64_t state_bytes(uint32_t n, int32_t precision) override { return round_state_bytes(n, precision); }

    std::shared_ptr<void> lane_setup() override { return std::make_shared<OwnStreams>(); }

    void* segment_alloc(uint64_t bytes) override {
        DeviceBuffer b(bytes);
        void* p = b.get();
        segments_.push_back(std::move(b));
        return p;
    }

    void segment_free(void* segment) noexcept override {
        for (auto it = segments_.begin(); it != segments_.end(); ++it)
            if (it->get() == segment) {
                segments_.erase(it);
                return;
            }
    }

    void plan(serve::Job& j, int32_t fusion, uint32_t width, std::shared_ptr<const void>& plan) override {
        const CircuitView cv{&j.circuit.view};
        if (!plan) plan = std::make_shared<const FusionPlan>(plan_fusion(cv, fusion, width));
        j.work = std::make_shared<std::vector<Block>>(build_blocks(cv, *static_cast<const FusionPlan*>(plan.get())));
    }

    void execute(int32_t precision, const std::vector<serve::Job*>& jobs, void* arena, uint32_t threads,
                 serve::PhaseSeconds& seconds) override {
        std::vector<RoundItem> items(jobs.size());
        uint32_t terms = 0;
        for (std::size_t k = 0; k < jobs.size(); ++k) {
            serve::Job& j = *jobs[k];
            RoundItem& it = items[k];
            it.n = j.circuit.view.num_qubits;
            it.blocks = static_cast<const std::vector<Block>*>(j.work.get());
            it.global_phase = j.global_phase;
            it.num_terms = j.num_terms();
            it.x = j.x.data();
            it.z = j.z.data();
            it.term_out = terms;
            terms += j.num_terms();
            it.shots = j.shots;
            it.seed = j.seed;
            it.targets = &j.targets;
            it.counts = &j.counts;
        }
        std::vector<double> raw(2 * static_cast<std::size_t>(terms));
        RoundStats rs;
        execute_round(precision, items, threads, raw.data(), &rs, arena);
        for (std::size_t k = 0; k < jobs.size(); ++k) {
            serve::Job& j = *jobs[k];
            const uint32_t at = items[k].term_out;
            j.raw_terms.assign(raw.begin() + 2 * at, raw.begin() + 2 * (at + j.num_terms()));
            j.work.reset();
        }
        seconds.compile += rs.compile_seconds;
        seconds.pass += rs.pass_seconds;
        seconds.expectation += rs.expectation_seconds;
        seconds.sample += rs.sample_seconds;
    }

private:
    std::vector<DeviceBuffer> segments_;  // touched under the server's lock only
};

}  // namespace
}  // namespace mh

using namespace mh;

struct mh_server {
    std::unique_ptr<serve::Server> impl;
};

extern "C" {

int mh_server_create(const mh_server_options* options, mh_server** out) {
    return guarded([&] {
        if (out == nullptr) fail(MH_ERR_INVALID, "null output");
        mh_server_options o{};
        if (options) o = *options;
        if (options == nullptr) o.fusion = MH_FUSE_CONTAINMENT;
        if (o.fusion != MH_FUSE_NONE && o.fusion != MH_FUSE_CONTAINMENT && o.fusion != MH_FUSE_WIDTH &&
            o.fusion != MH_FUSE_PAIRS)
            fail(MH_ERR_INVALID, "unknown fusion mode");
        context();
        auto s = std::make_unique<mh_server>();
        s->impl = std::make_unique<serve::Server>(o, std::make_unique<HipServerBackend>());
        *out = s.release();
    });
}

int mh_server_submit(mh_server* server, const mh_job* job, uint64_t* ticket) {
    return guarded([&] {
        if (server == nullptr || ticket == nullptr) fail(MH_ERR_INVALID, "null server or ticket");
        *ticket = server->impl->submit(job);
    });
}

int mh_server_wait(mh_server* server, uint64_t ticket, mh_job_result* result) {
    return guarded([&] {
        if (server == nullptr || result == nullptr) fail(MH_ERR_INVALID, "null server or result");
        server->impl->wait(ticket, result);
    });
}

int mh_server_replay(mh_server* server, uint32_t count, const mh_job* jobs, const double* arrival_seconds,
                     mh_job_result* results) {
    return guarded([&] {
        if (server == nullptr || (count > 0 && (jobs == nullptr || arrival_seconds == nullptr || results == nullptr)))
            fail(MH_ERR_INVALID, "null server, jobs, arrivals or results");
        server->impl->replay(count, jobs, arrival_seconds, results);
    });
}

int mh_server_stats_get(mh_server* server, mh_server_stats* out) {
    return guarded([&] {
        if (server == nullptr || out == nullptr) fail(MH_ERR_INVALID, "null server or output");
        server->impl->stats(out);
    });
}

int mh_server_destroy(mh_server* server) {
    return guarded([&] { delete server; });
}

int mh_derive_seeds(uint64_t seed, uint32_t count, uint64_t* out) {
    return guarded([&] {
        if (count > 0 && out == nullptr) fail(MH_ERR_INVALID, "null output");
        uint64_t state = seed;
        for (uint32_t i = 0; i < count; ++i) {
            state += 0x9E3779B97F4A7C15ULL;
            uint64_t z = state;
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
            out[i] = z ^ (z >> 31);
        }
    });
}

}  // extern "C"
