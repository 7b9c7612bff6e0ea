// Device context, error state, the memory pool and host<->device copies.
//
// The pool replaces the reference's DevicePool (src/kernels/cuda/device_pool.cpp:24-141),
// which keeps freed blocks in a cudaMallocAsync pool whose release threshold is
// the whole device. Here freed blocks are kept in a size-keyed free list and
// handed back on an exact-size match, so a job that allocates the same 16 GiB
// or 128 GiB state again reuses it without a hipMalloc; trim() returns them.
// Allocation failure after a trim is ResourceExceeded naming the bytes, the
// free and the retained figures (reference device_pool.cpp:73-86).
#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <unordered_map>

#include "mh_internal.h"

namespace mh {

namespace {
thread_local std::string g_last_error;
}

void set_last_error(const std::string& message) { g_last_error = message; }

void check_hip(hipError_t e, const char* what) {
    if (e != hipSuccess) {
        (void)hipGetLastError();
        fail(MH_ERR_INTERNAL, std::string(what) + ": " + hipGetErrorString(e));
    }
}

namespace {

// The stream set installed by an OwnStreams on this thread, if any.
thread_local Streams* t_streams = nullptr;

void create_streams(Streams& s) {
    MH_HIP(hipStreamCreateWithFlags(&s.main, hipStreamNonBlocking));
    MH_HIP(hipStreamCreateWithFlags(&s.second, hipStreamNonBlocking));
    MH_HIP(hipStreamCreateWithFlags(&s.third, hipStreamNonBlocking));
    MH_HIP(hipStreamCreateWithFlags(&s.fourth, hipStreamNonBlocking));
    MH_HIP(hipStreamCreateWithFlags(&s.copy, hipStreamNonBlocking));
    MH_HIP(hipStreamCreateWithFlags(&s.reduce, hipStreamNonBlocking));
}

}  // namespace

Context& context() {
    static Context* ctx = [] {
        int count = 0;
        if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
            (void)hipGetLastError();
            fail(MH_ERR_DEVICE, "no HIP device is visible");
        }
        auto* c = new Context();
        MH_HIP(hipSetDevice(c->device));
        MH_HIP(hipGetDeviceProperties(&c->props, c->device));
        if (std::strncmp(c->props.gcnArchName, "gfx950", 6) != 0) {
            fail(MH_ERR_DEVICE, std::string("device is ") + c->props.gcnArchName + ", this library is built for gfx950");
        }
        create_streams(c->defaults);
        return c;
    }();
    MH_HIP(hipSetDevice(ctx->device));
    return *ctx;
}

Streams& streams() {
    Context& ctx = context();
    return t_streams ? *t_streams : ctx.defaults;
}

OwnStreams::OwnStreams() {
    context();
    create_streams(own_);
    previous_ = t_streams;
    t_streams = &own_;
}

OwnStreams::~OwnStreams() {
    t_streams = previous_;
    for (hipStream_t s : {own_.reduce, own_.copy, own_.fourth, own_.third, own_.second, own_.main}) {
        (void)hipStreamSynchronize(s);
        (void)hipStreamDestroy(s);
    }
}

void synchronize() { MH_HIP(hipStreamSynchronize(streams().main)); }

std::size_t amp_bytes(int32_t precision) {
    if (precision == MH_C128) return 16;
    if (precision == MH_C64) return 8;
    fail(MH_ERR_INVALID, "precision must be MH_C128 or MH_C64");
}

// ---- pool

namespace {

struct Pool {
    std::mutex mu;
    std::multimap<uint64_t, void*> free_blocks;     // bytes -> block
    std::unordered_map<void*, uint64_t> live;        // block -> bytes
    uint64_t free_total = 0;
    uint64_t live_total = 0;
};

Pool& pool() {
    static Pool* p = new Pool();
    return *p;
}

// Blocks are rounded up to 2 MiB so states of one size class reuse each other.
uint64_t rounded(uint64_t bytes) {
    constexpr uint64_t kGrain = uint64_t{2} << 20;
    return (bytes + kGrain - 1) / kGrain * kGrain;
}

void trim_locked(Pool& p) {
    for (auto& [bytes, ptr] : p.free_blocks) {
        (void)bytes;
        (void)hipFree(ptr);
    }
    p.free_blocks.clear();
    p.free_total = 0;
}

}  // namespace

void* pool_alloc(uint64_t bytes) {
    Context& ctx = context();
    Pool& p = pool();
    const uint64_t size = rounded(std::max<uint64_t>(bytes, 1));
    std::lock_guard<std::mutex> lock(p.mu);
    auto it = p.free_blocks.find(size);
    if (it != p.free_blocks.end()) {
        void* ptr = it->second;
        p.free_blocks.erase(it);
        p.free_total -= size;
        p.live[ptr] = size;
        p.live_total += size;
        return ptr;
    }
    void* ptr = nullptr;
    hipError_t e = hipMalloc(&ptr, size);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        trim_locked(p);
        e = hipMalloc(&ptr, size);
    }
    if (e != hipSuccess) {
        (void)hipGetLastError();
        size_t free_b = 0, total_b = 0;
        (void)hipMemGetInfo(&free_b, &total_b);
        (void)hipGetLastError();
        fail(MH_ERR_RESOURCE, "device allocation of " + std::to_string(bytes) + " bytes failed (free " +
                                  std::to_string(free_b) + " of " + std::to_string(total_b) + " bytes, " +
                                  std::to_string(p.live_total) + " live in the pool)");
    }
    (void)ctx;
    p.live[ptr] = size;
    p.live_total += size;
    return ptr;
}

void pool_free(void* ptr) noexcept {
    if (ptr == nullptr) return;
    Pool& p = pool();
    std::lock_guard<std::mutex> lock(p.mu);
    auto it = p.live.find(ptr);
    if (it == p.live.end()) return;
    const uint64_t size = it->second;
    p.live.erase(it);
    p.live_total -= size;
    p.free_blocks.emplace(size, ptr);
    p.free_total += size;
}

uint64_t pool_reusable_bytes() {
    Pool& p = pool();
    std::lock_guard<std::mutex> lock(p.mu);
    return p.free_total;
}

uint64_t pool_live_bytes() {
    Pool& p = pool();
    std::lock_guard<std::mutex> lock(p.mu);
    return p.live_total;
}

void pool_trim() {
    context();
    Pool& p = pool();
    std::lock_guard<std::mutex> lock(p.mu);
    trim_locked(p);
}

// ---- copies

void upload(void* device_dst, const void* host_src, std::size_t bytes) {
    if (bytes == 0) return;
    context();
    MH_HIP(hipMemcpyAsync(device_dst, host_src, bytes, hipMemcpyHostToDevice, streams().main));
}

void download(void* host_dst, const void* device_src, std::size_t bytes) {
    if (bytes == 0) return;
    context();
    MH_HIP(hipMemcpyAsync(host_dst, device_src, bytes, hipMemcpyDeviceToHost, streams().main));
    MH_HIP(hipStreamSynchronize(streams().main));
}

}  // namespace mh

using namespace mh;

extern "C" {

const char* mh_last_error(void) { return g_last_error.c_str(); }

int mh_device_query(mh_device_info* out) {
    return guarded([&] {
        if (out == nullptr) fail(MH_ERR_INVALID, "mh_device_query: null output");
        Context& ctx = context();
        std::memset(out, 0, sizeof(*out));
        std::strncpy(out->name, ctx.props.name, sizeof(out->name) - 1);
        std::strncpy(out->arch, ctx.props.gcnArchName, sizeof(out->arch) - 1);
        out->compute_units = ctx.props.multiProcessorCount;
        out->warp_size = ctx.props.warpSize;
        out->max_threads_per_block = ctx.props.maxThreadsPerBlock;
        out->lds_bytes_per_block = static_cast<int32_t>(ctx.props.sharedMemPerBlock);
        out->clock_khz = ctx.props.clockRate;
        out->memory_clock_khz = ctx.props.memoryClockRate;
        out->memory_bus_width = ctx.props.memoryBusWidth;
        size_t free_b = 0, total_b = 0;
        MH_HIP(hipMemGetInfo(&free_b, &total_b));
        out->total_bytes = total_b;
        out->free_bytes = free_b;
    });
}

int mh_device_available_bytes(uint64_t* out) {
    return guarded([&] {
        context();
        size_t free_b = 0, total_b = 0;
        MH_HIP(hipMemGetInfo(&free_b, &total_b));
        *out = free_b + pool_reusable_bytes();
    });
}

int mh_pool_alloc(uint64_t bytes, void** out) {
    return guarded([&] {
        if (out == nullptr) fail(MH_ERR_INVALID, "mh_pool_alloc: null output");
        *out = pool_alloc(bytes);
    });
}

int mh_pool_free(void* ptr) {
    return guarded([&] { pool_free(ptr); });
}

int mh_pool_reusable_bytes(uint64_t* out) {
    return guarded([&] { *out = pool_reusable_bytes(); });
}

int mh_pool_trim(void) {
    return guarded([&] { pool_trim(); });
}

int mh_copy_to_host(const void* state, uint32_t num_qubits, int32_t precision, void* host) {
    return guarded([&] {
        if (num_qubits > 40) fail(MH_ERR_INVALID, "mh_copy_to_host: num_qubits above 40");
        download(host, state, (std::size_t{1} << num_qubits) * amp_bytes(precision));
    });
}

int mh_copy_from_host(void* state, uint32_t num_qubits, int32_t precision, const void* host) {
    return guarded([&] {
        if (num_qubits > 40) fail(MH_ERR_INVALID, "mh_copy_from_host: num_qubits above 40");
        upload(state, host, (std::size_t{1} << num_qubits) * amp_bytes(precision));
        synchronize();
    });
}

int mh_state_digest(const void* state, uint32_t num_qubits, int32_t precision, uint64_t* out) {
    return guarded([&] {
        if (num_qubits > 40) fail(MH_ERR_INVALID, "mh_state_digest: num_qubits above 40");
        const std::size_t total = (std::size_t{1} << num_qubits) * amp_bytes(precision);
        const std::size_t chunk = std::min<std::size_t>(total, std::size_t{256} << 20);
        void* pinned = nullptr;
        MH_HIP(hipHostMalloc(&pinned, chunk, hipHostMallocDefault));
        uint64_t h = 0xcbf29ce484222325ULL;
        try {
            for (std::size_t at = 0; at < total; at += chunk) {
                const std::size_t len = std::min(chunk, total - at);
                download(pinned, static_cast<const char*>(state) + at, len);
                const auto* b = static_cast<const unsigned char*>(pinned);
                for (std::size_t i = 0; i < len; ++i) {
                    h ^= b[i];
                    h *= 0x100000001b3ULL;
                }
            }
        } catch (...) {
            (void)hipHostFree(pinned);
            throw;
        }
        MH_HIP(hipHostFree(pinned));
        *out = h;
    });
}

}  // extern "C"
