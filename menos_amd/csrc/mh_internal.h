// Internal declarations shared by the engine's translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <complex>
#include <functional>
#include <memory>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "menos_hip.h"
#include "mh_error.h"

namespace mh {

using cplx = std::complex<double>;

void check_hip(hipError_t e, const char* what);
#define MH_HIP(call) ::mh::check_hip((call), #call)

void set_last_error(const std::string& message);

// Runs f, mapping exceptions to statuses (the C boundary of every entry point).
template <typename F>
int guarded(F&& f) noexcept {
    try {
        f();
        return MH_OK;
    } catch (const Error& e) {
        set_last_error(e.what());
        return e.status;
    } catch (const std::bad_alloc&) {
        set_last_error("host allocation failed");
        return MH_ERR_RESOURCE;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return MH_ERR_INTERNAL;
    }
}

// A set of streams of default priority: every kernel and copy the library
// issues goes to the calling thread's set.
struct Streams {
    static constexpr uint32_t kKernelStreams = 4;
    hipStream_t main = nullptr;    // kernels
    hipStream_t second = nullptr;  // more kernel streams: batched groups rotate over main, second, third, fourth
    hipStream_t third = nullptr;
    hipStream_t fourth = nullptr;
    hipStream_t copy = nullptr;    // table uploads that run ahead of the kernels
    hipStream_t reduce = nullptr;  // batched term reductions, beside the kernel streams

    hipStream_t kernel(uint32_t i) const {
        const hipStream_t k[kKernelStreams] = {main, second, third, fourth};
        return k[i % kKernelStreams];
    }
};

// The library's device context on device 0, created on first use, with the
// stream set of every thread that has none of its own.
struct Context {
    int device = 0;
    Streams defaults;
    hipDeviceProp_t props{};
};
Context& context();

// The calling thread's streams: those an OwnStreams on this thread installed,
// else the context's.
Streams& streams();

// Gives the calling thread a stream set of its own for the scope's lifetime,
// so that its work overlaps on the device with other threads' (the job
// server's lanes). The destructor waits for the set's work, then destroys it.
class OwnStreams {
public:
    OwnStreams();
    ~OwnStreams();
    OwnStreams(const OwnStreams&) = delete;
    OwnStreams& operator=(const OwnStreams&) = delete;

private:
    Streams own_;
    Streams* previous_ = nullptr;
};

// Waits for the calling thread's main stream.
void synchronize();

std::size_t amp_bytes(int32_t precision);

// ---- memory pool (device.hip)
void* pool_alloc(uint64_t bytes);
void pool_free(void* p) noexcept;
uint64_t pool_reusable_bytes();
void pool_trim();
uint64_t pool_live_bytes();

// A device buffer from the pool, released on destruction.
class DeviceBuffer {
public:
    DeviceBuffer() = default;
    explicit DeviceBuffer(uint64_t bytes) : ptr_(bytes ? pool_alloc(bytes) : nullptr), bytes_(bytes) {}
    ~DeviceBuffer() { reset(); }
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    DeviceBuffer(DeviceBuffer&& o) noexcept : ptr_(o.ptr_), bytes_(o.bytes_) { o.ptr_ = nullptr; o.bytes_ = 0; }
    DeviceBuffer& operator=(DeviceBuffer&& o) noexcept {
        if (this != &o) { reset(); ptr_ = o.ptr_; bytes_ = o.bytes_; o.ptr_ = nullptr; o.bytes_ = 0; }
        return *this;
    }
    void reset() noexcept { if (ptr_) pool_free(ptr_); ptr_ = nullptr; bytes_ = 0; }
    void* get() const { return ptr_; }
    uint64_t bytes() const { return bytes_; }
    template <typename T> T* as() const { return static_cast<T*>(ptr_); }
private:
    void* ptr_ = nullptr;
    uint64_t bytes_ = 0;
};

// Pinned host staging for small uploads (matrices, pass programs).
void upload(void* device_dst, const void* host_src, std::size_t bytes);
void download(void* host_dst, const void* device_src, std::size_t bytes);

// ---- kernels (gates.hip)
void launch_init(void* state, uint32_t n, int32_t precision, double re, double im);
// The same on a given stream.
void launch_init_on(void* state, uint32_t n, int32_t precision, double re, double im, hipStream_t stream);
// One gate, dispatched (diagonal / k=1 / k=2 / generic) or generic only.
void apply_gate(void* state, uint32_t n, int32_t precision, const cplx* matrix, const uint32_t* operands,
                uint32_t k, bool generic_only);

// A block ready to run: matrix (row-major, 4^k entries) and operands in the
// matrix's operand order.
struct Block {
    std::vector<uint32_t> operands;
    std::vector<cplx> matrix;
    bool diagonal = false;
};

struct PassStats {
    uint32_t passes = 0;
    std::vector<float> pass_ms;  // filled when timing is requested
    // Called after each completed pass with the passes completed so far;
    // true abandons (the queued work is waited for, then MH_ERR_ABANDONED).
    std::function<bool(uint64_t)> monitor;
};

// Applies the blocks in order, grouped into window passes.
// max_cost caps the per-amplitude FP work of one pass (0: the default).
void apply_blocks(void* state, uint32_t n, int32_t precision, const std::vector<Block>& blocks,
                  uint32_t window_qubits, bool time_passes, PassStats* stats, uint32_t max_cost);

// The default window: one wavefront's registers (10 qubits).
uint32_t window_default(int32_t precision) noexcept;

// Number of HBM passes the schedule makes for these blocks.
uint32_t count_passes(uint32_t n, int32_t precision, const std::vector<Block>& blocks, uint32_t window_qubits,
                      uint32_t max_cost);

bool matrix_is_diagonal(const cplx* m, std::size_t dim);

// ---- expectation (expectation.hip)
void expectation_terms(const void* state, uint32_t n, int32_t precision, const uint64_t* x, const uint64_t* z,
                       uint32_t terms, cplx* out);

// The terms of many states in two launches; results land at out[2 * (out_index + t)].
struct ExpectationItem {
    const void* state = nullptr;
    uint32_t n = 0;
    uint32_t terms = 0;
    const uint64_t* x = nullptr;
    const uint64_t* z = nullptr;
    uint32_t out_index = 0;
};
void expectation_batch(int32_t precision, const std::vector<ExpectationItem>& items, double* out);

// The same reduction split for a caller that launches it on streams of its
// own or keeps it resident (a captured graph): planned once on the host, its
// descriptors uploaded into a workspace of workspace_bytes(), then launched any
// number of times. States must stay where they were when it was planned.
class ExpectationProgram {
public:
    ExpectationProgram(int32_t precision, const std::vector<ExpectationItem>& items);
    ~ExpectationProgram();
    ExpectationProgram(const ExpectationProgram&) = delete;
    ExpectationProgram& operator=(const ExpectationProgram&) = delete;
    uint64_t workspace_bytes() const;
    // Copies the descriptors into the workspace on `stream`, not waited for
    // (the program outlives the copy).
    void upload_on(void* workspace, hipStream_t stream) const;
    // Launches the reduction kernels on `stream` (no copies, no allocation, no
    // waits: capturable); term t of item i lands at out[2 * (out_index + t)]
    // (device memory), or in the workspace's output region when out is null.
    void launch_on(void* workspace, hipStream_t stream, double* out = nullptr) const;

private:
    struct Impl;
    std::unique_ptr<Impl> impl_;
};

// ---- sampler (sampler.hip)
void sample_sorted(const void* state, uint32_t n, int32_t precision, const double* sorted_u, uint64_t shots,
                   uint64_t* basis_out);
// Keeps `state` (2^n amplitudes) on the device with its sampler tables
// (mh_region_sampler_*); synchronises.
mh_region_sampler* make_region_sampler(DeviceBuffer&& state, uint32_t n, int32_t precision);

// sample_sorted over a state of 2^n amplitudes that reaches the device in
// consecutive chunks of 2^chunk_log (index order), with the same indices bit
// for bit: each chunk's tables start from the exact running value the chunk
// before it ended on, and a uniform is drawn in the chunk whose range of
// running values holds it (one at or above the total takes the last index).
// Everything is queued on `stream`; a chunk's memory may be reused once the
// work queued by its add() is done.
class ChunkedSampler {
public:
    ChunkedSampler(uint32_t n, uint32_t chunk_log, int32_t precision, const double* sorted_u, uint64_t shots,
                   hipStream_t stream);
    ~ChunkedSampler();
    ChunkedSampler(const ChunkedSampler&) = delete;
    ChunkedSampler& operator=(const ChunkedSampler&) = delete;
    // The next chunk, at `chunk` (device memory, ready in stream order).
    void add(const void* chunk);
    // After the last chunk: waits, and returns each uniform's basis index.
    std::vector<uint64_t> basis();
    uint64_t device_bytes() const;

private:
    struct Impl;
    std::unique_ptr<Impl> impl_;
};

// ---- host IR (host_ir.cpp)
struct CircuitView {
    const mh_circuit* c;
    uint32_t num_ops() const { return c->num_ops; }
    int32_t opcode(uint32_t i) const { return c->opcodes[i]; }
    const uint32_t* qubits(uint32_t i) const { return c->qubits + c->qubit_offsets[i]; }
    uint32_t num_qubits_of(uint32_t i) const { return c->qubit_offsets[i + 1] - c->qubit_offsets[i]; }
    const double* params(uint32_t i) const { return c->params + c->param_offsets[i]; }
    uint32_t num_params(uint32_t i) const { return c->param_offsets[i + 1] - c->param_offsets[i]; }
};

constexpr int32_t kOpUnitary = 21;
constexpr int32_t kOpMeasure = 22;

std::vector<cplx> gate_matrix(int32_t opcode, const double* params, uint32_t num_params);
std::vector<cplx> op_matrix(const CircuitView& cv, uint32_t i);
bool op_is_diagonal_for_fusion(const CircuitView& cv, uint32_t i);
void validate_circuit(const mh_circuit* c);

struct FusionPlan {
    std::vector<uint32_t> order, block_begin, block_qubits, qubits_begin;
    uint64_t fused_blocks = 0, gates_absorbed = 0;
    uint32_t block_count() const { return block_begin.empty() ? 0 : static_cast<uint32_t>(block_begin.size() - 1); }
};

FusionPlan plan_fusion(const CircuitView& cv, int32_t fusion, uint32_t width);
std::vector<Block> build_blocks(const CircuitView& cv, const FusionPlan& plan);

}  // namespace mh
