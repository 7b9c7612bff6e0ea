// A batch's window-pass program on the host: every state's compiled passes,
// the tables the kernels read (packed per state) and the launch lists. Built by
// compile_batch_program (window.hip) and issued by issue_batch_program, either
// straight away (BatchPassRunner) or once into a captured HIP graph that is
// replayed with new matrix tables (graph.hip).
#pragma once

#include <cstdint>
#include <functional>
#include <vector>

#include "mh_internal.h"
#include "window.h"

namespace mh {

// The ops, sign terms and layout the window kernel reads (window.hip).
struct WOp;
struct SignTerm;
struct WBatchArgs;

// One launch of window_batch_kernel: pass `pass` of the states of launch
// group `group`, its argument list at args[args_at, args_at + len).
struct BatchLaunch {
    std::size_t args_at = 0, groups_at = 0, len = 0;
    uint32_t total_groups = 0;
    bool pairs = false;
    uint32_t group = 0, pass = 0;
};

// The tables of a program in device memory.
struct BatchDeviceTables {
    const uint64_t* bundles = nullptr;
    const WOp* ops = nullptr;
    const SignTerm* terms = nullptr;
    const void* mats = nullptr;
    const WBatchArgs* args = nullptr;
    const uint32_t* groups = nullptr;
};

// Element sizes of the packed tables, for byte counts outside window.hip.
std::size_t wop_bytes();
std::size_t sign_term_bytes();
std::size_t wbatch_args_bytes();

struct BatchProgramData;  // window.hip

// The host side of a compiled batch (opaque contents; see window.hip).
class BatchProgram {
public:
    BatchProgram();
    ~BatchProgram();
    BatchProgram(BatchProgram&&) noexcept;
    BatchProgram& operator=(BatchProgram&&) noexcept;
    BatchProgram(const BatchProgram&) = delete;
    BatchProgram& operator=(const BatchProgram&) = delete;

    int32_t precision() const;
    // Packed tables, as bytes, in the layout the device copies must have.
    const std::vector<unsigned char>& bundles() const;
    const std::vector<unsigned char>& ops() const;
    const std::vector<unsigned char>& terms() const;
    const std::vector<unsigned char>& mats() const;
    const std::vector<unsigned char>& args() const;
    const std::vector<unsigned char>& groups() const;
    // Whether another program of the same states has the same passes, ops,
    // sign terms and launch lists (its matrix table alone may differ).
    bool same_shape(const BatchProgram& other) const;
    bool has_generic() const;
    // Passes of item i, and the launch group it belongs to.
    uint32_t pass_count(std::size_t item) const;
    uint32_t launch_group_of(std::size_t item) const;
    std::size_t item_count() const;
    // Whether some state needs an initialisation kernel (its first pass is not
    // a window pass that can generate it).
    bool needs_initialisation() const;
    uint32_t launch_count() const;

    BatchProgramData& data();
    const BatchProgramData& data() const;

private:
    BatchProgramData* d_;
};

// Plans and compiles every item's passes on `threads` host threads and packs
// the tables and launch lists for `sched` (state addresses are taken from the
// items: the launch lists hold offsets from the lowest one).
BatchProgram compile_batch_program(int32_t precision, const BatchItem* items, std::size_t count, uint32_t threads,
                                   const BatchSchedule& sched);

// Called after each launch of issue_batch_program (pass l.pass of launch
// group l.group, with the generic launches of that pass) with the stream it
// went on: every state of that group with l.pass + 1 passes is then final.
using AfterLaunch = std::function<void(const BatchLaunch&, hipStream_t)>;

// Issues a compiled program's launches (no waits, no copies, no allocation),
// launch group g on kernel stream (first_stream + g) mod sched.kernel_streams
// (everything on the main stream when a pass is generic: such passes launch
// the generic kernel, which allocates and synchronises). Returns the launch count.
uint32_t issue_batch_program(const BatchProgram& prog, const BatchItem* items, const BatchDeviceTables& dev,
                             const BatchSchedule& sched, uint32_t first_stream, const AfterLaunch& after = {});

// What BatchPassRunner::add calls around a group's launches: `compiled` once
// the group's program is built (before its first launch), `after` after each
// launch (see AfterLaunch).
struct BatchHooks {
    std::function<void(const BatchProgram&)> compiled;
    AfterLaunch after;
};

}  // namespace mh
