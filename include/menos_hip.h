/* menos_hip.h -- the C ABI of the MI355X dense statevector engine.
 *
 * A drop-in for MENOS's CUDA dense backend: the same engine contract
 * (execution plan in; statevector, seeded counts and raw Pauli expectation
 * terms out; device peak reported to the memory ledger), re-implemented as
 * hand-written HIP kernels for gfx950 behind plain C entry points. No entry
 * point throws; each returns an mh_status and, on failure, leaves a message
 * for mh_last_error() on the calling thread. The C++ side maps statuses to
 * the reference's exception types: MH_ERR_RESOURCE -> ResourceExceeded,
 * MH_ERR_INTERNAL / MH_ERR_DEVICE -> InternalError (reference
 * src/kernels/cuda/cuda_error.h:11-29, device_pool.cpp:68-92).
 *
 * Conventions (reference src/kernels/dense_kernels.h:25-47): qubits are
 * little-endian (bit j of an amplitude index is qubit j); bit j of a gate
 * matrix's row/column index is operand j in the operand list's own order;
 * matrices are row-major 2^k x 2^k complex values given as interleaved
 * (re, im) doubles; a state is 2^n contiguous amplitudes, interleaved
 * (re, im), complex128 (MH_C128) or complex64 (MH_C64).
 *
 * The MH_C128 form is bitwise the reference CPU generic kernel
 * (dense_kernels.cpp:465-524) on the same plan, and its counts are the
 * reference sampler's (sample.cpp:24-75) for the same seed. The MH_C64 form
 * stores complex64 amplitudes, computes gates in FP32 and keeps norms,
 * expectations and probability prefix sums in FP64.
 *
 * Every call that touches device memory synchronises the streams it used
 * before it returns (reference dense_kernels.h:204-206), so a following call
 * sees its writes. Device pointers passed in are never freed by the callee.
 */
#ifndef MENOS_HIP_H
#define MENOS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum mh_status {
    MH_OK = 0,
    MH_ERR_INVALID = 1,  /* a precondition of the call does not hold        */
    MH_ERR_RESOURCE = 2, /* over budget or out of device memory (bytes named) */
    MH_ERR_INTERNAL = 3, /* a HIP call failed                                */
    MH_ERR_DEVICE = 4,   /* no usable gfx950 device                          */
    MH_ERR_ABANDONED = 5 /* the execution monitor asked to stop               */
} mh_status;

typedef enum mh_precision { MH_C128 = 0, MH_C64 = 1 } mh_precision;

/* The message of the calling thread's last failed call ("" if none). */
const char* mh_last_error(void);

/* ---- Device query. Replaces kernels::cuda device_info
 * (reference src/kernels/cuda/device_info.h:14-55). -------------------------- */
typedef struct mh_device_info {
    char name[128];
    char arch[64];             /* gcnArchName, e.g. "gfx950:sramecc+:xnack-" */
    int32_t compute_units;
    int32_t warp_size;
    int32_t max_threads_per_block;
    int32_t lds_bytes_per_block;
    int32_t clock_khz;
    int32_t memory_clock_khz;
    int32_t memory_bus_width;
    uint64_t total_bytes;
    uint64_t free_bytes;
} mh_device_info;

int mh_device_query(mh_device_info* out);
/* Free device bytes plus what the pool holds for reuse
 * (reference device_info.h: device_available_bytes). */
int mh_device_available_bytes(uint64_t* out);

/* ---- Memory pool. Replaces DevicePool (reference device_pool.h:34-78): freed
 * blocks are retained and reused, sized for 288 GB of HBM3E. --------------- */
int mh_pool_alloc(uint64_t bytes, void** out);
int mh_pool_free(void* ptr);
int mh_pool_reusable_bytes(uint64_t* out);
int mh_pool_trim(void);

/* ---- State primitives. Replace kernels::cuda::{initialize, apply_matrix,
 * apply_matrix_generic, expectation_of_term} (reference dense_kernels.h:198-258). */

/* state := phase * |0...0>; the phase exp(i*global_phase) is computed by the
 * caller on the host (reference dense_kernels.cpp:21-27). */
int mh_state_init(void* state, uint32_t num_qubits, int32_t precision, double phase_re, double phase_im);

/* One gate: a 2^k x 2^k matrix (host memory, 2*4^k doubles) at `operands`
 * (host memory), k <= 6, as one HBM pass. Diagonal blocks and blocks of 1 or 2
 * qubits run in the window-pass kernel (register windows, sign and diagonal
 * stages, controlled pairs), wider dense blocks in the per-group kernel; every
 * path is bitwise the generic kernel. */
int mh_apply_matrix(void* state, uint32_t num_qubits, int32_t precision, const double* matrix,
                    const uint32_t* operands, uint32_t k);

/* The generic kernel alone, whatever the matrix (the oracle of the others,
 * reference dense_kernels.h:227-235). */
int mh_apply_matrix_generic(void* state, uint32_t num_qubits, int32_t precision, const double* matrix,
                            const uint32_t* operands, uint32_t k);

/* A whole block list in one call: block b applies the matrix at
 * matrices + matrix_offsets[b] (in complex entries) to operands
 * operands + operand_offsets[b] .. operand_offsets[b+1]. Blocks are grouped into
 * window passes (many blocks per HBM read+write of the state; each window
 * sits in one wavefront's registers); results are
 * bitwise those of calling mh_apply_matrix block by block. window_qubits = 0
 * picks the default window. On return *passes (if non-null) holds the number
 * of HBM passes made. */
int mh_apply_blocks(void* state, uint32_t num_qubits, int32_t precision, uint32_t num_blocks,
                    const uint32_t* operand_offsets, const uint32_t* operands, const uint64_t* matrix_offsets,
                    const double* matrices, uint32_t window_qubits, uint32_t* passes);

/* The HBM passes mh_apply_blocks makes for a block list (the same arrays),
 * without touching the device: pass p applies blocks
 * [pass_begin[p], pass_begin[p+1]) (pass_begin has room for num_blocks + 1
 * entries). Splitting a block list at these boundaries and applying each part
 * with mh_apply_blocks gives the same state bit for bit. The reference's
 * counterpart is the disk tier's pass plan (src/engines/dense_tiered/pass_plan.h:12-102),
 * which groups blocks into windows the same way; its in-memory engines make
 * one pass per block. */
int mh_plan_passes(uint32_t num_qubits, int32_t precision, uint32_t num_blocks, const uint32_t* operand_offsets,
                   const uint32_t* operands, const uint64_t* matrix_offsets, const double* matrices,
                   uint32_t window_qubits, uint32_t max_pass_cost, uint32_t* pass_begin, uint32_t* num_passes);

/* out[i] = state[indices[i]] for i < count (indices and out in host memory,
 * amplitudes of the state's precision): a sample of a state too large to copy
 * back, for the per-pass sampled checks of 33- and 34-qubit runs. No
 * counterpart in the reference (its largest GPU states are copied back whole,
 * DeviceBuffer::copy_to_host, device_buffer.h:25-56). */
int mh_gather(const void* state, uint32_t num_qubits, int32_t precision, const uint64_t* indices, uint64_t count,
              void* out);

/* Raw expectation terms <psi|P_t|psi> (no coefficient), one per (x, z) mask
 * pair, out = 2 doubles per term. FP64 fixed-order reduction: reproducible
 * run to run; equal to the reference's sequential sum within rounding
 * (reference dense_kernels.cpp:544-569, CUDA contract dense_kernels.h:237-246). */
int mh_expectation(const void* state, uint32_t num_qubits, int32_t precision, const uint64_t* x_masks,
                   const uint64_t* z_masks, uint32_t num_terms, double* out);

/* mh_expectation without any device allocation: no counterpart in the
 * reference. For callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc): the
 * same kernels and the same terms bit for bit, with their descriptors, block
 * partials and outputs in a static device array of 32 MiB per device, held
 * until the process ends. Term sets that need more are refused with
 * MH_ERR_INVALID naming the bytes (a tile-reduced term takes 1 MiB of
 * partials at 30 qubits, 256 KiB at 28, 64 KiB at 26 and below). Every call on one device shares that array, so calls that overlap on
 * one device (from several host threads) race; the array keeps the last
 * call's contents and nothing in it needs to start at zero. */
int mh_expectation_noalloc(const void* state, uint32_t num_qubits, int32_t precision, const uint64_t* x_masks,
                           const uint64_t* z_masks, uint32_t num_terms, double* out);

/* The basis index every sorted uniform lands on under the reference's
 * sequential cumulative walk (reference sample.cpp:55-69): the first basis
 * whose running FP64 sum of |a|^2 exceeds u, or the last basis if none does.
 * Exact (bitwise the sequential sums) by an integer scan per binade.
 * sorted_uniforms: host, ascending; basis_out: host, `shots` entries. */
int mh_sample(const void* state, uint32_t num_qubits, int32_t precision, const double* sorted_uniforms,
              uint64_t shots, uint64_t* basis_out);

/* The first basis whose cumulative sum exceeds each uniform, for a
 * non-sole region's sampler (reference cumulative_distribution.cpp:83-92);
 * uniforms need not be sorted. */
int mh_sample_unsorted(const void* state, uint32_t num_qubits, int32_t precision, const double* uniforms,
                       uint64_t count, uint64_t* basis_out);

int mh_copy_to_host(const void* state, uint32_t num_qubits, int32_t precision, void* host);
int mh_copy_from_host(void* state, uint32_t num_qubits, int32_t precision, const void* host);
/* FNV-1a 64 over the state's bytes, computed on the host after a copy-back
 * in chunks (the reference's digest, tests/memory/memory_fixtures.cpp:599-645). */
int mh_state_digest(const void* state, uint32_t num_qubits, int32_t precision, uint64_t* out);
/* The largest |a_i - b_i| (complex modulus in FP64) between two device states
 * of the same size, of either precision: the complex64 form's amplitude error
 * against a complex128 run. No counterpart in the reference (its CUDA engine
 * has one precision); used by the error reports. */
int mh_state_max_abs_diff(const void* a, int32_t precision_a, const void* b, int32_t precision_b,
                          uint32_t num_qubits, double* out);

/* ---- The engine. Replaces CudaDenseEngine::{estimate, execute}
 * (reference src/engines/dense_cuda/cuda_dense_engine.cpp:182-412) behind the
 * SimulationEngine contract (include/menos/simulation_engine.h:322-367). ---- */

/* A circuit in the reference's flat batch layout (Circuit::append_batch,
 * src/ir/circuit.cpp:34-67). Opcodes are the reference's enum Opcode
 * (include/menos/circuit.h); a Unitary op (21) carries its row-major matrix
 * as 2*4^k doubles in params; a Measure op (22) carries (qubit, clbit) as its
 * two qubits entries. */
typedef struct mh_circuit {
    uint32_t num_qubits;
    uint32_t num_ops;
    const int32_t* opcodes;
    const uint32_t* qubit_offsets; /* num_ops + 1 */
    const uint32_t* qubits;
    const uint32_t* param_offsets; /* num_ops + 1 */
    const double* params;
} mh_circuit;

/* How the gate list becomes blocks. MH_FUSE_NONE: one block per gate.
 * MH_FUSE_CONTAINMENT: the reference's rule (plan_fusion,
 * src/ir/fuse_gates.cpp:108-201, with may_join at :53-65), its default.
 * MH_FUSE_PAIRS: the containment rule, except that a gate does not
 * join a block containing a CX if it would act non-diagonally on that CX's
 * control (nor a CX whose control such a gate touches), so those blocks stay
 * controlled pairs (a 2x2 block on the target per control value), which cost
 * half of a dense 2-qubit block; no reference counterpart (the reference
 * applies every fused block the same way). MH_FUSE_WIDTH: this engine's
 * width-capped fusion (fusion_width
 * operands at most, 2..6), order-preserving; the oracle applies the same
 * block list. */
typedef enum mh_fusion {
    MH_FUSE_NONE = 0,
    MH_FUSE_CONTAINMENT = 1,
    MH_FUSE_WIDTH = 2,
    MH_FUSE_PAIRS = 3,
    /* The engine's autotuner: among the containment plan and width-capped
     * fusion at widths 2..6 it takes the plan whose gate passes the calibrated
     * time model predicts fastest for this state size and precision (window
     * passes at the measured pass rate; a wider block's own pass at the larger
     * of its HBM time and its arithmetic at the measured FP64, FP32 or
     * matrix-core rate); ties go to the containment plan. The result and the
     * estimate report the plan taken (fusion_used, fusion_width_used). No
     * reference counterpart (the reference always uses its containment rule). */
    MH_FUSE_AUTO = 4
} mh_fusion;

/* The execution monitor (reference ExecutionMonitor::at_safe_boundary,
 * include/menos/simulation_engine.h:54-116, which apply_blocks calls after
 * every block, src/engines/dense/fused_block.h:96-99). mh_engine_execute calls
 * it after each HBM pass, on the calling thread, with the number of passes
 * completed so far (from 1): there the state holds every block of those passes
 * and nothing has been handed to the caller. The reference abandons by
 * throwing; across this ABI a nonzero return abandons instead: the state is
 * freed and the call returns MH_ERR_ABANDONED (the adapter rethrows what its
 * monitor threw). The next pass may already be queued when it runs, and is
 * waited for before the state is freed. */
typedef int (*mh_monitor_fn)(void* user_data, uint64_t boundaries_crossed);

typedef struct mh_execution_request {
    const mh_circuit* circuit;
    int32_t precision;
    int32_t fusion;          /* mh_fusion */
    uint32_t fusion_width;   /* MH_FUSE_WIDTH only */
    uint32_t window_qubits;  /* 0: default */
    double global_phase;
    uint64_t shots;          /* 0: no counts */
    uint64_t seed;
    const uint32_t* target_qubits; /* measure targets (qubit -> clbit) */
    const uint32_t* target_clbits;
    uint32_t num_targets;
    uint32_t bitstring_width;      /* <= 64 in this ABI */
    uint32_t num_terms;            /* 0: no expectation */
    const uint64_t* x_masks;
    const uint64_t* z_masks;
    uint64_t budget_bytes;         /* device budget; 0 = no limit */
    int32_t want_statevector;
    int32_t record_pass_times;     /* per-pass device times into the result */
    uint32_t max_pass_cost;        /* FP operations per amplitude one HBM pass may carry; 0 = default */
    mh_monitor_fn monitor;         /* null: no monitor */
    void* monitor_data;
    int32_t want_region_sampler;   /* keep the final state on the device as a region sampler (below) */
    uint32_t reserved0;
} mh_execution_request;

/* A device-resident region sampler: the reference's CumulativeDistribution
 * (src/engines/dense/cumulative_distribution.cpp:14-92), the sampler a
 * non-sole region hands the runtime (RegionCountsSampler,
 * include/menos/simulation_engine.h:47-52; runtime/combine_counts.cpp:10-36),
 * over a state that stays in device memory instead of a host copy of its 2^n
 * amplitudes. mh_engine_execute with want_region_sampler set keeps its final
 * state and returns the sampler in result->region_sampler; the caller owns it
 * and frees it (and the state) with mh_region_sampler_destroy. */
typedef struct mh_region_sampler mh_region_sampler;

typedef struct mh_execution_result {
    /* Caller-owned outputs, any may be null when not requested. */
    void* statevector;             /* 2^n amplitudes of the request's precision */
    uint64_t* outcomes;            /* up to `shots` distinct outcomes, ascending */
    uint64_t* counts;
    uint64_t num_outcomes;         /* written */
    double* expectation_terms;     /* 2 doubles per term */
    float* pass_ms;                /* record_pass_times: one per pass (capacity pass_ms_capacity) */
    uint32_t pass_ms_capacity;
    /* Filled by the engine. */
    uint32_t num_blocks;
    uint32_t num_passes;
    uint64_t fused_blocks;
    uint64_t gates_absorbed;
    uint64_t actual_bytes;         /* device bytes of the state */
    uint64_t peak_device_bytes;    /* the memory ledger's device peak for this call */
    double execute_seconds;        /* host wall time of execute */
    double gate_seconds;           /* device time of the gate passes */
    double sample_seconds;
    double expectation_seconds;
    int32_t fusion_used;           /* the plan run: the request's, or the autotuner's choice */
    uint32_t fusion_width_used;
    mh_region_sampler* region_sampler; /* want_region_sampler: the caller's to destroy; else null */
} mh_execution_result;

typedef struct mh_cost_estimate {
    int32_t feasible;
    char infeasible_reason[256];
    uint64_t device_bytes;   /* state + scratch (reference plan.h DeviceCost) */
    uint64_t host_bytes;
    uint64_t work_units;     /* 2^n + sum over blocks 2^(n+k) + outputs (dense_engine.cpp:309-331) */
    uint32_t blocks;
    uint32_t passes;         /* HBM passes of the window schedule */
    uint64_t launches;
    uint64_t h2d_bytes;
    uint64_t d2h_bytes;
    double seconds;          /* the calibrated time model's prediction */
    int32_t fusion_used;     /* the plan costed: the request's, or the autotuner's choice */
    uint32_t fusion_width_used;
} mh_cost_estimate;

/* Never fails on an infeasible request: it reports feasible = 0 and a reason
 * (reference simulation_engine.h:349-359). Host work only (plans the blocks and
 * passes, applies the time model); it needs no device. The predicted seconds
 * are the calibrated profile's:
 *   launches * launch_seconds + passes * 2 * state / pass_bytes_per_second
 *   + state / hbm (initialisation) + 3 * state / hbm if sampling
 *   + terms * state / hbm + h2d / h2d_rate + d2h / d2h_rate. */
int mh_engine_estimate(const mh_execution_request* request, mh_cost_estimate* out);
int mh_engine_execute(const mh_execution_request* request, mh_execution_result* result);

/* ---- The host-offloaded tier ------------------------------------------------
 * The dense algorithm over a state held in pinned host memory and run on the
 * device through windows: the reference's Tier::HostOffloaded
 * (include/menos/plan.h:88-98, measured there and not built), on the pass plan
 * of its disk tier (src/engines/dense_tiered/pass_plan.h:12-102) and the shape
 * of its chunked executor (chunked_executor.h:54-75). A state of n qubits is
 * 2^(n-L) chunks of 2^L amplitudes (qubits below L local, the rest global); a
 * device window of 2^W amplitudes brings 2^(W-L) chunks together, so a pass
 * may hold any set of at most W - L global qubits. The passes follow the
 * reference's greedy rule over the plan's blocks; each block runs on the
 * windows with its operands remapped, by the in-memory engine's kernels. The
 * statevector and counts are mh_engine_execute's for the same request bit for
 * bit; raw terms agree with its terms to rounding (FP64 sums, in another
 * order). The request's fields mean what they mean there; budget_bytes bounds
 * the device bytes (two windows and the outputs' tables), and
 * want_region_sampler is refused (no state stays on the device). */
typedef struct mh_tier_pass {
    uint64_t global_mask;   /* bit (g - chunk_qubits) for each global qubit g of the pass's set */
    uint32_t first_block;   /* the plan's blocks [first_block, end_block) */
    uint32_t end_block;
} mh_tier_pass;

/* The reference's tier seams (simulation_engine.h:189-190). */
typedef struct mh_tier_request {
    uint32_t chunk_qubits;    /* L; 0: the chunk with the fewest passes (choose_pass_plan, pass_plan.h:82-101),
                                 at least W - 6 (at most 64 chunks to a window) */
    uint32_t window_qubits;   /* W; 0: the largest the budget holds (no budget: min(n, 28)) */
    mh_tier_pass* passes;     /* optional: receives up to pass_capacity of the plan's passes */
    uint32_t pass_capacity;
    uint32_t reserved0;
} mh_tier_request;

/* The reference's TierCost (include/menos/plan.h:177-196) and the device side. */
typedef struct mh_tier_cost {
    int32_t feasible;
    char infeasible_reason[256];
    uint64_t passes;              /* gate passes over the whole state */
    uint32_t chunk_qubits;
    uint32_t window_qubits;
    uint64_t link_bytes;          /* host<->device bytes: each pass's write-back and read, the outputs' reads */
    uint64_t host_state_bytes;    /* the pinned state, 2^n amplitudes (the disk tier's file_bytes) */
    uint64_t link_write_bytes;    /* the device->host part of link_bytes */
    uint64_t global_blocks;       /* blocks with an operand at or above the chunk */
    uint32_t max_global_operands;
    uint32_t blocks;
    uint64_t device_bytes;        /* two windows, the sampler's tables, the expectation partials */
    double seconds;               /* the calibrated time model (mh_calibration's h2d/d2h and pass rates) */
    int32_t fusion_used;
    uint32_t fusion_width_used;
} mh_tier_cost;

/* Host work only; never fails on an infeasible request (feasible = 0 and a
 * reason), as DiskDenseEngine::estimate (disk_dense_engine.cpp:307-341). */
int mh_tier_estimate(const mh_execution_request* request, const mh_tier_request* tier, mh_tier_cost* out);
/* Runs the plan mh_tier_estimate reports (written to `plan` when not null);
 * num_passes counts tier passes, pass_ms their wall times, actual_bytes the
 * windows; the monitor is called at the end of each pass (chunked_executor.cpp:
 * 216-221), with the passes completed. */
int mh_tier_execute(const mh_execution_request* request, const mh_tier_request* tier, mh_execution_result* result,
                    mh_tier_cost* plan);

/* Basis indices for `count` uniforms in [0, 1), in any order: index i is the
 * first basis state whose running sum fl(c + |a|^2) in index order exceeds
 * uniforms[i] (the last one if none), as CumulativeDistribution::sample_one
 * (cumulative_distribution.cpp:83-92) finds it, bit for bit. Reads only the
 * sampler's device tables; calls on one sampler may run from several threads.
 * A uniform outside [0, 1) is MH_ERR_INVALID. */
int mh_region_sampler_draw(const mh_region_sampler* sampler, const double* uniforms, uint64_t count,
                           uint64_t* basis_out);
/* The state's qubit count and the device bytes the sampler holds. */
int mh_region_sampler_info(const mh_region_sampler* sampler, uint32_t* num_qubits, uint64_t* device_bytes);
int mh_region_sampler_destroy(mh_region_sampler* sampler);

/* The fusion plan the engine would run (FusionPlan layout, reference
 * src/ir/fuse_gates.h:38-76). Outputs sized for the worst case by the caller:
 * order[num_ops], block_begin[num_ops+2], block_qubits[sum of operand counts],
 * qubits_begin[num_ops+2]. */
int mh_plan_fusion(const mh_circuit* circuit, int32_t fusion, uint32_t fusion_width, uint32_t* num_blocks,
                   uint32_t* order, uint32_t* block_begin, uint32_t* block_qubits, uint32_t* qubits_begin,
                   uint64_t* fused_blocks, uint64_t* gates_absorbed);

/* A named gate's matrix as the engine generates it on the host (reference
 * src/ir/gate_table.cpp:19-265); returns k through *k. */
int mh_gate_matrix(int32_t opcode, const double* params, double* out, uint32_t* k);

/* ---- Calibration for the time model (reference src/plan/cuda_calibration.cpp:42-153,
 * HardwareProfile cuda_* fields, hardware_profile.h:52-77). ------------------ */
typedef struct mh_calibration {
    double hbm_bytes_per_second;     /* measured copy bandwidth (read + write) */
    double pass_bytes_per_second;    /* a window pass at 24 qubits, c128 */
    double launch_seconds;           /* an empty launch, synchronised */
    double h2d_bytes_per_second;     /* pinned */
    double d2h_bytes_per_second;     /* pinned */
    double fp64_ops_per_second;      /* non-fused DMUL/DADD */
    double mfma_flops_per_second;    /* complex64 6-qubit blocks on the matrix cores (v_mfma_f32_32x32x2_f32) */
} mh_calibration;

/* Measures the device (about a second) and installs the result as the time
 * model's profile. */
int mh_calibrate(mh_calibration* out);
/* Installs a profile measured earlier or made up for a test (the reference's
 * planner takes its HardwareProfile the same way, hardware_profile.h:52-77),
 * and reads the one in force. Neither touches the device. */
int mh_set_calibration(const mh_calibration* profile);
int mh_get_calibration(mh_calibration* out);

/* ---- Batched execution: many independent circuits per device launch.
 * Replaces the reference's sequential batch (run_batch_with_options,
 * src/api/batch.cpp:10-27, one full run per circuit with no device
 * residency) for expectation-value workloads (the Estimator's path,
 * python/menos/qiskit/estimator.py:91-133). Every circuit is planned and
 * compiled on host threads; the states of a round of circuits share one
 * device arena; pass p of every circuit in the round runs in one launch;
 * every circuit's terms are reduced in two launches. Results are the
 * single-circuit engine's bit for bit: complex128 statevectors as the
 * reference's, expectation terms as mh_expectation's. -------------------- */
typedef struct mh_batch_request {
    uint32_t num_circuits;
    const mh_circuit* circuits;    /* each with at least 10 qubits */
    int32_t precision;
    int32_t fusion;                /* mh_fusion */
    uint32_t fusion_width;
    const double* global_phases;   /* per circuit, or null (all 0) */
    const uint32_t* term_offsets;  /* num_circuits + 1: circuit c owns terms [term_offsets[c], term_offsets[c+1]) */
    const uint64_t* x_masks;
    const uint64_t* z_masks;
    uint64_t round_bytes;          /* device bytes of one round's states; 0: half the available memory */
    uint32_t host_threads;         /* planning threads; 0: the hardware's, at most 16 */
    int32_t want_statevectors;     /* copy every final state into result->statevectors */
    uint32_t max_pass_cost;        /* FP operations per amplitude one HBM pass may carry; 0 = default */
    /* The device schedule. group_bytes = 0: pass p of every circuit of a planning group in one launch.
     * group_bytes > 0: circuits split, largest first, into cache groups of at most group_bytes of state,
     * each running all its passes back to back (so its states stay in the 256 MB Infinity Cache), the
     * groups rotating over kernel_streams streams (0: 2; at most 4). cached_states: ordinary (cached)
     * state loads and stores instead of non-temporal ones. Results do not depend on any of these. */
    uint64_t group_bytes;
    uint32_t kernel_streams;
    int32_t cached_states;
} mh_batch_request;

typedef struct mh_batch_result {
    double* expectation_terms;     /* 2 doubles per term, in term order */
    void* statevectors;            /* want_statevectors: the states back to back, in circuit order */
    double* completion_seconds;    /* per circuit, or null: from the call's start to its results on the host */
    uint32_t rounds;
    uint32_t launches;
    double plan_seconds;           /* host: fusion plans and block matrices */
    double device_seconds;         /* device: init, passes and reductions of every round,
                                      with the host's pass programs (compile_seconds) */
    double compile_seconds;        /* host: window-pass programs and launch tables */
    double total_seconds;
} mh_batch_result;

int mh_batch_execute(const mh_batch_request* request, mh_batch_result* result);

/* ---- A resident batch replayed from a captured HIP graph: the variational
 * estimator's loop (the same circuits with new parameters, again and again).
 * Replaces re-running run_batch_with_options (src/api/batch.cpp:10-27) with
 * the plan cache's parameter-only reuse (plan_job_cached /
 * reparameterize_direct_plan, src/plan/plan_cache.cpp:233-276).
 *
 * mh_batch_graph_create plans and compiles every circuit of the request once,
 * keeps every state resident in one arena (all of them must fit the device:
 * MH_ERR_RESOURCE otherwise) with every table the kernels read, and captures
 * the whole round (matrix upload from pinned memory, initialisation, every
 * pass, the term reductions, the terms back to pinned memory) into one graph.
 * Uses the request's circuits, precision, fusion, global phases, terms, host
 * threads, pass cost and schedule fields; want_statevectors must be 0, and
 * blocks must be at most 2-qubit dense or diagonal.
 * mh_batch_graph_update takes new parameters for the same circuits (same
 * opcodes, qubits and parameter counts) and global phases (null: 0): block
 * matrices and pass programs are rebuilt on host threads and the new matrix
 * table is staged for the next run. A circuit whose structure differs, or
 * whose new matrices change its pass program (an entry that becomes or stops
 * being exactly 0 or +-1), is refused with MH_ERR_INVALID and nothing changes.
 * mh_batch_graph_run replays the graph once and writes the terms (2 doubles
 * per term, term order) into result->expectation_terms; it fills rounds,
 * launches (those the graph holds), the last update's plan and compile
 * seconds, and total_seconds. Terms equal mh_batch_execute's bit for bit for
 * the same circuits and schedule. One thread at a time per graph. */
typedef struct mh_batch_graph mh_batch_graph;
int mh_batch_graph_create(const mh_batch_request* request, mh_batch_graph** out);
int mh_batch_graph_update(mh_batch_graph* graph, const mh_circuit* circuits, const double* global_phases);
/* The same with the parameters alone: every circuit's parameters, in circuit
 * order, concatenated (mh_batch_graph_param_count of them; each circuit's
 * params laid out as in its mh_circuit), and the global phases (null: 0). */
int mh_batch_graph_set_params(mh_batch_graph* graph, const double* params, const double* global_phases);
int mh_batch_graph_param_count(const mh_batch_graph* graph, uint64_t* out);
int mh_batch_graph_run(mh_batch_graph* graph, mh_batch_result* result);
int mh_batch_graph_destroy(mh_batch_graph* graph);

/* ---- Serving: a resident job server. Replaces the per-pub sequential
 * `run` loops of the reference's Python primitives (MENOSEstimator._run_pub,
 * python/menos/qiskit/estimator.py:91-146: one run per observable;
 * MENOSSampler._run_pub, sampler.py:93-129) for a stream of jobs.
 *
 * Continuous batching: jobs are queued as they are submitted, and several
 * lanes (threads with streams of their own) take them. An idle lane admits
 * the queued jobs, in arrival order, whose states fit the device memory the
 * other lanes leave within the budget (the ledger admission, reference
 * device_buffer.cpp:22-27), runs them as one batched round (mh_batch_execute's
 * kernels) and completes them; jobs that arrive meanwhile are admitted by the
 * next idle lane and their work overlaps on the device with the rounds
 * already running. Within a round, jobs whose state is at most an eighth of
 * its largest complete first. A job whose state alone exceeds the budget
 * completes with MH_ERR_RESOURCE. A job's results do not depend on which
 * lane or round ran it, nor on the other jobs beside it.
 *
 * Fusion plans are cached by circuit structure (opcodes, qubits, parameter
 * counts, unitary entries), so a parameter-only update of a circuit seen
 * before reuses its plan, as the reference's plan cache does
 * (plan_job_cached / reparameterize_direct_plan, src/plan/plan_cache.cpp:233-276,
 * planner.cpp:1601-1652). Results are the single-circuit engine's bit for bit:
 * raw terms as mh_engine_execute's, counts as its counts for the same seed,
 * and each observable's value combined from its raw terms in term order from
 * (0, 0) as the reference's combine_expectation does
 * (src/runtime/combine_expectation.cpp:7-25). --------------------------- */
typedef struct mh_server mh_server;

typedef struct mh_server_options {
    uint64_t budget_bytes;          /* device bytes the states of all running rounds may hold; 0: 256 GB, capped to what the device has */
    uint32_t max_round_jobs;        /* 0: no cap */
    uint32_t host_threads;          /* planning threads; 0: the hardware's, at most 16 */
    int32_t fusion;                 /* mh_fusion */
    uint32_t fusion_width;
    uint32_t plan_cache_capacity;   /* cached fusion plans; 0: 256 */
    uint64_t reserve_bytes;         /* device bytes for round states allocated at creation (at most the
                                       budget); 0: the arena grows with the rounds */
    int32_t complex64_pairs;        /* nonzero: complex64 jobs use MH_FUSE_PAIRS (complex128 jobs keep `fusion`) */
    uint32_t lanes;                 /* rounds that may run at once; 0: 3, at most 8 */
    uint32_t reserved0;
    uint64_t large_job_bytes;       /* nonzero: a job whose state has at least these bytes runs in a round of its
                                       own, and while one such round runs the others wait in arrival order as
                                       smaller jobs pass them (two large states sharing the device both run
                                       about twice as long); 0: arrival order only */
} mh_server_options;

typedef enum mh_job_kind { MH_JOB_ESTIMATOR = 0, MH_JOB_SAMPLER = 1 } mh_job_kind;

typedef struct mh_job {
    int32_t kind;                   /* mh_job_kind */
    int32_t precision;              /* a sampler job must be MH_C128 */
    const mh_circuit* circuit;      /* 10..30 qubits; copied at submission */
    double global_phase;
    /* Estimator: observables, each a sum of Pauli terms with complex coefficients. */
    uint32_t num_observables;
    const uint32_t* observable_offsets; /* num_observables + 1 offsets into the term arrays */
    const uint64_t* x_masks;
    const uint64_t* z_masks;
    const double* coefficients;     /* (re, im) per term */
    /* Sampler: shots drawn with `seed` (the pub's SplitMix64 seed, mh_derive_seeds). */
    uint64_t shots;
    uint64_t seed;
    uint32_t num_targets;
    const uint32_t* target_qubits;  /* measure targets (qubit -> clbit) */
    const uint32_t* target_clbits;
    uint32_t bitstring_width;       /* <= 64 */
} mh_job;

typedef struct mh_job_result {
    /* Caller-owned outputs, sized from the job; any may be null. */
    double* observable_values;      /* 2 doubles (re, im) per observable */
    double* raw_terms;              /* 2 doubles per term */
    uint64_t* outcomes;             /* capacity `shots`: distinct outcomes, ascending */
    uint64_t* counts;
    /* Filled by the server. */
    int32_t status;                 /* MH_OK or the job's failure */
    char error[256];
    uint64_t num_outcomes;
    uint32_t round;                 /* the round that ran it (1-based) */
    int32_t plan_cache_hit;
    double submitted_seconds;       /* from the server's start */
    double started_seconds;         /* its round's start */
    double completed_seconds;       /* its results ready */
    uint32_t lane;                  /* the lane that ran it (0-based) */
} mh_job_result;

typedef struct mh_server_stats {
    uint64_t jobs_completed;
    uint64_t jobs_failed;
    uint32_t rounds;
    uint64_t plan_cache_hits;
    uint64_t plan_cache_misses;
    uint64_t peak_round_bytes;      /* the largest round's states */
    uint64_t budget_bytes;          /* the budget in force */
    double busy_seconds;            /* wall time inside rounds, summed over lanes, split into: */
    double plan_seconds;            /*   host: fusion plans (or cache hits) and block matrices */
    double arena_seconds;           /*   growing the round-state arena */
    double compile_seconds;         /*   host: window-pass programs */
    double pass_seconds;            /*   device: initialisation and gate passes */
    double expectation_seconds;     /*   device: term reductions */
    double sample_seconds;          /*   device and host: sampler draws and counts */
    uint32_t lanes;                 /* lanes in force */
    uint32_t peak_concurrent_rounds;/* the most rounds that ran at once */
    uint64_t peak_held_bytes;       /* the ledger's peak over all lanes */
} mh_server_stats;

int mh_server_create(const mh_server_options* options, mh_server** out);
/* Copies the job (circuit, terms, targets) and returns at once with its ticket. */
int mh_server_submit(mh_server* server, const mh_job* job, uint64_t* ticket);
/* Blocks until the ticket's job is complete and writes its result; a ticket
 * is waited for once. Returns MH_OK when the result was written (the job's
 * own outcome is result->status). */
int mh_server_wait(mh_server* server, uint64_t ticket, mh_job_result* result);
/* Submits job i at arrival_seconds[i] (ascending, from the call's start) on a
 * submitter thread, then waits for every job. For arrival-process benchmarks. */
int mh_server_replay(mh_server* server, uint32_t count, const mh_job* jobs, const double* arrival_seconds,
                     mh_job_result* results);
int mh_server_stats_get(mh_server* server, mh_server_stats* out);
/* Completes the queued jobs, stops the worker and frees the server. */
int mh_server_destroy(mh_server* server);

/* The first `count` outputs of SplitMix64 seeded with `seed`: the per-circuit
 * seeds of run_batch and the per-pub seeds of the sampler (reference
 * src/common/splitmix64.h:15-28, src/api/batch.cpp:18-24,
 * src/python/bindings.cpp:581-595). */
int mh_derive_seeds(uint64_t seed, uint32_t count, uint64_t* out);

#ifdef __cplusplus
}
#endif

#endif /* MENOS_HIP_H */
