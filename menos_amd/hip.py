"""ctypes bindings of libmenos_hip.so (the C ABI in include/menos_hip.h).

The library is built in-tree by ``menos_amd.build`` and loaded from this
package's directory; loading fails loudly if it is missing (there is no other
implementation to fall back to).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import workloads as W

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmenos_hip.so")

C128, C64 = 0, 1
FUSE_NONE, FUSE_CONTAINMENT, FUSE_WIDTH, FUSE_PAIRS, FUSE_AUTO = 0, 1, 2, 3, 4
MH_OK, MH_ERR_INVALID, MH_ERR_RESOURCE, MH_ERR_INTERNAL, MH_ERR_DEVICE, MH_ERR_ABANDONED = range(6)

# int (*)(void* user_data, uint64_t boundaries_crossed): nonzero abandons.
MONITOR_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_uint64)

# Every entry point include/menos_hip.h declares.
EXPORTS = (
    "mh_last_error", "mh_device_query", "mh_device_available_bytes", "mh_pool_alloc", "mh_pool_free",
    "mh_pool_reusable_bytes", "mh_pool_trim", "mh_state_init", "mh_apply_matrix", "mh_apply_matrix_generic",
    "mh_apply_blocks", "mh_expectation", "mh_sample", "mh_sample_unsorted", "mh_copy_to_host",
    "mh_copy_from_host", "mh_state_digest", "mh_engine_estimate", "mh_engine_execute", "mh_plan_fusion",
    "mh_gate_matrix", "mh_calibrate", "mh_batch_execute", "mh_state_max_abs_diff", "mh_server_create",
    "mh_server_submit", "mh_server_wait", "mh_server_replay", "mh_server_stats_get", "mh_server_destroy",
    "mh_derive_seeds", "mh_set_calibration", "mh_get_calibration", "mh_expectation_noalloc",
    "mh_plan_passes", "mh_gather", "mh_batch_graph_create", "mh_batch_graph_update", "mh_batch_graph_run",
    "mh_batch_graph_destroy", "mh_batch_graph_set_params", "mh_batch_graph_param_count",
    "mh_region_sampler_draw", "mh_region_sampler_info", "mh_region_sampler_destroy",
    "mh_tier_estimate", "mh_tier_execute",
)
JOB_ESTIMATOR, JOB_SAMPLER = 0, 1

_u32p = ctypes.POINTER(ctypes.c_uint32)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_i32p = ctypes.POINTER(ctypes.c_int32)
_f64p = ctypes.POINTER(ctypes.c_double)
_f32p = ctypes.POINTER(ctypes.c_float)


class HipError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"[status {status}] {message}")
        self.status = status


class ResourceExceeded(HipError):
    """Over the device budget or out of device memory (reference ResourceExceeded)."""


class DeviceInfo(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 128), ("arch", ctypes.c_char * 64), ("compute_units", ctypes.c_int32),
                ("warp_size", ctypes.c_int32), ("max_threads_per_block", ctypes.c_int32),
                ("lds_bytes_per_block", ctypes.c_int32), ("clock_khz", ctypes.c_int32),
                ("memory_clock_khz", ctypes.c_int32), ("memory_bus_width", ctypes.c_int32),
                ("total_bytes", ctypes.c_uint64), ("free_bytes", ctypes.c_uint64)]


class CCircuit(ctypes.Structure):
    _fields_ = [("num_qubits", ctypes.c_uint32), ("num_ops", ctypes.c_uint32), ("opcodes", _i32p),
                ("qubit_offsets", _u32p), ("qubits", _u32p), ("param_offsets", _u32p), ("params", _f64p)]


class CRequest(ctypes.Structure):
    _fields_ = [("circuit", ctypes.POINTER(CCircuit)), ("precision", ctypes.c_int32), ("fusion", ctypes.c_int32),
                ("fusion_width", ctypes.c_uint32), ("window_qubits", ctypes.c_uint32),
                ("global_phase", ctypes.c_double), ("shots", ctypes.c_uint64), ("seed", ctypes.c_uint64),
                ("target_qubits", _u32p), ("target_clbits", _u32p), ("num_targets", ctypes.c_uint32),
                ("bitstring_width", ctypes.c_uint32), ("num_terms", ctypes.c_uint32), ("x_masks", _u64p),
                ("z_masks", _u64p), ("budget_bytes", ctypes.c_uint64), ("want_statevector", ctypes.c_int32),
                ("record_pass_times", ctypes.c_int32), ("max_pass_cost", ctypes.c_uint32),
                ("monitor", MONITOR_FN), ("monitor_data", ctypes.c_void_p),
                ("want_region_sampler", ctypes.c_int32), ("reserved0", ctypes.c_uint32)]


class CResult(ctypes.Structure):
    _fields_ = [("statevector", ctypes.c_void_p), ("outcomes", _u64p), ("counts", _u64p),
                ("num_outcomes", ctypes.c_uint64), ("expectation_terms", _f64p), ("pass_ms", _f32p),
                ("pass_ms_capacity", ctypes.c_uint32), ("num_blocks", ctypes.c_uint32),
                ("num_passes", ctypes.c_uint32), ("fused_blocks", ctypes.c_uint64),
                ("gates_absorbed", ctypes.c_uint64), ("actual_bytes", ctypes.c_uint64),
                ("peak_device_bytes", ctypes.c_uint64), ("execute_seconds", ctypes.c_double),
                ("gate_seconds", ctypes.c_double), ("sample_seconds", ctypes.c_double),
                ("expectation_seconds", ctypes.c_double), ("fusion_used", ctypes.c_int32),
                ("fusion_width_used", ctypes.c_uint32), ("region_sampler", ctypes.c_void_p)]


class CCost(ctypes.Structure):
    _fields_ = [("feasible", ctypes.c_int32), ("infeasible_reason", ctypes.c_char * 256),
                ("device_bytes", ctypes.c_uint64), ("host_bytes", ctypes.c_uint64), ("work_units", ctypes.c_uint64),
                ("blocks", ctypes.c_uint32), ("passes", ctypes.c_uint32), ("launches", ctypes.c_uint64),
                ("h2d_bytes", ctypes.c_uint64), ("d2h_bytes", ctypes.c_uint64), ("seconds", ctypes.c_double),
                ("fusion_used", ctypes.c_int32), ("fusion_width_used", ctypes.c_uint32)]


class Calibration(ctypes.Structure):
    _fields_ = [("hbm_bytes_per_second", ctypes.c_double), ("pass_bytes_per_second", ctypes.c_double),
                ("launch_seconds", ctypes.c_double), ("h2d_bytes_per_second", ctypes.c_double),
                ("d2h_bytes_per_second", ctypes.c_double), ("fp64_ops_per_second", ctypes.c_double),
                ("mfma_flops_per_second", ctypes.c_double)]


_lib = None


def lib() -> ctypes.CDLL:
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(f"{LIB_PATH} is not built: run `python -m menos_amd.build`")
        _lib = ctypes.CDLL(LIB_PATH)
        _lib.mh_last_error.restype = ctypes.c_char_p
    return _lib


def check(status: int) -> None:
    if status != MH_OK:
        msg = lib().mh_last_error().decode(errors="replace")
        raise (ResourceExceeded if status == MH_ERR_RESOURCE else HipError)(status, msg)


def _ptr(a, t):
    return a.ctypes.data_as(t) if a is not None and a.size else ctypes.cast(None, t)


def device_info() -> DeviceInfo:
    info = DeviceInfo()
    check(lib().mh_device_query(ctypes.byref(info)))
    return info


def available_bytes() -> int:
    """Free device bytes plus what the pool holds for reuse."""
    v = ctypes.c_uint64()
    check(lib().mh_device_available_bytes(ctypes.byref(v)))
    return v.value


def amp_dtype(precision: int):
    return np.complex128 if precision == C128 else np.complex64


class DeviceState:
    """A statevector in device memory, allocated from the engine's pool."""

    def __init__(self, num_qubits: int, precision: int = C128):
        self.n, self.precision = num_qubits, precision
        self.nbytes = (1 << num_qubits) * (16 if precision == C128 else 8)
        p = ctypes.c_void_p()
        check(lib().mh_pool_alloc(ctypes.c_uint64(self.nbytes), ctypes.byref(p)))
        self.ptr = p

    def max_abs_diff(self, other: "DeviceState") -> float:
        """Largest |a_i - b_i| against another device state of the same size."""
        out = ctypes.c_double()
        check(lib().mh_state_max_abs_diff(self.ptr, self.precision, other.ptr, other.precision, self.n,
                                          ctypes.byref(out)))
        return out.value

    def free(self):
        if self.ptr:
            lib().mh_pool_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def init(self, phase: complex = 1.0):
        check(lib().mh_state_init(self.ptr, self.n, self.precision, ctypes.c_double(phase.real),
                                  ctypes.c_double(complex(phase).imag)))

    def upload(self, host: np.ndarray):
        h = np.ascontiguousarray(host, amp_dtype(self.precision))
        check(lib().mh_copy_from_host(self.ptr, self.n, self.precision, h.ctypes.data_as(ctypes.c_void_p)))

    def download(self) -> np.ndarray:
        out = np.empty(1 << self.n, amp_dtype(self.precision))
        check(lib().mh_copy_to_host(self.ptr, self.n, self.precision, out.ctypes.data_as(ctypes.c_void_p)))
        return out

    def digest(self) -> int:
        h = ctypes.c_uint64()
        check(lib().mh_state_digest(self.ptr, self.n, self.precision, ctypes.byref(h)))
        return h.value

    def apply_matrix(self, matrix, operands, generic=False):
        m = np.ascontiguousarray(matrix, np.complex128)
        ops = np.asarray(operands, np.uint32)
        fn = lib().mh_apply_matrix_generic if generic else lib().mh_apply_matrix
        check(fn(self.ptr, self.n, self.precision, m.ctypes.data_as(_f64p), _ptr(ops, _u32p), len(ops)))

    def apply_blocks(self, blocks, window_qubits=0) -> int:
        """blocks: [(matrix, operands)]. Returns the HBM pass count."""
        offs, ops, moffs, mats = _block_arrays(blocks)
        passes = ctypes.c_uint32()
        check(lib().mh_apply_blocks(self.ptr, self.n, self.precision, len(blocks), _ptr(offs, _u32p),
                                    _ptr(ops, _u32p), _ptr(moffs, _u64p), mats.ctypes.data_as(_f64p),
                                    window_qubits, ctypes.byref(passes)))
        return passes.value

    def gather(self, indices) -> np.ndarray:
        """The amplitudes at `indices` (host array of the state's dtype)."""
        idx = np.ascontiguousarray(indices, np.uint64)
        out = np.empty(max(idx.size, 1), amp_dtype(self.precision))
        check(lib().mh_gather(self.ptr, self.n, self.precision, _ptr(idx, _u64p), ctypes.c_uint64(idx.size),
                              out.ctypes.data_as(ctypes.c_void_p)))
        return out[:idx.size]

    def expectation(self, terms, noalloc=False) -> np.ndarray:
        """Raw terms; noalloc=True takes mh_expectation_noalloc (its static
        device workspace, no allocation)."""
        xs = np.array([t[0] for t in terms], np.uint64)
        zs = np.array([t[1] for t in terms], np.uint64)
        out = np.zeros(2 * max(len(terms), 1), np.float64)
        fn = lib().mh_expectation_noalloc if noalloc else lib().mh_expectation
        check(fn(self.ptr, self.n, self.precision, _ptr(xs, _u64p), _ptr(zs, _u64p), len(terms),
                 out.ctypes.data_as(_f64p)))
        return out.view(np.complex128)[:len(terms)].copy()

    def sample(self, sorted_uniforms) -> np.ndarray:
        u = np.ascontiguousarray(sorted_uniforms, np.float64)
        out = np.zeros(max(u.size, 1), np.uint64)
        check(lib().mh_sample(self.ptr, self.n, self.precision, _ptr(u, _f64p), ctypes.c_uint64(u.size),
                              out.ctypes.data_as(_u64p)))
        return out[:u.size]

    def sample_unsorted(self, uniforms) -> np.ndarray:
        u = np.ascontiguousarray(uniforms, np.float64)
        out = np.zeros(max(u.size, 1), np.uint64)
        check(lib().mh_sample_unsorted(self.ptr, self.n, self.precision, _ptr(u, _f64p), ctypes.c_uint64(u.size),
                                       out.ctypes.data_as(_u64p)))
        return out[:u.size]


def _block_arrays(blocks):
    """The C ABI's flat block arrays: operand offsets, operands, matrix
    offsets (complex entries), matrices."""
    offs, ops, moffs, mats = [0], [], [], []
    at = 0
    for m, q in blocks:
        ops.extend(q)
        offs.append(len(ops))
        moffs.append(at)
        flat = np.ascontiguousarray(m, np.complex128).reshape(-1)
        mats.append(flat)
        at += flat.size
    return (np.array(offs, np.uint32), np.array(ops, np.uint32), np.array(moffs, np.uint64),
            np.concatenate(mats) if mats else np.zeros(1, np.complex128))


def plan_passes(num_qubits: int, blocks, precision=C128, window_qubits=0, max_pass_cost=0) -> list[tuple[int, int]]:
    """The HBM passes mh_apply_blocks makes: [(first block, end block)]."""
    offs, ops, moffs, mats = _block_arrays(blocks)
    begin = np.zeros(len(blocks) + 1, np.uint32)
    count = ctypes.c_uint32()
    check(lib().mh_plan_passes(num_qubits, precision, len(blocks), _ptr(offs, _u32p), _ptr(ops, _u32p),
                               _ptr(moffs, _u64p), mats.ctypes.data_as(_f64p), window_qubits, max_pass_cost,
                               _ptr(begin, _u32p), ctypes.byref(count)))
    return [(int(begin[p]), int(begin[p + 1])) for p in range(count.value)]


def gate_matrix(opcode: int, params=()) -> np.ndarray:
    p = np.asarray(params, np.float64)
    out = np.zeros(64, np.complex128)
    k = ctypes.c_uint32()
    check(lib().mh_gate_matrix(ctypes.c_int32(opcode), _ptr(p, _f64p), out.ctypes.data_as(_f64p), ctypes.byref(k)))
    d = 1 << k.value
    return out[:d * d].reshape(d, d)


class _CircuitKeep:
    def __init__(self, circuit: W.Circuit):
        self.arrays = circuit.flat()
        op, qo, qs, po, ps = self.arrays
        self.c = CCircuit(circuit.num_qubits, len(op), _ptr(op, _i32p), _ptr(qo, _u32p), _ptr(qs, _u32p),
                          _ptr(po, _u32p), _ptr(ps, _f64p))


def plan_fusion(circuit: W.Circuit, fusion=FUSE_CONTAINMENT, width=0):
    """(order, block_begin, block_qubits, qubits_begin, fused_blocks, gates_absorbed)."""
    keep = _CircuitKeep(circuit)
    n_ops = len(circuit.ops)
    nq = sum(len(o[1]) for o in circuit.ops) + 1
    order = np.zeros(max(n_ops, 1), np.uint32)
    bb = np.zeros(n_ops + 2, np.uint32)
    bq = np.zeros(nq, np.uint32)
    qb = np.zeros(n_ops + 2, np.uint32)
    nb = ctypes.c_uint32()
    fb, ga = ctypes.c_uint64(), ctypes.c_uint64()
    check(lib().mh_plan_fusion(ctypes.byref(keep.c), fusion, width, ctypes.byref(nb), _ptr(order, _u32p),
                               _ptr(bb, _u32p), _ptr(bq, _u32p), _ptr(qb, _u32p), ctypes.byref(fb),
                               ctypes.byref(ga)))
    b = nb.value
    return order[:n_ops], bb[:b + 1], bq[:qb[b]], qb[:b + 1], fb.value, ga.value


class Result:
    def __init__(self, **kw):
        self.__dict__.update(kw)


class HipDenseEngine:
    """The {Dense, Hip, InMemory} engine (SimulationEngine::estimate / execute)."""

    name = "dense_hip"

    @staticmethod
    def _request(circuit, keep, *, precision, fusion, fusion_width, window_qubits, shots, seed, targets, width,
                 terms, budget_bytes, statevector, record_pass_times, max_pass_cost=0):
        targets = circuit.measure_targets() if targets is None else targets
        width = circuit.num_clbits if width is None else width
        tq = np.array([t[0] for t in targets], np.uint32)
        tc = np.array([t[1] for t in targets], np.uint32)
        xs = np.array([t[0] for t in terms], np.uint64)
        zs = np.array([t[1] for t in terms], np.uint64)
        req = CRequest(ctypes.pointer(keep.c), precision, fusion, fusion_width, window_qubits, circuit.global_phase,
                       shots, seed, _ptr(tq, _u32p), _ptr(tc, _u32p), len(targets), width, len(terms),
                       _ptr(xs, _u64p), _ptr(zs, _u64p), budget_bytes, int(bool(statevector)),
                       int(bool(record_pass_times)), max_pass_cost)
        return req, (tq, tc, xs, zs)

    def estimate(self, circuit: W.Circuit, *, precision=C128, fusion=FUSE_CONTAINMENT, fusion_width=0,
                 window_qubits=0, shots=0, seed=0, targets=None, width=None, terms=(), budget_bytes=0,
                 statevector=False, max_pass_cost=0):
        keep = _CircuitKeep(circuit)
        req, hold = self._request(circuit, keep, precision=precision, fusion=fusion, fusion_width=fusion_width,
                                  window_qubits=window_qubits, shots=shots, seed=seed, targets=targets, width=width,
                                  terms=terms, budget_bytes=budget_bytes, statevector=statevector,
                                  record_pass_times=False, max_pass_cost=max_pass_cost)
        cost = CCost()
        check(lib().mh_engine_estimate(ctypes.byref(req), ctypes.byref(cost)))
        return cost

    def execute(self, circuit: W.Circuit, *, precision=C128, fusion=FUSE_CONTAINMENT, fusion_width=0,
                window_qubits=0, shots=0, seed=0, targets=None, width=None, terms=(), budget_bytes=0,
                statevector=True, record_pass_times=False, max_passes=4096, max_pass_cost=0,
                monitor=None, region_sampler=False) -> Result:
        """``monitor(boundaries_crossed)`` runs after every HBM pass; it
        abandons the run by raising (as the reference's at_safe_boundary
        does), which this call re-raises after the engine has freed the state.
        ``region_sampler``: the final state stays on the device as
        ``Result.region_sampler`` (a RegionSampler)."""
        keep = _CircuitKeep(circuit)
        req, hold = self._request(circuit, keep, precision=precision, fusion=fusion, fusion_width=fusion_width,
                                  window_qubits=window_qubits, shots=shots, seed=seed, targets=targets, width=width,
                                  terms=terms, budget_bytes=budget_bytes, statevector=statevector,
                                  record_pass_times=record_pass_times, max_pass_cost=max_pass_cost)
        raised = []
        if monitor is not None:
            def _boundary(_data, crossed):
                try:
                    monitor(int(crossed))
                    return 0
                except BaseException as e:  # handed back to the caller below
                    raised.append(e)
                    return 1
            req.monitor = MONITOR_FN(_boundary)
        req.want_region_sampler = int(bool(region_sampler))
        sv = np.zeros(1 << circuit.num_qubits, amp_dtype(precision)) if statevector else None
        outc = np.zeros(max(shots, 1), np.uint64)
        cnt = np.zeros(max(shots, 1), np.uint64)
        ex = np.zeros(2 * max(len(terms), 1), np.float64)
        pms = np.zeros(max_passes, np.float32)
        res = CResult(sv.ctypes.data_as(ctypes.c_void_p) if sv is not None else None,
                      outc.ctypes.data_as(_u64p), cnt.ctypes.data_as(_u64p), 0, ex.ctypes.data_as(_f64p),
                      pms.ctypes.data_as(_f32p), max_passes)
        status = lib().mh_engine_execute(ctypes.byref(req), ctypes.byref(res))
        if status == MH_ERR_ABANDONED and raised:
            raise raised[0]
        check(status)
        counts = {int(outc[i]): int(cnt[i]) for i in range(res.num_outcomes)} if shots else None
        return Result(statevector=sv, counts=counts,
                      expectation=ex.view(np.complex128)[:len(terms)].copy() if terms else None,
                      pass_ms=pms[:min(res.num_passes, max_passes)] if record_pass_times else None,
                      num_blocks=res.num_blocks, num_passes=res.num_passes, fused_blocks=res.fused_blocks,
                      gates_absorbed=res.gates_absorbed, actual_bytes=res.actual_bytes,
                      peak_device_bytes=res.peak_device_bytes, execute_seconds=res.execute_seconds,
                      gate_seconds=res.gate_seconds, sample_seconds=res.sample_seconds,
                      expectation_seconds=res.expectation_seconds, fusion_used=res.fusion_used,
                      fusion_width_used=res.fusion_width_used,
                      region_sampler=RegionSampler(res.region_sampler) if res.region_sampler else None)


class CTierPass(ctypes.Structure):
    _fields_ = [("global_mask", ctypes.c_uint64), ("first_block", ctypes.c_uint32), ("end_block", ctypes.c_uint32)]


class CTierRequest(ctypes.Structure):
    _fields_ = [("chunk_qubits", ctypes.c_uint32), ("window_qubits", ctypes.c_uint32),
                ("passes", ctypes.POINTER(CTierPass)), ("pass_capacity", ctypes.c_uint32),
                ("reserved0", ctypes.c_uint32)]


class CTierCost(ctypes.Structure):
    _fields_ = [("feasible", ctypes.c_int32), ("infeasible_reason", ctypes.c_char * 256),
                ("passes", ctypes.c_uint64), ("chunk_qubits", ctypes.c_uint32), ("window_qubits", ctypes.c_uint32),
                ("link_bytes", ctypes.c_uint64), ("host_state_bytes", ctypes.c_uint64),
                ("link_write_bytes", ctypes.c_uint64), ("global_blocks", ctypes.c_uint64),
                ("max_global_operands", ctypes.c_uint32), ("blocks", ctypes.c_uint32),
                ("device_bytes", ctypes.c_uint64), ("seconds", ctypes.c_double), ("fusion_used", ctypes.c_int32),
                ("fusion_width_used", ctypes.c_uint32)]


class HipHostTierEngine:
    """The {Dense, Hip, HostOffloaded} engine (mh_tier_*): the state in pinned
    host memory, the passes run on the device through two windows. Takes
    HipDenseEngine's arguments, plus the tier seams ``tier_chunk_qubits`` and
    ``tier_window_qubits`` (0: chosen); ``plan`` holds the pass plan."""

    name = "dense_hip_host_offloaded"

    @staticmethod
    def _tier(tier_chunk_qubits, tier_window_qubits, capacity):
        passes = (CTierPass * max(capacity, 1))()
        return CTierRequest(tier_chunk_qubits, tier_window_qubits, passes, capacity), passes

    @staticmethod
    def _plan(cost, passes):
        count = min(int(cost.passes), len(passes))
        return [(int(passes[i].global_mask), int(passes[i].first_block), int(passes[i].end_block))
                for i in range(count)]

    def estimate(self, circuit: W.Circuit, *, tier_chunk_qubits=0, tier_window_qubits=0, precision=C128,
                 fusion=FUSE_CONTAINMENT, fusion_width=0, shots=0, seed=0, targets=None, width=None, terms=(),
                 budget_bytes=0, statevector=False, max_passes=4096):
        keep = _CircuitKeep(circuit)
        req, hold = HipDenseEngine._request(circuit, keep, precision=precision, fusion=fusion,
                                            fusion_width=fusion_width, window_qubits=0, shots=shots, seed=seed,
                                            targets=targets, width=width, terms=terms, budget_bytes=budget_bytes,
                                            statevector=statevector, record_pass_times=False)
        tier, passes = self._tier(tier_chunk_qubits, tier_window_qubits, max_passes)
        cost = CTierCost()
        check(lib().mh_tier_estimate(ctypes.byref(req), ctypes.byref(tier), ctypes.byref(cost)))
        cost.plan = self._plan(cost, passes) if cost.feasible else []
        return cost

    def execute(self, circuit: W.Circuit, *, tier_chunk_qubits=0, tier_window_qubits=0, precision=C128,
                fusion=FUSE_CONTAINMENT, fusion_width=0, window_qubits=0, shots=0, seed=0, targets=None,
                width=None, terms=(), budget_bytes=0, statevector=True, record_pass_times=False,
                max_passes=4096, max_pass_cost=0, monitor=None) -> Result:
        """``monitor(passes_completed)`` runs at the end of every tier pass and
        abandons the run by raising, as in HipDenseEngine.execute."""
        keep = _CircuitKeep(circuit)
        req, hold = HipDenseEngine._request(circuit, keep, precision=precision, fusion=fusion,
                                            fusion_width=fusion_width, window_qubits=window_qubits, shots=shots,
                                            seed=seed, targets=targets, width=width, terms=terms,
                                            budget_bytes=budget_bytes, statevector=statevector,
                                            record_pass_times=record_pass_times, max_pass_cost=max_pass_cost)
        raised = []
        if monitor is not None:
            def _boundary(_data, crossed):
                try:
                    monitor(int(crossed))
                    return 0
                except BaseException as e:  # handed back to the caller below
                    raised.append(e)
                    return 1
            req.monitor = MONITOR_FN(_boundary)
        tier, passes = self._tier(tier_chunk_qubits, tier_window_qubits, max_passes)
        sv = np.zeros(1 << circuit.num_qubits, amp_dtype(precision)) if statevector else None
        outc = np.zeros(max(shots, 1), np.uint64)
        cnt = np.zeros(max(shots, 1), np.uint64)
        ex = np.zeros(2 * max(len(terms), 1), np.float64)
        pms = np.zeros(max_passes, np.float32)
        res = CResult(sv.ctypes.data_as(ctypes.c_void_p) if sv is not None else None,
                      outc.ctypes.data_as(_u64p), cnt.ctypes.data_as(_u64p), 0, ex.ctypes.data_as(_f64p),
                      pms.ctypes.data_as(_f32p), max_passes)
        cost = CTierCost()
        status = lib().mh_tier_execute(ctypes.byref(req), ctypes.byref(tier), ctypes.byref(res), ctypes.byref(cost))
        if status == MH_ERR_ABANDONED and raised:
            raise raised[0]
        check(status)
        counts = {int(outc[i]): int(cnt[i]) for i in range(res.num_outcomes)} if shots else None
        return Result(statevector=sv, counts=counts,
                      expectation=ex.view(np.complex128)[:len(terms)].copy() if terms else None,
                      pass_ms=pms[:min(res.num_passes, max_passes)] if record_pass_times else None,
                      num_blocks=res.num_blocks, num_passes=res.num_passes, actual_bytes=res.actual_bytes,
                      peak_device_bytes=res.peak_device_bytes, execute_seconds=res.execute_seconds,
                      gate_seconds=res.gate_seconds, sample_seconds=res.sample_seconds,
                      expectation_seconds=res.expectation_seconds, fusion_used=res.fusion_used,
                      fusion_width_used=res.fusion_width_used, cost=cost, plan=self._plan(cost, passes))


class RegionSampler:
    """A final state kept on the device with its cumulative-distribution tables
    (mh_region_sampler_*): the reference's CumulativeDistribution without the
    host copy."""

    def __init__(self, handle):
        self.ptr = ctypes.c_void_p(handle)
        n = ctypes.c_uint32()
        nbytes = ctypes.c_uint64()
        check(lib().mh_region_sampler_info(self.ptr, ctypes.byref(n), ctypes.byref(nbytes)))
        self.num_qubits = n.value
        self.device_bytes = nbytes.value

    def draw(self, uniforms) -> np.ndarray:
        """The basis index of each uniform in [0, 1), in the given order."""
        u = np.ascontiguousarray(uniforms, np.float64)
        out = np.zeros(max(u.size, 1), np.uint64)
        check(lib().mh_region_sampler_draw(self.ptr, _ptr(u, _f64p), u.size, out.ctypes.data_as(_u64p)))
        return out[:u.size]

    def close(self):
        if self.ptr:
            check(lib().mh_region_sampler_destroy(self.ptr))
            self.ptr = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class CBatchRequest(ctypes.Structure):
    _fields_ = [("num_circuits", ctypes.c_uint32), ("circuits", ctypes.POINTER(CCircuit)),
                ("precision", ctypes.c_int32), ("fusion", ctypes.c_int32), ("fusion_width", ctypes.c_uint32),
                ("global_phases", _f64p), ("term_offsets", _u32p), ("x_masks", _u64p), ("z_masks", _u64p),
                ("round_bytes", ctypes.c_uint64), ("host_threads", ctypes.c_uint32),
                ("want_statevectors", ctypes.c_int32), ("max_pass_cost", ctypes.c_uint32),
                ("group_bytes", ctypes.c_uint64), ("kernel_streams", ctypes.c_uint32),
                ("cached_states", ctypes.c_int32)]


class CBatchResult(ctypes.Structure):
    _fields_ = [("expectation_terms", _f64p), ("statevectors", ctypes.c_void_p), ("completion_seconds", _f64p),
                ("rounds", ctypes.c_uint32), ("launches", ctypes.c_uint32), ("plan_seconds", ctypes.c_double),
                ("device_seconds", ctypes.c_double), ("compile_seconds", ctypes.c_double),
                ("total_seconds", ctypes.c_double)]


def batch_execute(circuits, terms, *, precision=C128, fusion=FUSE_CONTAINMENT, fusion_width=0, round_bytes=0,
                  host_threads=0, statevectors=False, max_pass_cost=0, group_bytes=0, kernel_streams=0,
                  cached_states=False):
    """Runs every circuit and reduces its Pauli terms (a list of (x, z) masks per
    circuit) through mh_batch_execute. Returns a Result with `expectation` (one
    complex array per circuit), `completion_seconds`, optionally `statevectors`
    (one array per circuit) and the call's timings."""
    keeps = [_CircuitKeep(c) for c in circuits]
    arr = (CCircuit * max(len(keeps), 1))(*[k.c for k in keeps])
    offs = np.zeros(len(circuits) + 1, np.uint32)
    for i, t in enumerate(terms):
        offs[i + 1] = offs[i] + len(t)
    xs = np.array([x for t in terms for x, _ in t] or [0], np.uint64)
    zs = np.array([z for t in terms for _, z in t] or [0], np.uint64)
    out = np.zeros(2 * max(int(offs[-1]), 1), np.float64)
    done = np.zeros(max(len(circuits), 1), np.float64)
    sizes = [1 << c.num_qubits for c in circuits]
    sv = np.zeros(sum(sizes) if statevectors else 1, amp_dtype(precision))
    phases = np.array([c.global_phase for c in circuits] or [0.0], np.float64)
    req = CBatchRequest(len(circuits), arr, precision, fusion, fusion_width, _ptr(phases, _f64p), _ptr(offs, _u32p),
                        _ptr(xs, _u64p), _ptr(zs, _u64p), round_bytes, host_threads, 1 if statevectors else 0,
                        max_pass_cost, group_bytes, kernel_streams, 1 if cached_states else 0)
    res = CBatchResult(_ptr(out, _f64p), sv.ctypes.data_as(ctypes.c_void_p), _ptr(done, _f64p))
    check(lib().mh_batch_execute(ctypes.byref(req), ctypes.byref(res)))
    e = out[:2 * int(offs[-1])].view(np.complex128)
    r = Result(expectation=[e[offs[i]:offs[i + 1]].copy() for i in range(len(circuits))],
               completion_seconds=done[:len(circuits)].copy(), rounds=res.rounds, launches=res.launches,
               plan_seconds=res.plan_seconds, device_seconds=res.device_seconds, total_seconds=res.total_seconds,
               compile_seconds=res.compile_seconds)
    if statevectors:
        at, r.statevectors = 0, []
        for s in sizes:
            r.statevectors.append(sv[at:at + s].copy())
            at += s
    return r


class BatchGraph:
    """A resident batch replayed from a captured HIP graph (mh_batch_graph_*):
    create once for a list of circuits and their terms, update() with new
    parameters for the same circuits, run() to get the terms."""

    def __init__(self, circuits, terms, *, precision=C128, fusion=FUSE_CONTAINMENT, fusion_width=0, host_threads=0,
                 max_pass_cost=0, group_bytes=0, kernel_streams=0, cached_states=False):
        self._keeps = [_CircuitKeep(c) for c in circuits]
        self._arr = (CCircuit * max(len(self._keeps), 1))(*[k.c for k in self._keeps])
        self.offs = np.zeros(len(circuits) + 1, np.uint32)
        for i, t in enumerate(terms):
            self.offs[i + 1] = self.offs[i] + len(t)
        self._xs = np.array([x for t in terms for x, _ in t] or [0], np.uint64)
        self._zs = np.array([z for t in terms for _, z in t] or [0], np.uint64)
        self._phases = np.array([c.global_phase for c in circuits] or [0.0], np.float64)
        req = CBatchRequest(len(circuits), self._arr, precision, fusion, fusion_width, _ptr(self._phases, _f64p),
                            _ptr(self.offs, _u32p), _ptr(self._xs, _u64p), _ptr(self._zs, _u64p), 0, host_threads, 0,
                            max_pass_cost, group_bytes, kernel_streams, 1 if cached_states else 0)
        self.ptr = ctypes.c_void_p()
        check(lib().mh_batch_graph_create(ctypes.byref(req), ctypes.byref(self.ptr)))
        self.count = len(circuits)

    def update(self, circuits):
        """New parameters for the same circuits (same structure)."""
        keeps = [_CircuitKeep(c) for c in circuits]
        arr = (CCircuit * max(len(keeps), 1))(*[k.c for k in keeps])
        phases = np.array([c.global_phase for c in circuits] or [0.0], np.float64)
        check(lib().mh_batch_graph_update(self.ptr, arr, _ptr(phases, _f64p)))

    def param_count(self) -> int:
        n = ctypes.c_uint64()
        check(lib().mh_batch_graph_param_count(self.ptr, ctypes.byref(n)))
        return n.value

    def set_params(self, params, global_phases=None):
        """New parameters for every circuit, concatenated in circuit order."""
        p = np.ascontiguousarray(params, np.float64)
        ph = None if global_phases is None else np.ascontiguousarray(global_phases, np.float64)
        check(lib().mh_batch_graph_set_params(self.ptr, _ptr(p, _f64p) if p.size else None,
                                              _ptr(ph, _f64p) if ph is not None else None))

    def run(self):
        """Replays the graph: (one complex array of terms per circuit, CBatchResult)."""
        out = np.zeros(2 * max(int(self.offs[-1]), 1), np.float64)
        res = CBatchResult(_ptr(out, _f64p), None, None)
        check(lib().mh_batch_graph_run(self.ptr, ctypes.byref(res)))
        e = out[:2 * int(self.offs[-1])].view(np.complex128)
        return [e[self.offs[i]:self.offs[i + 1]].copy() for i in range(self.count)], res

    def close(self):
        if self.ptr:
            check(lib().mh_batch_graph_destroy(self.ptr))
            self.ptr = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class CServerOptions(ctypes.Structure):
    _fields_ = [("budget_bytes", ctypes.c_uint64), ("max_round_jobs", ctypes.c_uint32),
                ("host_threads", ctypes.c_uint32), ("fusion", ctypes.c_int32), ("fusion_width", ctypes.c_uint32),
                ("plan_cache_capacity", ctypes.c_uint32), ("reserve_bytes", ctypes.c_uint64),
                ("complex64_pairs", ctypes.c_int32), ("lanes", ctypes.c_uint32), ("reserved0", ctypes.c_uint32),
                ("large_job_bytes", ctypes.c_uint64)]


class CJob(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("precision", ctypes.c_int32), ("circuit", ctypes.POINTER(CCircuit)),
                ("global_phase", ctypes.c_double), ("num_observables", ctypes.c_uint32),
                ("observable_offsets", _u32p), ("x_masks", _u64p), ("z_masks", _u64p), ("coefficients", _f64p),
                ("shots", ctypes.c_uint64), ("seed", ctypes.c_uint64), ("num_targets", ctypes.c_uint32),
                ("target_qubits", _u32p), ("target_clbits", _u32p), ("bitstring_width", ctypes.c_uint32)]


class CJobResult(ctypes.Structure):
    _fields_ = [("observable_values", _f64p), ("raw_terms", _f64p), ("outcomes", _u64p), ("counts", _u64p),
                ("status", ctypes.c_int32), ("error", ctypes.c_char * 256), ("num_outcomes", ctypes.c_uint64),
                ("round", ctypes.c_uint32), ("plan_cache_hit", ctypes.c_int32),
                ("submitted_seconds", ctypes.c_double), ("started_seconds", ctypes.c_double),
                ("completed_seconds", ctypes.c_double), ("lane", ctypes.c_uint32)]


class CServerStats(ctypes.Structure):
    _fields_ = [("jobs_completed", ctypes.c_uint64), ("jobs_failed", ctypes.c_uint64), ("rounds", ctypes.c_uint32),
                ("plan_cache_hits", ctypes.c_uint64), ("plan_cache_misses", ctypes.c_uint64),
                ("peak_round_bytes", ctypes.c_uint64), ("budget_bytes", ctypes.c_uint64),
                ("busy_seconds", ctypes.c_double), ("plan_seconds", ctypes.c_double), ("arena_seconds", ctypes.c_double),
                ("compile_seconds", ctypes.c_double), ("pass_seconds", ctypes.c_double),
                ("expectation_seconds", ctypes.c_double), ("sample_seconds", ctypes.c_double),
                ("lanes", ctypes.c_uint32), ("peak_concurrent_rounds", ctypes.c_uint32),
                ("peak_held_bytes", ctypes.c_uint64)]


def derive_seeds(seed: int, count: int) -> list[int]:
    """The first `count` SplitMix64 outputs of `seed` (mh_derive_seeds)."""
    out = np.zeros(max(count, 1), np.uint64)
    check(lib().mh_derive_seeds(ctypes.c_uint64(seed), ctypes.c_uint32(count), _ptr(out, _u64p)))
    return [int(v) for v in out[:count]]


class Job:
    """A served job and the host arrays its C form points into.

    Estimator: ``observables`` is a list of observables, each a list of
    ``(x_mask, z_mask, coefficient)`` terms. Sampler: ``shots`` outcomes drawn
    with ``seed``; ``targets`` are (qubit, clbit) pairs (the circuit's final
    measurements when omitted) and ``width`` the bit-string width."""

    def __init__(self, kind, circuit, *, precision=C128, observables=(), shots=0, seed=0, targets=None,
                 width=None):
        self.kind, self.circuit, self.precision = kind, circuit, precision
        self._keep = _CircuitKeep(circuit)
        self.num_observables = len(observables)
        offs = [0]
        xs, zs, cs = [], [], []
        for obs in observables:
            for x, z, c in obs:
                xs.append(x)
                zs.append(z)
                cs.extend((complex(c).real, complex(c).imag))
            offs.append(len(xs))
        self.num_terms = len(xs)
        self._offs = np.array(offs, np.uint32)
        self._x = np.array(xs or [0], np.uint64)
        self._z = np.array(zs or [0], np.uint64)
        self._c = np.array(cs or [0.0, 0.0], np.float64)
        if targets is None:
            targets = circuit.measure_targets()
        if width is None:
            width = max([c for _, c in targets], default=-1) + 1
        self.shots = shots
        self._tq = np.array([q for q, _ in targets] or [0], np.uint32)
        self._tc = np.array([c for _, c in targets] or [0], np.uint32)
        self.c = CJob(kind, precision, ctypes.pointer(self._keep.c), circuit.global_phase, self.num_observables,
                      _ptr(self._offs, _u32p), _ptr(self._x, _u64p), _ptr(self._z, _u64p), _ptr(self._c, _f64p),
                      shots, seed, len(targets), _ptr(self._tq, _u32p), _ptr(self._tc, _u32p), width)

    def result_buffers(self):
        vals = np.zeros(2 * max(self.num_observables, 1), np.float64)
        raw = np.zeros(2 * max(self.num_terms, 1), np.float64)
        outc = np.zeros(max(self.shots, 1), np.uint64)
        cnts = np.zeros(max(self.shots, 1), np.uint64)
        r = CJobResult(_ptr(vals, _f64p), _ptr(raw, _f64p), _ptr(outc, _u64p), _ptr(cnts, _u64p))
        return r, (vals, raw, outc, cnts)

    def collect(self, r: CJobResult, bufs) -> Result:
        vals, raw, outc, cnts = bufs
        out = Result(status=r.status, error=r.error.decode(), round=r.round, lane=r.lane,
                     plan_cache_hit=bool(r.plan_cache_hit),
                     submitted=r.submitted_seconds, started=r.started_seconds, completed=r.completed_seconds,
                     latency=r.completed_seconds - r.submitted_seconds)
        if r.status == MH_OK:
            out.values = vals[:2 * self.num_observables].view(np.complex128).copy()
            out.raw_terms = raw[:2 * self.num_terms].view(np.complex128).copy()
            m = int(r.num_outcomes)
            out.counts = {int(k): int(v) for k, v in zip(outc[:m], cnts[:m])} if self.kind == JOB_SAMPLER else None
        return out


class Server:
    """A resident job server (mh_server_*): submit jobs, wait for results.
    ``large_job_bytes``: jobs whose states reach it run one at a time in rounds
    of their own while smaller jobs pass them (0: arrival order only)."""

    def __init__(self, *, budget_bytes=0, max_round_jobs=0, host_threads=0, fusion=FUSE_CONTAINMENT,
                 fusion_width=0, plan_cache_capacity=0, reserve_bytes=0, complex64_pairs=False, lanes=0,
                 large_job_bytes=0):
        o = CServerOptions(budget_bytes, max_round_jobs, host_threads, fusion, fusion_width, plan_cache_capacity,
                           reserve_bytes, int(bool(complex64_pairs)), lanes, 0, large_job_bytes)
        self._h = ctypes.c_void_p()
        check(lib().mh_server_create(ctypes.byref(o), ctypes.byref(self._h)))
        self._pending = {}

    def submit(self, job: Job) -> int:
        t = ctypes.c_uint64()
        check(lib().mh_server_submit(self._h, ctypes.byref(job.c), ctypes.byref(t)))
        self._pending[t.value] = job
        return t.value

    def wait(self, ticket: int) -> Result:
        job = self._pending.pop(ticket)
        r, bufs = job.result_buffers()
        check(lib().mh_server_wait(self._h, ctypes.c_uint64(ticket), ctypes.byref(r)))
        return job.collect(r, bufs)

    def replay(self, jobs, arrivals) -> list:
        """Submits jobs[i] at arrivals[i] seconds (from the call's start) and
        waits for all of them; returns their results in order."""
        n = len(jobs)
        arr = (CJob * max(n, 1))(*[j.c for j in jobs])
        res = (CJobResult * max(n, 1))()
        bufs = []
        for i, j in enumerate(jobs):
            r, b = j.result_buffers()
            res[i] = r
            bufs.append(b)
        t = np.ascontiguousarray(arrivals, np.float64)
        check(lib().mh_server_replay(self._h, ctypes.c_uint32(n), arr, _ptr(t if n else np.zeros(1), _f64p), res))
        return [j.collect(res[i], bufs[i]) for i, j in enumerate(jobs)]

    def stats(self) -> CServerStats:
        s = CServerStats()
        check(lib().mh_server_stats_get(self._h, ctypes.byref(s)))
        return s

    def close(self):
        if self._h:
            check(lib().mh_server_destroy(self._h))
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def calibrate() -> Calibration:
    c = Calibration()
    check(lib().mh_calibrate(ctypes.byref(c)))
    return c


def set_calibration(c: Calibration) -> None:
    check(lib().mh_set_calibration(ctypes.byref(c)))


def get_calibration() -> Calibration:
    c = Calibration()
    check(lib().mh_get_calibration(ctypes.byref(c)))
    return c
