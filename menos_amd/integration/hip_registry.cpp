// Registry entry and dense backend choice for the HIP engine (see hip_registry.h).
#include "hip_registry.h"

#include <algorithm>

#include "common/error.h"

namespace menos::engines::dense_hip {

std::vector<const SimulationEngine*> with_hip_engine(std::vector<const SimulationEngine*> engines) {
    // Function-local static and runtime-gated, as the CUDA entry
    // (execution_engines.cpp:22-30): a HIP-enabled binary on a machine with no
    // device behaves like one without HIP.
    static const HipDenseEngine kHipDense;
    if (hip_device_available()) engines.push_back(&kHipDense);
    return engines;
}

const SimulationEngine* find_engine(const std::vector<const SimulationEngine*>& engines, Engine engine,
                                    Backend backend, Tier tier) {
    for (const SimulationEngine* e : engines)
        if (e->engine_id() == engine && e->backend() == backend && e->tier() == tier) return e;
    return nullptr;
}

DenseBackendChoice choose_dense_backend(const DenseBackendInputs& in) {
    DenseBackendChoice out;
    const double on_cpu = std::max(in.cpu_seconds, in.observed_cpu_seconds);
    out.seconds = on_cpu;
    out.feasible = in.cpu_feasible;
    if (in.requested == BackendRequest::Hip) {
        if (!in.hip_engine_registered)
            throw common::UnsupportedFeature("backend=hip was requested, but this build of MENOS has no HIP support");
        if (!in.hip.present)
            throw common::UnsupportedFeature("backend=hip was requested, but no usable HIP device is available");
        if (in.noise_reason.has_value())
            throw common::UnsupportedFeature(
                "backend=hip was requested, but the circuit is noisy and the HIP backend runs noiseless dense "
                "simulation only");
        out.backend = kBackendHip;
        out.seconds = std::max(in.hip_seconds, in.observed_hip_seconds);
        if (!in.hip_estimate.feasible) {
            out.feasible = false;
            out.infeasible_reason = "backend=hip requested: " + in.hip_estimate.infeasible_reason;
        } else if (in.hip_estimate.device.device_bytes > in.hip.device_available_bytes) {
            out.feasible = false;
            out.infeasible_reason = "backend=hip requested: the state needs " +
                                    std::to_string(in.hip_estimate.device.device_bytes) +
                                    " bytes of device memory, more than the " +
                                    std::to_string(in.hip.device_available_bytes) + " available";
        } else {
            out.feasible = true;
        }
        return out;
    }
    if (in.requested == BackendRequest::Cpu || !in.hip_engine_registered || !in.hip.present ||
        !in.hip.calibrated || !in.cpu_calibrated)
        return out;
    const bool fits = in.hip_estimate.device.device_bytes <= in.hip.device_available_bytes;
    if (!(in.hip_estimate.feasible && in.cpu_feasible && in.cpu_tier == Tier::InMemory && fits)) return out;
    const double on_hip = std::max(in.hip_seconds, in.observed_hip_seconds);
    if (on_hip < on_cpu) {  // strictly faster: ties go to the CPU
        out.backend = kBackendHip;
        out.seconds = on_hip;
    }
    return out;
}

}  // namespace menos::engines::dense_hip
