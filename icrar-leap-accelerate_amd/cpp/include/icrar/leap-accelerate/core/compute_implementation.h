// icrar::ComputeImplementation (src/icrar/leap-accelerate/core/compute_implementation.h:33-37)
// extended with hip; the string forms follow compute_implementation.cc:31-56.
#pragma once

#include <string>

namespace icrar {

enum class ComputeImplementation {
    cpu,   // the reference's Eigen implementation
    cuda,  // the reference's CUDA implementation
    hip,   // this library: gfx950 HIP kernels
};

inline bool TryParseComputeImplementation(const std::string& value, ComputeImplementation& out) {
    if (value == "cpu" || value == "casa" || value == "eigen") {
        out = ComputeImplementation::cpu;
        return true;
    }
    if (value == "cuda") {
        out = ComputeImplementation::cuda;
        return true;
    }
    if (value == "hip") {
        out = ComputeImplementation::hip;
        return true;
    }
    return false;
}

}  // namespace icrar
