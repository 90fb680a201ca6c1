// icrar::LeapCalibratorFactory (src/icrar/leap-accelerate/algorithm/LeapCalibratorFactory.cc:30-48)
// with the hip implementation.  The reference's cpu and cuda calibrators are not
// part of this library; asking for them raises the reference's
// invalid_argument_exception, as the reference does for cuda without CUDA_ENABLED.
#pragma once

#include <memory>

#include "icrar/leap-accelerate/algorithm/ILeapCalibrator.h"
#include "icrar/leap-accelerate/algorithm/hip/HipLeapCalibrator.h"
#include "icrar/leap-accelerate/core/compute_implementation.h"
#include "icrar/leap-accelerate/exception/exception.h"

namespace icrar {

class LeapCalibratorFactory {
public:
    static std::unique_ptr<ILeapCalibrator> Create(ComputeImplementation impl) {
        if (impl == ComputeImplementation::hip) return std::make_unique<hip::HipLeapCalibrator>();
        if (impl == ComputeImplementation::cpu)
            throw invalid_argument_exception("cpu implementation is not part of this library", "impl", __FILE__,
                                             __LINE__);
        if (impl == ComputeImplementation::cuda)
            throw invalid_argument_exception("cuda build option not enabled", "impl", __FILE__, __LINE__);
        throw invalid_argument_exception("invalid argument", "impl", __FILE__, __LINE__);
    }
};

}  // namespace icrar
