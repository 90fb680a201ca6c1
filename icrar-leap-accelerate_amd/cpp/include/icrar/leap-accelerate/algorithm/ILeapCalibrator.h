// icrar::ILeapCalibrator (src/icrar/leap-accelerate/algorithm/ILeapCalibrator.h:46-69).
// The Calibrate signature is the reference's with std::optional in place of
// boost::optional.
#pragma once

#include <optional>
#include <vector>

#include "icrar/leap-accelerate/common/Slice.h"
#include "icrar/leap-accelerate/common/SphericalDirection.h"
#include "icrar/leap-accelerate/model/cpu/calibration/CalibrationCollection.h"

namespace icrar {

class MeasurementSet;

class ILeapCalibrator {
public:
    ILeapCalibrator() = default;
    ILeapCalibrator(const ILeapCalibrator&) = delete;
    ILeapCalibrator& operator=(const ILeapCalibrator&) = delete;
    virtual ~ILeapCalibrator() = default;

    /**
     * Performs LEAP calibration of the given directions over the solution
     * intervals of `solutionInterval`.
     */
    virtual cpu::CalibrationCollection Calibrate(const icrar::MeasurementSet& ms,
                                                 const std::vector<SphericalDirection>& directions,
                                                 const Slice& solutionInterval, double minimumBaselineThreshold,
                                                 std::optional<unsigned int> referenceAntenna,
                                                 bool isFileSystemCacheEnabled) = 0;
};

}  // namespace icrar
