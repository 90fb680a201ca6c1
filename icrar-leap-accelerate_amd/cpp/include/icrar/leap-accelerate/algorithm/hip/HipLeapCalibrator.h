// icrar::hip::HipLeapCalibrator: the ILeapCalibrator of the gfx950 library.
//
// It drops in beside cpu::CpuLeapCalibrator (algorithm/cpu/CpuLeapCalibrator.h:49-88)
// and replaces cuda::CudaLeapCalibrator (algorithm/cuda/CudaLeapCalibrator.h:72-161):
// it owns a device context (stream, workspace, pseudo-inverse cache) for its
// lifetime, and every call goes through the C ABI of include/leap_hip.h.
#pragma once

#include <functional>
#include <optional>
#include <vector>

#include "icrar/leap-accelerate/algorithm/ILeapCalibrator.h"
#include "leap_hip.h"

namespace icrar {
namespace hip {

class HipLeapCalibrator : public ILeapCalibrator {
    leap_ctx* m_ctx = nullptr;
    leap_options m_options{};

public:
    // device: HIP device ordinal.  options: precision / strategy (defaults to
    // the reference's arithmetic, FP64 one direction at a time).
    explicit HipLeapCalibrator(int device = 0, const leap_options* options = nullptr);
    ~HipLeapCalibrator() override;

    cpu::CalibrationCollection Calibrate(const icrar::MeasurementSet& ms,
                                         const std::vector<SphericalDirection>& directions,
                                         const Slice& solutionInterval, double minimumBaselineThreshold,
                                         std::optional<unsigned int> referenceAntenna,
                                         bool isFileSystemCacheEnabled) override;

    // Streaming form: each solution interval's calibration is handed to
    // `output` as soon as it is computed.
    void Calibrate(const icrar::MeasurementSet& ms, const std::vector<SphericalDirection>& directions,
                   const Slice& solutionInterval, double minimumBaselineThreshold,
                   std::optional<unsigned int> referenceAntenna, bool isFileSystemCacheEnabled,
                   const std::function<void(cpu::Calibration)>& output);

    const leap_options& GetOptions() const { return m_options; }
    void SetOptions(const leap_options& o) { m_options = o; }
    leap_ctx* GetContext() const { return m_ctx; }
};

}  // namespace hip
}  // namespace icrar
