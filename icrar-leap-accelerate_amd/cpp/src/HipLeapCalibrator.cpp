// icrar::hip::HipLeapCalibrator over the C ABI (no HIP headers needed here).
#include "icrar/leap-accelerate/algorithm/hip/HipLeapCalibrator.h"

#include <string>

#include "icrar/leap-accelerate/exception/exception.h"
#include "icrar/leap-accelerate/ms/MeasurementSet.h"

namespace icrar {
namespace hip {
namespace {

[[noreturn]] void raise(int status, leap_ctx* ctx) {
    const std::string msg = leap_hip_last_error(ctx);
    switch (status) {
        case LEAP_ERR_INVALID_ARGUMENT:
            throw invalid_argument_exception(msg, "calibrate", __FILE__, __LINE__);
        case LEAP_ERR_NOT_IMPLEMENTED:
            throw not_implemented_exception(__FILE__, __LINE__);
        case LEAP_ERR_HIP:
        case LEAP_ERR_OUT_OF_MEMORY:
            throw device_exception(msg, __FILE__, __LINE__);
        default:
            throw icrar::exception(msg, __FILE__, __LINE__);
    }
}

struct Sink {
    const std::vector<SphericalDirection>* directions;
    const std::function<void(cpu::Calibration)>* output;
    std::exception_ptr error;
};

int on_solution(void* user, const leap_solution* s) {
    Sink* sink = static_cast<Sink*>(user);
    try {
        cpu::Calibration cal(s->start_epoch, s->end_epoch);
        for (uint32_t d = 0; d < s->n_directions; ++d) {
            const double* c = s->calibrations + size_t(d) * s->n_antennas;
            cal.GetBeamCalibrations().emplace_back((*sink->directions)[d],
                                                   std::vector<double>(c, c + s->n_antennas));
        }
        (*sink->output)(std::move(cal));
        return 0;
    } catch (...) {
        sink->error = std::current_exception();
        return 1;
    }
}

}  // namespace

HipLeapCalibrator::HipLeapCalibrator(int device, const leap_options* options) {
    if (options) m_options = *options;
    const int rc = leap_hip_create(&m_ctx, device);
    if (rc != LEAP_OK) raise(rc, nullptr);
}

HipLeapCalibrator::~HipLeapCalibrator() { leap_hip_destroy(m_ctx); }

void HipLeapCalibrator::Calibrate(const icrar::MeasurementSet& ms, const std::vector<SphericalDirection>& directions,
                                  const Slice& solutionInterval, double minimumBaselineThreshold,
                                  std::optional<unsigned int> referenceAntenna, bool isFileSystemCacheEnabled,
                                  const std::function<void(cpu::Calibration)>& output) {
    const leap_ms_view view = ms.View();
    std::vector<double> dirs;
    dirs.reserve(directions.size() * 2);
    for (const auto& d : directions) {
        dirs.push_back(d(0));
        dirs.push_back(d(1));
    }
    Sink sink{&directions, &output, nullptr};
    const int rc = leap_hip_calibrate(m_ctx, &view, dirs.data(), uint32_t(directions.size()), solutionInterval.start,
                                      solutionInterval.interval, solutionInterval.end, minimumBaselineThreshold,
                                      referenceAntenna ? int32_t(*referenceAntenna) : -1,
                                      isFileSystemCacheEnabled ? 1 : 0, &m_options, &on_solution, &sink);
    if (sink.error) std::rethrow_exception(sink.error);
    if (rc != LEAP_OK) raise(rc, m_ctx);
}

cpu::CalibrationCollection HipLeapCalibrator::Calibrate(const icrar::MeasurementSet& ms,
                                                        const std::vector<SphericalDirection>& directions,
                                                        const Slice& solutionInterval,
                                                        double minimumBaselineThreshold,
                                                        std::optional<unsigned int> referenceAntenna,
                                                        bool isFileSystemCacheEnabled) {
    std::vector<cpu::Calibration> out;
    Calibrate(ms, directions, solutionInterval, minimumBaselineThreshold, referenceAntenna, isFileSystemCacheEnabled,
              [&out](cpu::Calibration c) { out.push_back(std::move(c)); });
    return cpu::CalibrationCollection(std::move(out));
}

}  // namespace hip
}  // namespace icrar
