// cpu::CalibrationCollection (src/icrar/leap-accelerate/model/cpu/calibration/CalibrationCollection.h:41-82):
// the Calibrate result, serialised as a JSON array of calibrations
// (schema CalibrationCollection.json).
#pragma once

#include <ostream>
#include <vector>

#include "icrar/leap-accelerate/model/cpu/calibration/Calibration.h"

namespace icrar {
namespace cpu {

class CalibrationCollection {
    std::vector<Calibration> m_calibrations;

public:
    CalibrationCollection() = default;
    explicit CalibrationCollection(std::vector<Calibration> calibrations) : m_calibrations(std::move(calibrations)) {}

    const std::vector<Calibration>& GetCalibrations() const { return m_calibrations; }

    void Write(JsonWriter& w) const {
        w.StartArray();
        for (const auto& c : m_calibrations) c.Write(w);
        w.EndArray();
    }

    void Serialize(std::ostream& os) const {
        JsonWriter w;
        Write(w);
        os << w.str() << std::endl;
    }
};

}  // namespace cpu
}  // namespace icrar
