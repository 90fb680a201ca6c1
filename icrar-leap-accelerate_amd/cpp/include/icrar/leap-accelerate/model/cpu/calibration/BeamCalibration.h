// cpu::BeamCalibration (src/icrar/leap-accelerate/model/cpu/calibration/BeamCalibration.h:45-124):
// the phase calibration of every antenna for one direction, with the same
// JSON form {"direction":[ra,dec],"beamCalibration":[...]}.
#pragma once

#include <vector>

#include "icrar/leap-accelerate/common/SphericalDirection.h"
#include "icrar/leap-accelerate/model/cpu/calibration/json_writer.h"

namespace icrar {
namespace cpu {

class BeamCalibration {
    SphericalDirection m_direction;
    std::vector<double> m_calibration;  // one value per antenna column of A

public:
    BeamCalibration(SphericalDirection direction, std::vector<double> calibration)
        : m_direction(direction), m_calibration(std::move(calibration)) {}

    const SphericalDirection& GetDirection() const { return m_direction; }
    const std::vector<double>& GetPhaseCalibration() const { return m_calibration; }

    void Write(JsonWriter& w) const {
        w.StartObject();
        w.Key("direction");
        w.StartArray();
        for (double v : m_direction) w.Double(v);
        w.EndArray();
        w.Key("beamCalibration");
        w.StartArray();
        for (double v : m_calibration) w.Double(v);
        w.EndArray();
        w.EndObject();
    }
};

}  // namespace cpu
}  // namespace icrar
