// cpu::Calibration (src/icrar/leap-accelerate/model/cpu/calibration/Calibration.h:42-111):
// one solution interval, JSON {"epoch":{"start":..,"end":..},"calibration":[...]}.
#pragma once

#include <ostream>
#include <vector>

#include "icrar/leap-accelerate/model/cpu/calibration/BeamCalibration.h"

namespace icrar {
namespace cpu {

class Calibration {
    double m_startEpoch;
    double m_endEpoch;
    std::vector<BeamCalibration> m_beamCalibrations;

public:
    Calibration(double startEpoch, double endEpoch) : m_startEpoch(startEpoch), m_endEpoch(endEpoch) {}
    Calibration(double startEpoch, double endEpoch, std::vector<BeamCalibration> beams)
        : m_startEpoch(startEpoch), m_endEpoch(endEpoch), m_beamCalibrations(std::move(beams)) {}

    double GetStartEpoch() const { return m_startEpoch; }
    double GetEndEpoch() const { return m_endEpoch; }
    const std::vector<BeamCalibration>& GetBeamCalibrations() const { return m_beamCalibrations; }
    std::vector<BeamCalibration>& GetBeamCalibrations() { return m_beamCalibrations; }

    void Write(JsonWriter& w) const {
        w.StartObject();
        w.Key("epoch");
        w.StartObject();
        w.Key("start");
        w.Double(m_startEpoch);
        w.Key("end");
        w.Double(m_endEpoch);
        w.EndObject();
        w.Key("calibration");
        w.StartArray();
        for (const auto& b : m_beamCalibrations) b.Write(w);
        w.EndArray();
        w.EndObject();
    }

    void Serialize(std::ostream& os) const {
        JsonWriter w;
        Write(w);
        os << w.str() << std::endl;
    }
};

}  // namespace cpu
}  // namespace icrar
