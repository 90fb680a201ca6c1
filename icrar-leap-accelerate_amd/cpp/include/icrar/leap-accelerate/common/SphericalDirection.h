// icrar::SphericalDirection: (right ascension, declination) in radians
// (src/icrar/leap-accelerate/common/SphericalDirection.h:32, an Eigen::Vector2d
// there; element access by operator() is kept).
#pragma once

#include <cstddef>
#include <string>
#include <vector>

#include "icrar/leap-accelerate/common/json_reader.h"
#include "icrar/leap-accelerate/exception/exception.h"

namespace icrar {

struct SphericalDirection {
    double v[2] = {0.0, 0.0};
    SphericalDirection() = default;
    SphericalDirection(double ra, double dec) : v{ra, dec} {}
    double& operator()(std::size_t i) { return v[i]; }
    double operator()(std::size_t i) const { return v[i]; }
    double& operator[](std::size_t i) { return v[i]; }
    double operator[](std::size_t i) const { return v[i]; }
    const double* begin() const { return v; }
    const double* end() const { return v + 2; }
    bool operator==(const SphericalDirection& o) const { return v[0] == o.v[0] && v[1] == o.v[1]; }
};

// ParseDirections (common/SphericalDirection.cc:31-63): [[ra, dec], ...] in radians.
inline std::vector<SphericalDirection> ParseDirections(const JsonValue& doc) {
    if (doc.kind != JsonValue::Array) throw icrar::exception("expected an array", __FILE__, __LINE__);
    std::vector<SphericalDirection> result;
    for (const auto& e : doc.arr) {
        if (e.kind != JsonValue::Array || e.arr.size() != 2 || !e.arr[0].IsNumber() || !e.arr[1].IsNumber())
            throw icrar::exception("expected an array of 2 numbers", __FILE__, __LINE__);
        result.emplace_back(e.arr[0].num, e.arr[1].num);
    }
    return result;
}

inline std::vector<SphericalDirection> ParseDirections(const std::string& json) {
    return ParseDirections(JsonReader::Parse(json));
}

}  // namespace icrar
