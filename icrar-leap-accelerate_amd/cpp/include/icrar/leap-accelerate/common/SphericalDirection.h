// icrar::SphericalDirection: (right ascension, declination) in radians
// (src/icrar/leap-accelerate/common/SphericalDirection.h:32, an Eigen::Vector2d
// there; element access by operator() is kept).
#pragma once

#include <cstddef>

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

}  // namespace icrar
