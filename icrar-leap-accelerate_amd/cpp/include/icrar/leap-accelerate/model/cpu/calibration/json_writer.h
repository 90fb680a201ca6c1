// Compact JSON writer for the calibration results.
//
// The reference serialises with rapidjson::Writer (BeamCalibration.h:84-123),
// whose Double() prints the shortest digit string that reads back to the same
// double and places the decimal point as follows, with value = digits * 10^k
// and kk = number of digits + k:
//   0 <= k and kk <= 21  ->  digits, k zeros, ".0"        (5020320156.0)
//   0 < kk <= 21         ->  point after kk digits         (2.58353835690993)
//   -6 < kk <= 0         ->  "0." then -kk zeros, digits    (0.000125)
//   one digit            ->  "de<kk-1>"                     (1e-7)
//   otherwise            ->  "d.ddd" "e<kk-1>"              (6.25039512024925e-15)
// with no '+' in exponents.  The digits here come from std::to_chars (shortest
// round trip).
#pragma once

#include <charconv>
#include <cmath>
#include <cstdlib>
#include <string>
#include <vector>

namespace icrar {

inline std::string FormatJsonDouble(double x) {
    if (std::isnan(x)) return "NaN";
    if (std::isinf(x)) return x > 0 ? "Infinity" : "-Infinity";
    if (x == 0.0) return std::signbit(x) ? "-0.0" : "0.0";
    char buf[64];
    auto res = std::to_chars(buf, buf + sizeof(buf), x, std::chars_format::scientific);
    std::string s(buf, res.ptr);
    std::string out;
    if (s[0] == '-') {
        out = "-";
        s.erase(0, 1);
    }
    const size_t epos = s.find('e');
    const int e = std::atoi(s.c_str() + epos + 1);
    std::string digits;
    for (size_t i = 0; i < epos; ++i)
        if (s[i] != '.') digits += s[i];
    const int n = int(digits.size());
    const int k = e - (n - 1);
    const int kk = n + k;
    if (k >= 0 && kk <= 21) {
        out += digits + std::string(size_t(k), '0') + ".0";
    } else if (kk > 0 && kk <= 21) {
        out += digits.substr(0, size_t(kk)) + "." + digits.substr(size_t(kk));
    } else if (kk > -6 && kk <= 0) {
        out += "0." + std::string(size_t(-kk), '0') + digits;
    } else if (n == 1) {
        out += digits + "e" + std::to_string(kk - 1);
    } else {
        out += digits.substr(0, 1) + "." + digits.substr(1) + "e" + std::to_string(kk - 1);
    }
    return out;
}

class JsonWriter {
    std::string m_s;
    std::vector<bool> m_first;  // per open container: no element written yet
    bool m_after_key = false;

    void value_prefix() {
        if (m_after_key) {
            m_after_key = false;
            return;
        }
        if (!m_first.empty()) {
            if (!m_first.back()) m_s += ',';
            m_first.back() = false;
        }
    }

public:
    void StartObject() {
        value_prefix();
        m_s += '{';
        m_first.push_back(true);
    }
    void EndObject() {
        m_s += '}';
        m_first.pop_back();
    }
    void StartArray() {
        value_prefix();
        m_s += '[';
        m_first.push_back(true);
    }
    void EndArray() {
        m_s += ']';
        m_first.pop_back();
    }
    void Key(const std::string& k) {
        value_prefix();
        m_s += '"' + k + "\":";
        m_after_key = true;
    }
    void Double(double v) {
        value_prefix();
        m_s += FormatJsonDouble(v);
    }
    const std::string& str() const { return m_s; }
};

}  // namespace icrar
