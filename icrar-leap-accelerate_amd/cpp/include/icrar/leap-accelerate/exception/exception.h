// Exceptions of the C++ API: the icrar::exception hierarchy of
// src/icrar/leap-accelerate/exception/exception.h:35-92 (same class names,
// constructor arguments and message layout "file:line message").
#pragma once

#include <exception>
#include <sstream>
#include <string>

namespace icrar {

class exception : public std::exception {
    std::string m_message;

public:
    exception(const std::string& msg, const std::string& file, int line) {
        std::stringstream ss;
        ss << file << ":" << line << " " << msg << std::endl;
        m_message = ss.str();
    }
    const char* what() const noexcept override { return m_message.c_str(); }
};

class invalid_argument_exception : public icrar::exception {
public:
    invalid_argument_exception(const std::string& msg, const std::string& arg, const std::string& file, int line)
        : exception(arg + " invalid argument exception: " + msg, file, line) {}
};

class file_exception : public icrar::exception {
public:
    file_exception(const std::string& msg, const std::string& filename, const std::string& file, int line)
        : exception(filename + " file exception: " + msg, file, line) {}
};

class json_exception : public icrar::exception {
public:
    json_exception(const std::string& msg, const std::string& file, int line)
        : exception("json exception: " + msg, file, line) {}
};

class not_implemented_exception : public icrar::exception {
public:
    not_implemented_exception(const std::string& file, int line) : exception("not implemented", file, line) {}
};

// Raised for failures of the device runtime; the reference terminates the
// process instead (cuda/helper_cuda.cuh:613-624).
class device_exception : public icrar::exception {
public:
    device_exception(const std::string& msg, const std::string& file, int line)
        : exception("device exception: " + msg, file, line) {}
};

}  // namespace icrar
