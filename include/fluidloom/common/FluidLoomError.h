// Error types of the C++ API (same names and hierarchy as the reference's
// include/fluidloom/common/FluidLoomError.h:14-81; HipError takes the place
// of OpenCLError and carries the C-ABI status).
#pragma once
#include <sstream>
#include <stdexcept>
#include <string>

namespace fluidloom {

class FluidLoomError : public std::runtime_error {
public:
    FluidLoomError(const std::string& message, const std::string& file, int line)
        : std::runtime_error(format(message, file, line)), m_file(file), m_line(line) {}
    const std::string& getFile() const { return m_file; }
    int getLine() const { return m_line; }

private:
    static std::string format(const std::string& m, const std::string& f, int l) {
        std::ostringstream o;
        o << "[FluidLoomError] " << m << " at " << f << ":" << l;
        return o.str();
    }
    std::string m_file;
    int m_line;
};

class BackendError : public FluidLoomError {
public:
    BackendError(const std::string& msg, const std::string& file, int line)
        : FluidLoomError("[Backend] " + msg, file, line) {}
};

// A failed fl_* call (status from fluidloom_hip.h plus its message).
class HipError : public BackendError {
public:
    HipError(int status, const std::string& msg, const std::string& file, int line)
        : BackendError(msg + " (status " + std::to_string(status) + ")", file, line), m_status(status) {}
    int getErrorCode() const { return m_status; }

private:
    int m_status;
};

}  // namespace fluidloom

#define FL_THROW(ExceptionType, msg) throw ExceptionType((msg), __FILE__, __LINE__)
// Converts a non-zero fl_status into HipError with fl_last_error_string().
#define FL_CHECK_STATUS(call)                                                              \
    do {                                                                                   \
        int fl_st_ = (call);                                                               \
        if (fl_st_ != 0) throw ::fluidloom::HipError(fl_st_, fl_last_error_string(), __FILE__, __LINE__); \
    } while (0)
