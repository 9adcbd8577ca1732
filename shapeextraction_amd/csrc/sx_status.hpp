// Status handling shared by every translation unit of the library (host only,
// no HIP dependency so that plain C++ sources can include it).
//
// The reference prints errors and continues (src/CudaCommon.hpp:19-76); this
// library returns a status code from every C entry point and keeps a
// thread-local message for sx_last_error().
#pragma once

#include <string>

#include "../../include/shapes_c.h"

namespace sx {

inline thread_local std::string g_last_error;

inline int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

}  // namespace sx
