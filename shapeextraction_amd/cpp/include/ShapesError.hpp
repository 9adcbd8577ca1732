// Error handling of the drop-in C++ API.
//
// The reference prints CUDA/cuSOLVER errors and continues
// (src/CudaCommon.hpp:19-76).  This API throws shapes::Error, carrying the C
// ABI status and message, on every failure instead.
#pragma once

#include <stdexcept>
#include <string>

#include "shapes_c.h"

namespace shapes {

class Error : public std::runtime_error {
public:
    Error(int status, const std::string& what) : std::runtime_error(what), status_(status) {}
    int status() const { return status_; }

private:
    int status_;
};

inline void check(int status, const char* what) {
    if (status != SX_OK) throw Error(status, std::string(what) + ": " + sx_last_error());
}

}  // namespace shapes
