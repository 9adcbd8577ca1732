// Morton codes (src/MortonUtils.hpp:9-20 of the reference), host side.
//
// pointToCode / codeToPoint compute exactly what the GPU encode and decode
// kernels compute (same float operation order, no contraction): x in bits 3i,
// y in 3i+1, z in 3i+2 of a CODE_LEN-bit code.
#pragma once

#include <cstdint>

#include "PointCloud.hpp"

typedef uint64_t Code_t;
constexpr int CODE_LEN = 63;

Code_t pointToCode(const float x, const float y, const float z, const float min_coord, const float range);
Point codeToPoint(const Code_t code, const float min_coord, const float range);
