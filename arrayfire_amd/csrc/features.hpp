// The object behind an af_features handle (src/api/c/features.cpp): n and
// the five f32 arrays it owns.
#pragma once
#include "../../include/af/defines.h"

namespace afhip {

struct Features {
    dim_t n = 0;
    af_array x = nullptr, y = nullptr, score = nullptr, orientation = nullptr, size = nullptr;
};

}  // namespace afhip
