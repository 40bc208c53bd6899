// Project autotuner (see autotune.cpp).
#pragma once
#include <functional>
#include <string>

#include "runtime.hpp"

namespace afhip {

// Cached choice for key, or -1.
int autotune_lookup(const std::string &key);

// Returns the cached choice for key, or times launch(c) for c in [0, ncand)
// on the library stream (one warm-up and three timed runs each), caches the
// fastest and returns it. launch must be safe to repeat.
int autotune_pick(const std::string &key, int ncand, const std::function<void(int)> &launch);

}  // namespace afhip
