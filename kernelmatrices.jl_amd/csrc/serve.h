// Serving loop internals (C ABI in include/kmhodlr.h).
#pragma once
#include "common.h"

namespace kmh {
uint64_t request_footprint(int64_t n, const kmh_build_opts& o, int nv);
}  // namespace kmh
