// Serving loop internals (C ABI in include/kmhodlr.h).
#pragma once
#include "common.h"

namespace kmh {
uint64_t request_footprint(int64_t n, const kmh_build_opts& o, int nv);
// Grow the context's pool by one chunk so that it holds a free block of at
// least `bytes` (nothing when one exists); the chunk stays with the pool.
void ctx_reserve(Ctx& c, size_t bytes);
}  // namespace kmh
