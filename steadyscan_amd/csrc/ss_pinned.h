// ss_pinned.h -- page-locked host staging for the library's host-side loops.
// No HIP headers here, so that host-only translation units can use it.
#pragma once

#include <cstddef>

namespace ss {

// Page-locked host memory (device-to-host copies into it run at DMA speed);
// nullptr when none can be had.  Release with pinned_free.
void *pinned_alloc(size_t bytes);
void pinned_free(void *p);

}  // namespace ss
