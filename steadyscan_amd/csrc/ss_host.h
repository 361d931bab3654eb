// ss_host.h -- what the other translation units of the library need from the
// handle, whose definition lives with the alignment engine.
#pragma once

#include <string>

#include <hip/hip_runtime.h>

#include "steadyscan.h"

namespace ss {

// Records msg as the handle's last error and returns st.
ss_status handle_fail(ss_handle *h, ss_status st, const std::string &msg);
// The handle's device and stream (default priority, non-blocking).
int handle_device(const ss_handle *h);
hipStream_t handle_stream(const ss_handle *h);
// Device work memory of at least `bytes` kept by the handle across calls
// (grown on demand, freed by ss_destroy); nullptr on failure (st set).
void *handle_scratch(ss_handle *h, size_t bytes, ss_status *st);
// Bytes of work memory the handle holds now.
size_t handle_scratch_held(const ss_handle *h);
// Looks up the device address of ss_resample_noalloc's static work area on
// `device` (which must be current), once per device; later calls return the
// first result.
hipError_t resample_scratch_init(int device);

}  // namespace ss
