// xfer.hpp -- large copies between pageable host memory and the device.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>

namespace dfm {

// Copies bytes between pageable host memory and device memory (toDevice:
// host src -> device dst, else device src -> host dst), synchronously with
// respect to all prior device work.  Copies of 64 MiB and more go through
// two pinned staging buffers: the device moves one chunk while host threads
// move the previous one between the staging buffer and the pageable memory
// (which also spreads its first-touch page faults over those threads).
hipError_t copyLarge(void* dst, const void* src, size_t bytes, bool toDevice);

}  // namespace dfm
