// lib_internal.hpp -- declarations shared by the translation units of
// libaphid_gpu.so (not part of the C ABI).
#pragma once
#include <hip/hip_runtime.h>

#include <string>

struct aphid_unet;

// sets the message returned by aphid_last_error() and returns `code`
int aphid_fail(int code, const std::string& msg);

namespace unet_internal {
// Runs the network's forward pass on host features feat[V][3] (cell grid,
// z fastest).  On success *out points to the device output [V][8] (FP32),
// valid until the next call on the network, and the dims are the cell grid.
int forward_host(aphid_unet* net, const float* feat, const float** out, int* nx, int* ny, int* nz);
}  // namespace unet_internal
