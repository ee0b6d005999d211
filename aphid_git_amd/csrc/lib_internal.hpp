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

// Training with a loss computed outside the network (the solver's residual
// of the network's guess).  All on the network's stream, which each call
// synchronises before returning.
int samples(const aphid_unet* net);             // training samples stored
int zero_grads(aphid_unet* net);                // clear the parameter gradients
// forward pass on stored training sample idx; *out = device output [V][8]
int forward_sample(aphid_unet* net, int idx, const float** out);
float* out_grad(aphid_unet* net);               // device d loss / d output [V][8]
int backward_sample(aphid_unet* net, int idx);  // accumulate parameter gradients
// backward pass of the last forward_host; grad (optional, host) receives
// d loss / d params
int backward_input(aphid_unet* net, float* grad);
int adam_step(aphid_unet* net, float lr);       // one Adam update of the parameters
}  // namespace unet_internal
