// Internal interface of the U-Net (unet.hip) used by the cavity solver's
// learned warm start.
#pragma once

#include <hip/hip_runtime.h>

#include "flursys_hip.h"

namespace fl {

// Resolution per axis of the network's cubic input, and its context.
int unet_resolution(const fl_unet *net);
fl_ctx *unet_context(const fl_unet *net);
// One inference on device memory: `in` and `out` hold n^3 FP32 values at the
// network's resolution (x fastest); out = the predicted solution of the
// rediscretised Poisson problem with diagonal `diag`, zero mean.  Runs on
// `stream`; the network's buffers are shared, so calls must not overlap.
fl_status unet_infer_device(fl_unet *net, hipStream_t stream, const float *in, double diag,
                            float *out);

}  // namespace fl
