// Interface between the solver (vfem_device.hip) and the learned coarse-grid
// correction (vfem_unet.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "vfem.h"

namespace vfem {
namespace unet {
constexpr int kNumLayers = 11;
}

// Inference buffers for one [4][d][h][w] input (batch 1); not during graph capture.
int unet_prepare(vfem_unet *net, int d, int h, int w);
// device input [4][d][h][w] (3 residual channels scaled by 1/s, occupancy code) and
// output [3][d][h][w] of the prepared shape
float *unet_input(vfem_unet *net);
const float *unet_output(vfem_unet *net);
// forward pass of the prepared shape on `stream` (graph-capturable: no allocation,
// no memset)
void unet_forward_enqueue(vfem_unet *net, hipStream_t stream);
const char *unet_last_error();

}  // namespace vfem
