// launch.h -- host-side launchers of the kernels in kernels.hip.
#pragma once
#include <hip/hip_runtime.h>
#include "jobs.h"

namespace sw {
hipError_t launch_world(const WorldArgs &W, hipStream_t s);
hipError_t launch_vertex(const BatchDev &B, uint32_t blocks, uint32_t w0, uint32_t n_words, hipStream_t s);
// marks, tile rectangles and counts, and (last block) the scan into bin offsets
hipError_t launch_front(const BatchDev &B, uint32_t blocks, hipStream_t s);
// the fill over `blocks` blocks, plus one block that writes the claim order;
// `done`, if given, is recorded by the launch's own completion
hipError_t launch_bin_fill(const BatchDev &B, uint32_t blocks, hipStream_t s, hipEvent_t done = nullptr);
hipError_t launch_raster(const BatchDev &B, int tw, int th, uint32_t max_groups, bool generic, hipStream_t s);
// A wave of one job with transparency layers (16x16 tiles): the in-order
// layer raster, then the flatten over the screen and the copy back.
hipError_t launch_layers(const BatchDev &B, bool generic, uint32_t pixels, hipStream_t s);
hipError_t launch_dof(const BatchDev &B, uint32_t blocks, int max_radius, hipStream_t s);
// Resolve every kernel of this file on the current device (code object
// loaded, each kernel's symbol looked up) without launching any, so that no
// kernel's first use happens inside a stream capture.
hipError_t resolve_kernels();
}  // namespace sw
