// Synthetic luma frame generation (host- and device-visible).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "vp8g_device.h"

namespace vp8g {

struct SynthRect {
  uint64_t seed;
  int x0, y0, w, h;
};

struct SynthArgs {
  uint8_t* out;                // [n][h][w]
  uint8_t* clean;              // [n][h][w] before the Gaussian stage, or null
  const int32_t* vel;          // [64][2] {vx, vy} px per frame
  const SynthRect* rects;      // [n][rects_per_frame]
  const uint64_t* noise_seeds; // [n]
  uint64_t texture_seed;
  int w, h, n, rects_per_frame;
  double sigma;
};

hipError_t launch_synth_luma(const SynthArgs& a, hipStream_t s);

}  // namespace vp8g
