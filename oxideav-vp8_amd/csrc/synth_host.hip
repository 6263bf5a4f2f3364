// C ABI of the GPU synthetic-frame generator (proposer training data).
#include <hip/hip_runtime.h>
#include <string.h>

#include <vector>

#include "../../include/vp8g.h"
#include "synth.h"
#include "vp8g_ctx.h"

using namespace vp8g;

extern "C" int vp8g_synth_luma_dev(vp8g_ctx* c, const vp8g_synth_spec* spec, uint8_t* out, uint8_t* clean) {
  if (!c || !spec || !out || !spec->vel || !spec->noise_seeds) return VP8G_E_ARG;
  if (spec->width == 0 || spec->height == 0 || spec->width > 16384 || spec->height > 16384 || spec->n == 0 ||
      spec->n > 256 || spec->rects_per_frame > 1024 || (spec->rects_per_frame && !spec->rects))
    return VP8G_E_ARG;
  for (size_t i = 0; i < spec->n * spec->rects_per_frame; i++) {
    const vp8g_synth_rect& r = spec->rects[i];
    if (r.w == 0 || r.h == 0 || r.x0 + r.w > spec->width || r.y0 + r.h > spec->height) return VP8G_E_ARG;
  }
  if (!(spec->sigma >= 0.0) || spec->sigma > 1000.0) return VP8G_E_ARG;
  if (hipSetDevice(c->device) != hipSuccess) return VP8G_E_DEVICE;
  const size_t nr = spec->n * spec->rects_per_frame;
  std::vector<SynthRect> rects(nr);
  for (size_t i = 0; i < nr; i++) {
    const vp8g_synth_rect& r = spec->rects[i];
    rects[i] = SynthRect{r.seed, (int)r.x0, (int)r.y0, (int)r.w, (int)r.h};
  }
  const size_t o_vel = 0, o_seeds = 512, o_rects = o_seeds + ((8 * spec->n + 255) & ~(size_t)255);
  const size_t bytes = o_rects + sizeof(SynthRect) * (nr ? nr : 1);
  uint8_t* dev = nullptr;
  if (hipMalloc(&dev, bytes) != hipSuccess) return VP8G_E_DEVICE;
  hipStream_t s = c->stream;
  bool good = hipMemcpyAsync(dev + o_vel, spec->vel, 512, hipMemcpyHostToDevice, s) == hipSuccess &&
              hipMemcpyAsync(dev + o_seeds, spec->noise_seeds, 8 * spec->n, hipMemcpyHostToDevice, s) == hipSuccess &&
              (nr == 0 || hipMemcpyAsync(dev + o_rects, rects.data(), sizeof(SynthRect) * nr, hipMemcpyHostToDevice,
                                         s) == hipSuccess);
  if (good) {
    SynthArgs a;
    a.out = out;
    a.clean = clean;
    a.vel = reinterpret_cast<const int32_t*>(dev + o_vel);
    a.rects = reinterpret_cast<const SynthRect*>(dev + o_rects);
    a.noise_seeds = reinterpret_cast<const uint64_t*>(dev + o_seeds);
    a.texture_seed = spec->texture_seed;
    a.w = (int)spec->width;
    a.h = (int)spec->height;
    a.n = (int)spec->n;
    a.rects_per_frame = (int)spec->rects_per_frame;
    a.sigma = spec->sigma;
    good = launch_synth_luma(a, s) == hipSuccess;
  }
  // The staging buffer is released once the kernel has consumed it.
  const bool synced = hipStreamSynchronize(s) == hipSuccess;
  (void)hipFree(dev);
  return good && synced ? VP8G_OK : VP8G_E_DEVICE;
}
