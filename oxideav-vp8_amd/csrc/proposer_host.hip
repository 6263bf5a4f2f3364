// C ABI of the learned motion proposer: parameters, inference into an
// alignment table, and training steps, all on the context's stream.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "../../include/vp8g.h"
#include "proposer.h"
#include "vp8g_ctx.h"

using namespace vp8g;

struct vp8g_proposer {
  vp8g_ctx* ctx = nullptr;
  float* params = nullptr;  // [kParamCount]
  float* grads = nullptr;   // [kParamCount + 1]; the last slot holds the loss
  float* m = nullptr;
  float* v = nullptr;
  uint16_t* wf[kLayers] = {};  // bf16 [taps][cout][cin]
  uint16_t* wb[kLayers] = {};  // bf16 [taps][cin][cout], taps flipped
  int step = 0;
  uint8_t* arena = nullptr;  // activations, gradients, partials
  size_t arena_bytes = 0;
};

namespace {

bool ok(hipError_t e) { return e == hipSuccess; }
size_t up256(size_t v) { return (v + 255) & ~(size_t)255; }

int ensure(vp8g_proposer* p, size_t bytes) {
  if (bytes <= p->arena_bytes) return VP8G_OK;
  if (p->arena) {
    if (!ok(hipStreamSynchronize(p->ctx->stream))) return VP8G_E_DEVICE;
    (void)hipFree(p->arena);
    p->arena = nullptr;
    p->arena_bytes = 0;
  }
  if (!ok(hipMalloc(&p->arena, bytes))) return VP8G_E_DEVICE;
  p->arena_bytes = bytes;
  return VP8G_OK;
}

// Per-image layout inside the arena (bf16 channels per cell): features x0
// (kFeat) | a0..a4 (32, 32, 64, 32, 32) | g0..g4 (same), then the reduction
// partials and (for propose) the downsampled planes.
struct Layout {
  size_t x0;
  size_t a[kLayers];
  size_t g[kLayers];
  size_t partials;
  size_t planes;
  size_t total;
};

Layout layout(int images, int cells, int cus, bool with_planes) {
  Layout L{};
  const size_t per = (size_t)images * cells * 2;  // bytes per channel
  size_t o = 0;
  L.x0 = o;
  o = up256(o + per * kFeat);
  for (int i = 0; i < kLayers; i++) {
    L.a[i] = o;
    o = up256(o + per * kConv[i].cout);
  }
  for (int i = 0; i < kLayers; i++) {
    L.g[i] = o;
    o = up256(o + per * kConv[i].cout);
  }
  size_t wmax = 0;
  for (int i = 0; i < kLayers; i++) wmax = std::max(wmax, (size_t)conv_weights(i) + kConv[i].cout);
  const long long hb = head_blocks((long long)images * cells);
  const size_t part_floats = std::max((size_t)(2 * cus) * wmax, (size_t)hb * kHeadPartials);
  L.partials = o;
  o = up256(o + part_floats * 4);
  L.planes = o;
  if (with_planes) o = up256(o + (size_t)images * cells * 32);  // 2 planes x 16 px per cell
  L.total = o;
  return L;
}

int pack_all(vp8g_proposer* p) {
  for (int l = 0; l < kLayers; l++) {
    PackArgs a{p->params + conv_offset(l), p->wf[l], l > 0 ? p->wb[l] : nullptr, kConv[l].taps, kConv[l].cin,
               kConv[l].cout};
    if (!ok(launch_pack_weights(a, p->ctx->stream))) return VP8G_E_DEVICE;
  }
  return VP8G_OK;
}

// fuse_first (inference): the features kernel applies the first layer
// itself and the 128-channel features are never stored; training keeps them
// for the first layer's weight gradient.
int forward(vp8g_proposer* p, const uint8_t* planes, int images, int cw, int ch, const Layout& L,
            bool fuse_first = false) {
  uint8_t* b = p->arena;
  hipStream_t s = p->ctx->stream;
  FeatArgs f{planes, reinterpret_cast<uint16_t*>(b + L.x0), images, cw, ch, nullptr, nullptr, nullptr};
  if (fuse_first) {
    f.w0 = p->wf[0];
    f.b0 = p->params + conv_offset(0) + conv_weights(0);
    f.out0 = reinterpret_cast<uint16_t*>(b + L.a[0]);
  }
  if (!ok(launch_features(f, s))) return VP8G_E_DEVICE;
  const uint16_t* in = fuse_first ? f.out0 : f.out;
  for (int l = fuse_first ? 1 : 0; l < kLayers; l++) {
    ConvArgs a{};
    a.in = in;
    a.out = reinterpret_cast<uint16_t*>(b + L.a[l]);
    a.w = p->wf[l];
    a.bias = p->params + conv_offset(l) + conv_weights(l);
    a.images = images;
    a.cw = cw;
    a.ch = ch;
    if (!ok(launch_conv(a, kConv[l].taps, kConv[l].cin, kConv[l].cout, EPI_BIAS_RELU, p->ctx->cus, s)))
      return VP8G_E_DEVICE;
    in = a.out;
  }
  return VP8G_OK;
}

int backward(vp8g_proposer* p, const int32_t* labels, int images, int cw, int ch, const Layout& L) {
  uint8_t* b = p->arena;
  hipStream_t s = p->ctx->stream;
  float* part = reinterpret_cast<float*>(b + L.partials);
  const int cells = cw * ch;
  HeadArgs h{};
  h.a4 = reinterpret_cast<const uint16_t*>(b + L.a[kLayers - 1]);
  h.params = p->params + kHeadOffset;
  h.images = images;
  h.cells = cells;
  h.labels = labels;
  h.g4 = reinterpret_cast<uint16_t*>(b + L.g[kLayers - 1]);
  h.partials = part;
  h.inv_count = 1.0f / (2.0f * (float)images * (float)cells);
  if (!ok(launch_head(h, s))) return VP8G_E_DEVICE;
  const int hb = head_blocks((long long)images * cells);
  if (!ok(launch_reduce_partials(part, hb, kHeadPartials, p->grads + kHeadOffset, s))) return VP8G_E_DEVICE;
  for (int l = kLayers - 1; l >= 0; l--) {
    const uint16_t* x = reinterpret_cast<const uint16_t*>(b + (l == 0 ? L.x0 : L.a[l - 1]));
    const uint16_t* dy = reinterpret_cast<const uint16_t*>(b + L.g[l]);
    WgradArgs w{dy, x, part, images, cw, ch};
    const int slots = wgrad_slots(w, p->ctx->cus);
    const int n = conv_weights(l) + kConv[l].cout;
    if (!ok(launch_wgrad(w, kConv[l].taps, kConv[l].cin, kConv[l].cout, slots, s)) ||
        !ok(launch_reduce_partials(part, slots, n, p->grads + conv_offset(l), s)))
      return VP8G_E_DEVICE;
    if (l == 0) break;
    // Gradient of layer l's input pre-activation: the transposed conv,
    // masked where relu cut.
    ConvArgs a{};
    a.in = dy;
    a.out = reinterpret_cast<uint16_t*>(b + L.g[l - 1]);
    a.w = p->wb[l];
    a.mask = reinterpret_cast<const uint16_t*>(b + L.a[l - 1]);
    a.images = images;
    a.cw = cw;
    a.ch = ch;
    if (!ok(launch_conv(a, kConv[l].taps, kConv[l].cout, kConv[l].cin, EPI_MASK, p->ctx->cus, s)))
      return VP8G_E_DEVICE;
  }
  return VP8G_OK;
}

bool dims_ok(int images, int cw, int ch) {
  return images > 0 && images <= 4096 && cw > 0 && ch > 0 && cw <= 1024 && ch <= 1024;
}

}  // namespace

extern "C" {

int vp8g_proposer_create(vp8g_ctx* c, vp8g_proposer** out) {
  if (!c || !out) return VP8G_E_ARG;
  *out = nullptr;
  if (!ok(hipSetDevice(c->device))) return VP8G_E_DEVICE;
  vp8g_proposer* p = new vp8g_proposer();
  p->ctx = c;
  bool good = ok(hipMalloc(&p->params, kParamCount * 4)) && ok(hipMalloc(&p->grads, (kParamCount + 1) * 4)) &&
              ok(hipMalloc(&p->m, kParamCount * 4)) && ok(hipMalloc(&p->v, kParamCount * 4));
  for (int l = 0; l < kLayers && good; l++)
    good = ok(hipMalloc(&p->wf[l], conv_weights(l) * 2)) && ok(hipMalloc(&p->wb[l], conv_weights(l) * 2));
  good = good && ok(hipMemsetAsync(p->params, 0, kParamCount * 4, c->stream)) &&
         ok(hipMemsetAsync(p->grads, 0, (kParamCount + 1) * 4, c->stream)) &&
         ok(hipMemsetAsync(p->m, 0, kParamCount * 4, c->stream)) &&
         ok(hipMemsetAsync(p->v, 0, kParamCount * 4, c->stream)) && pack_all(p) == VP8G_OK &&
         ok(hipStreamSynchronize(c->stream));
  if (!good) {
    vp8g_proposer_destroy(p);
    return VP8G_E_DEVICE;
  }
  *out = p;
  return VP8G_OK;
}

void vp8g_proposer_destroy(vp8g_proposer* p) {
  if (!p) return;
  (void)hipSetDevice(p->ctx->device);
  (void)hipStreamSynchronize(p->ctx->stream);
  float* f[4] = {p->params, p->grads, p->m, p->v};
  for (float* q : f)
    if (q) (void)hipFree(q);
  for (int l = 0; l < kLayers; l++) {
    if (p->wf[l]) (void)hipFree(p->wf[l]);
    if (p->wb[l]) (void)hipFree(p->wb[l]);
  }
  if (p->arena) (void)hipFree(p->arena);
  delete p;
}

size_t vp8g_proposer_param_count(void) { return (size_t)kParamCount; }

int vp8g_proposer_set_params(vp8g_proposer* p, const float* host, size_t n) {
  if (!p || !host || n != (size_t)kParamCount) return VP8G_E_ARG;
  if (!ok(hipSetDevice(p->ctx->device))) return VP8G_E_DEVICE;
  hipStream_t s = p->ctx->stream;
  const bool good = ok(hipMemcpyAsync(p->params, host, n * 4, hipMemcpyHostToDevice, s)) &&
                    ok(hipMemsetAsync(p->m, 0, n * 4, s)) && ok(hipMemsetAsync(p->v, 0, n * 4, s)) &&
                    pack_all(p) == VP8G_OK && ok(hipStreamSynchronize(s));
  p->step = 0;
  return good ? VP8G_OK : VP8G_E_DEVICE;
}

int vp8g_proposer_get_params(vp8g_proposer* p, float* host, size_t n) {
  if (!p || !host || n != (size_t)kParamCount) return VP8G_E_ARG;
  if (!ok(hipSetDevice(p->ctx->device))) return VP8G_E_DEVICE;
  return ok(hipMemcpyAsync(host, p->params, n * 4, hipMemcpyDeviceToHost, p->ctx->stream)) &&
                 ok(hipStreamSynchronize(p->ctx->stream))
             ? VP8G_OK
             : VP8G_E_DEVICE;
}

int vp8g_proposer_get_grads(vp8g_proposer* p, float* host, size_t n, float* loss) {
  if (!p || !host || n != (size_t)kParamCount) return VP8G_E_ARG;
  if (!ok(hipSetDevice(p->ctx->device))) return VP8G_E_DEVICE;
  float l = 0.f;
  const bool good = ok(hipMemcpyAsync(host, p->grads, n * 4, hipMemcpyDeviceToHost, p->ctx->stream)) &&
                    ok(hipMemcpyAsync(&l, p->grads + kParamCount, 4, hipMemcpyDeviceToHost, p->ctx->stream)) &&
                    ok(hipStreamSynchronize(p->ctx->stream));
  if (loss) *loss = l;
  return good ? VP8G_OK : VP8G_E_DEVICE;
}

// The downsampled planes of a group: [pairs][2] (each pair's anchor and
// neighbour), or with `shared_anchor` the anchor once and then the
// neighbours (inference).
static int prepare(vp8g_proposer* p, const vp8g_i420* frames, size_t n, size_t center, uint8_t* planes,
                   bool shared_anchor) {
  PrepArgs a{};
  a.shared_anchor = shared_anchor ? 1 : 0;
  const vp8g_i420& c = frames[center];
  a.anchor = c.y;
  a.anchor_stride = (uint32_t)c.y_stride;
  int k = 0;
  for (size_t i = 0; i < n; i++) {
    if (i == center) continue;
    a.neighbours[k] = frames[i].y;
    a.neighbour_stride[k] = (uint32_t)frames[i].y_stride;
    k++;
  }
  a.pairs = k;
  a.w = (int)c.width;
  a.h = (int)c.height;
  a.cw = (a.w + 15) / 16;
  a.ch = (a.h + 15) / 16;
  a.out = planes;
  return ok(launch_prep_downsample(a, p->ctx->stream)) ? VP8G_OK : VP8G_E_DEVICE;
}

int vp8g_proposer_prepare_dev(vp8g_proposer* p, const vp8g_i420* frames, size_t n, size_t center,
                              uint8_t* planes) {
  if (!p || !planes || ((uintptr_t)planes & 15)) return VP8G_E_ARG;
  int st = vp8g_validate_group(frames, n, center, nullptr);
  if (st != VP8G_OK) return st;
  if (n < 2) return VP8G_OK;
  if (!ok(hipSetDevice(p->ctx->device))) return VP8G_E_DEVICE;
  return prepare(p, frames, n, center, planes, false);
}

int vp8g_proposer_forward_dev(vp8g_proposer* p, const uint8_t* planes, int images, int cw, int ch,
                              uint16_t* a4_out) {
  if (!p || !planes || ((uintptr_t)planes & 15) || !dims_ok(images, cw, ch)) return VP8G_E_ARG;
  if (!ok(hipSetDevice(p->ctx->device))) return VP8G_E_DEVICE;
  const Layout L = layout(images, cw * ch, p->ctx->cus, false);
  int st = ensure(p, L.total);
  if (st != VP8G_OK) return st;
  st = forward(p, planes, images, cw, ch, L);
  if (st == VP8G_OK && a4_out)
    st = ok(hipMemcpyAsync(a4_out, p->arena + L.a[kLayers - 1], (size_t)images * cw * ch * 64, hipMemcpyDeviceToDevice,
                           p->ctx->stream))
             ? VP8G_OK
             : VP8G_E_DEVICE;
  return st;
}

int vp8g_proposer_propose_dev(vp8g_proposer* p, const vp8g_i420* frames, size_t n, size_t center, int clamp,
                              int32_t* table) {
  if (!p || !table || ((uintptr_t)table & 3) || clamp < 0 || clamp > 16384) return VP8G_E_ARG;
  int st = vp8g_validate_group(frames, n, center, nullptr);
  if (st != VP8G_OK) return st;
  if (n < 2) return VP8G_OK;
  if (!ok(hipSetDevice(p->ctx->device))) return VP8G_E_DEVICE;
  const int w = (int)frames[center].width, h = (int)frames[center].height;
  const int cw = (w + 15) / 16, ch = (h + 15) / 16, images = (int)n - 1;
  const Layout L = layout(images, cw * ch, p->ctx->cus, true);
  st = ensure(p, L.total);
  if (st != VP8G_OK) return st;
  uint8_t* planes = p->arena + L.planes;
  // The inference convs address their input with 32-bit buffer offsets
  // (64 channels of every cell below 2^31 bytes); larger batches take
  // forward() and head_kernel, which give the same table.
  const bool big = (long long)images * cw * ch * 64 * 2 > 0x7FFFFFFFLL;
  st = prepare(p, frames, n, center, planes, !big);
  if (st != VP8G_OK) return st;
  int frame_of_image[64];
  {
    int k = 0;
    for (size_t i = 0; i < n; i++)
      if (i != center) frame_of_image[k++] = (int)i;
  }
  if (big) {
    st = forward(p, planes, images, cw, ch, L, true);
    if (st != VP8G_OK) return st;
    HeadArgs hd{};
    hd.a4 = reinterpret_cast<const uint16_t*>(p->arena + L.a[kLayers - 1]);
    hd.params = p->params + kHeadOffset;
    hd.images = images;
    hd.cells = cw * ch;
    hd.table = table;
    hd.clamp = clamp;
    for (int k = 0; k < images; k++) hd.frame_of_image[k] = frame_of_image[k];
    return ok(launch_head(hd, p->ctx->stream)) ? VP8G_OK : VP8G_E_DEVICE;
  }
  // Inference: the features kernel applies the first layer, then one launch
  // runs the four 3x3 layers and the head per 16 x 16-cell tile, writing the
  // table (the same activations and vectors as forward() followed by
  // head_kernel).
  uint8_t* b = p->arena;
  hipStream_t s = p->ctx->stream;
  FeatArgs f{planes, nullptr, images, cw, ch, p->wf[0], p->params + conv_offset(0) + conv_weights(0),
             reinterpret_cast<uint16_t*>(b + L.a[0]), 1};
  if (!ok(launch_features(f, s))) return VP8G_E_DEVICE;
  ConvStackArgs a{};
  a.in = f.out0;
  for (int l = 1; l < kLayers; l++) {
    a.w[l - 1] = p->wf[l];
    a.bias[l - 1] = p->params + conv_offset(l) + conv_weights(l);
  }
  a.head_params = p->params + kHeadOffset;
  a.table = table;
  for (int k = 0; k < images; k++) a.frame_of_image[k] = frame_of_image[k];
  a.images = images;
  a.cw = cw;
  a.ch = ch;
  a.clamp = clamp;
  return ok(launch_conv_stack(a, s)) ? VP8G_OK : VP8G_E_DEVICE;
}

int vp8g_proposer_backward_dev(vp8g_proposer* p, const uint8_t* planes, const int32_t* labels, int images,
                               int cw, int ch) {
  if (!p || !planes || !labels || ((uintptr_t)planes & 15) || !dims_ok(images, cw, ch)) return VP8G_E_ARG;
  if (!ok(hipSetDevice(p->ctx->device))) return VP8G_E_DEVICE;
  const Layout L = layout(images, cw * ch, p->ctx->cus, false);
  int st = ensure(p, L.total);
  if (st == VP8G_OK) st = forward(p, planes, images, cw, ch, L);
  if (st == VP8G_OK) st = backward(p, labels, images, cw, ch, L);
  return st;
}

int vp8g_proposer_train_step_dev(vp8g_proposer* p, const uint8_t* planes, const int32_t* labels, int images,
                                 int cw, int ch, float lr, float* loss_dev) {
  if (!(lr >= 0.f) || lr > 1.f) return VP8G_E_ARG;
  int st = vp8g_proposer_backward_dev(p, planes, labels, images, cw, ch);
  if (st != VP8G_OK) return st;
  p->step++;
  AdamArgs a{};
  a.params = p->params;
  a.grads = p->grads;
  a.m = p->m;
  a.v = p->v;
  a.n = kParamCount;
  a.lr = lr;
  a.b1 = 0.9f;
  a.b2 = 0.999f;
  a.eps = 1e-8f;
  a.c1 = (float)(1.0 / (1.0 - pow(0.9, p->step)));
  a.c2 = (float)(1.0 / (1.0 - pow(0.999, p->step)));
  hipStream_t s = p->ctx->stream;
  if (!ok(launch_adam(a, s)) || pack_all(p) != VP8G_OK) return VP8G_E_DEVICE;
  if (loss_dev && !ok(hipMemcpyAsync(loss_dev, p->grads + kParamCount, 4, hipMemcpyDeviceToDevice, s)))
    return VP8G_E_DEVICE;
  return VP8G_OK;
}

}  // extern "C"
