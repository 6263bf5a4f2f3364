// Learned motion proposer: kernel argument blocks, parameter layout and
// launchers (host- and device-visible).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace vp8g {

constexpr float kMvScale = 8.0f;     // the head predicts mv / 8
constexpr float kHuberDelta = 0.25f; // in mv / 8 units (2 px)
constexpr int kHeadPartials = 67;    // dW5[2][32], db5[2], loss
constexpr int32_t kIgnoreLabel = -32768;  // label mvx marking a block left out of the loss
constexpr float kLeak = 0.0625f;          // leaky ReLU slope for negative inputs

// Input features per block: the 81 costs of a +-4 downsampled-pixel search
// (channel 9 * (ddy + 4) + (ddx + 4)), then the anchor's and the
// neighbour's 4x4 downsampled pixels, then zeros.
constexpr int kFeat = 128;
constexpr int kSearch = 4;
constexpr int kCostChannels = (2 * kSearch + 1) * (2 * kSearch + 1);

// Parameter layout (fp32, in this order): per conv layer the weights
// [taps][cout][cin] (3x3: tap = 3 * dy + dx) then the bias [cout]; then the
// head W5[2][32], b5[2].
struct ConvShape {
  int taps, cin, cout;
};
constexpr int kLayers = 5;
constexpr ConvShape kConv[kLayers] = {{1, kFeat, 32}, {9, 32, 32}, {9, 32, 64}, {9, 64, 32}, {9, 32, 32}};
constexpr int conv_weights(int l) { return kConv[l].taps * kConv[l].cin * kConv[l].cout; }
constexpr int conv_offset(int l) {
  return l == 0 ? 0 : conv_offset(l - 1) + conv_weights(l - 1) + kConv[l - 1].cout;
}
constexpr int kHeadOffset = conv_offset(kLayers);
constexpr int kParamCount = kHeadOffset + 66;

enum { EPI_BIAS_RELU = 0, EPI_MASK = 1 };

// 4x downsampling: plane (pair, f) pixel (x, y) = rounded mean of the 4x4
// full-resolution pixels at (4x.., 4y..) clamped to the frame; f = 0 is the
// anchor, 1 the pair's neighbour.
struct PrepArgs {
  const uint8_t* anchor;
  uint32_t anchor_stride;
  const uint8_t* neighbours[64];
  uint32_t neighbour_stride[64];
  int pairs, w, h, cw, ch;
  uint8_t* out;  // [pairs][2][4 ch][4 cw], or with shared_anchor [1 + pairs][4 ch][4 cw]
  // The anchor downsampled once as plane 0, neighbour p as plane 1 + p
  // (inference); otherwise each pair holds its own anchor copy (training).
  int shared_anchor;
};

struct FeatArgs {
  const uint8_t* planes;  // [images][2][4 ch][4 cw], or with shared_anchor [1 + images][4 ch][4 cw]
  uint16_t* out;          // [images][ch][cw][kFeat] bf16 (when w0 is null)
  int images, cw, ch;
  // Inference: the first layer (1x1 conv kFeat -> 32, bias, leaky ReLU)
  // applied in the same kernel, out0 = [images][ch][cw][32] bf16 instead of
  // the features.
  const uint16_t* w0;     // [32][kFeat] bf16, or null
  const float* b0;        // [32]
  uint16_t* out0;
  int shared_anchor;      // planes hold one anchor plane and then the neighbours
};

struct ConvArgs {
  const uint16_t* in;    // [images][ch][cw][cin] bf16
  uint16_t* out;         // [images][ch][cw][cout] bf16
  const uint16_t* w;     // [9][cout][cin] bf16
  const float* bias;     // [cout] (EPI_BIAS_RELU)
  const uint16_t* mask;  // [images][ch][cw][cout] bf16 (EPI_MASK)
  int images, cw, ch;
};

// Inference form of a 3x3 conv (bias + leaky ReLU): persistent workgroups
// that stage the weights once and overlap each tile's input loads with the
// previous tile's products.  With `head_params` set (the last layer) the
// activations are not stored: the head's vector goes to the alignment table.
// Inference of the four 3x3 layers and the head in one launch
// (conv_stack_kernel): a0 in, the alignment table out.
struct ConvStackArgs {
  const uint16_t* in;        // a0 [images][ch][cw][32] bf16
  const uint16_t* w[4];      // layers 1..4: [9][cout][cin] bf16
  const float* bias[4];      // layers 1..4: [cout]
  const float* head_params;  // W5[2][32], b5[2]
  int32_t* table;            // alignment table
  int frame_of_image[64];    // table frame index of each image
  int images, cw, ch, clamp;
};

struct WgradArgs {
  const uint16_t* dy;    // [images][ch][cw][cout]
  const uint16_t* x;     // [images][ch][cw][cin]
  float* partials;       // [slots][taps * cout * cin + cout]
  int images, cw, ch;
};

struct HeadArgs {
  const uint16_t* a4;        // [images][cells][32]
  const float* params;       // W5[2][32], b5[2]
  int images, cells;
  // inference
  int32_t* table;            // alignment table, or null
  int frame_of_image[64];    // table frame index of each image
  int clamp;
  // training
  const int32_t* labels;     // [images][cells][2] {mvx, mvy}, or null
  uint16_t* g4;              // [images][cells][32] gradient of a4's pre-activation
  float* partials;           // [blocks][kHeadPartials]
  float inv_count;           // 1 / (images * cells * 2)
};

struct AdamArgs {
  float* params;
  const float* grads;
  float* m;
  float* v;
  int n;
  float lr, b1, b2, eps, c1, c2;  // c1 = 1 / (1 - b1^t), c2 = 1 / (1 - b2^t)
};

struct PackArgs {
  const float* master;  // [taps][cout][cin]
  uint16_t* fwd;        // [taps][cout][cin]
  uint16_t* bwd;        // [taps][cin][cout], tap flipped (or null)
  int taps, cin, cout;
};

hipError_t launch_prep_downsample(const PrepArgs& a, hipStream_t s);
hipError_t launch_features(const FeatArgs& a, hipStream_t s);
hipError_t launch_conv(const ConvArgs& a, int taps, int cin, int cout, int epi, int cus, hipStream_t s);
int wgrad_slots(const WgradArgs& a, int cus);
hipError_t launch_wgrad(const WgradArgs& a, int taps, int cin, int cout, int slots, hipStream_t s);
hipError_t launch_reduce_partials(const float* partials, int slots, int n, float* out, hipStream_t s);
int head_blocks(long long cells);
hipError_t launch_head(const HeadArgs& a, hipStream_t s);
hipError_t launch_conv_stack(const ConvStackArgs& a, hipStream_t s);
hipError_t launch_adam(const AdamArgs& a, hipStream_t s);
hipError_t launch_pack_weights(const PackArgs& a, hipStream_t s);

}  // namespace vp8g
