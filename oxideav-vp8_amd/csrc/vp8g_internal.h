// Kernel argument blocks and launchers shared between the kernel sources
// and the C-ABI layer.  Host- and device-visible.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "../../include/vp8g.h"

namespace vp8g {

constexpr uint32_t kWMax = 16;              // centre weight, src/arnr.rs:37
constexpr uint32_t kSadPerPixelCutoff = 12; // src/arnr.rs:42
constexpr int kMaxStrength = 6;             // src/arnr.rs:69
constexpr int kMaxFrames = 64;              // frames per group carried in kernel arguments

// One I420 frame as the kernels see it (device pointers).
struct FrameDesc {
  const uint8_t* y;
  const uint8_t* u;
  const uint8_t* v;
  uint32_t y_stride;
  uint32_t uv_stride;
};

// One ARNR group (lookahead window) as the kernels see it.
struct ArnrGroup {
  uint8_t* out_y;   // packed w*h
  uint8_t* out_u;   // packed cw*ch
  uint8_t* out_v;
  int32_t* mv_out;  // optional n*brows*bcols*{mvx,mvy,sad}
  int w, h, cw, ch, bcols, brows;
  int n, center;
  int filter_on;    // 0: strength 0 or a single frame -> exact centre copy
  int aligned4;     // every luma plane pointer and stride is a multiple of 4
  int aligned16;    // ... of 16
  int aligned4c;    // every chroma plane pointer and stride is a multiple of 4
  int uniform_strides;  // all frames share y_stride and uv_stride; plane offsets < 2^31
  int strips_per_row;
};

// Single-group launch: frames travel in the kernel arguments.
struct ArnrArgs {
  FrameDesc frames[kMaxFrames];
  ArnrGroup g;
  uint32_t wt4[64]; // blend weight per |ref - ctr| as 256 bytes (values <= 16)
  int xcd_remap;    // map workgroups to contiguous MB bands per XCD
  int mbs_per_wg;   // M: anchor MBs per workgroup (one strip of an MB row): 1, 2, 4 or 8
  // strip / strips_per_row == umulhi(strip, spr_mul) for strips below 2^20
  // (spr_mul = ceil(2^32 / strips_per_row); 0 when strips_per_row is 1)
  uint32_t spr_mul;
  // The launch covers strips [strip_base, strip_base + strip_count) (whole
  // MB rows of the group, for the banded host entry point).
  int strip_base, strip_count;
};
// Launch shape: a workgroup holds M anchor MBs x K waves per MB (K splits the
// neighbour frames), 64 * M * K threads, M * K <= 16.

// Batched launch: many groups of any size and length in one persistent
// kernel; workgroups take (group, strip) units from an atomic ticket.
// Units are split into kBatchChunks contiguous chunks with one ticket each;
// workgroup i starts on chunk i % kBatchChunks (its XCD's share) and then
// drains the others.
constexpr int kBatchChunks = 8;

struct ArnrBatchJob {
  ArnrGroup g;
  int frame_base;   // index of the group's first frame in ArnrBatchArgs::frames
  int unit_base;    // first unit (strip) of this group
  uint32_t wt4[64];
};

struct ArnrBatchArgs {
  const ArnrBatchJob* jobs;
  const FrameDesc* frames;
  const uint16_t* unit_job;  // job index of every unit
  int total_units;
  uint32_t* ticket;          // kBatchChunks counters, zeroed before each launch
  int mbs_per_wg;
};

hipError_t launch_arnr_batch(const ArnrBatchArgs& a, int k, int workgroups, hipStream_t stream);
hipError_t launch_zero_u32(uint32_t* p, int n, hipStream_t stream);
int arnr_batch_workgroups_per_cu(int k, int mbs_per_wg);

constexpr size_t kMaxLdsPerWorkgroup = 65536;
size_t arnr_lds_bytes(int k, int mbs_per_wg, bool table);
int arnr_supported_mbs(int m);  // strip width rounded down to 1, 2, 4 or 8
hipError_t launch_arnr_fused(const ArnrArgs& a, int k, hipStream_t stream);
hipError_t launch_arnr_table(const ArnrArgs& a, const int32_t* table, int k, hipStream_t stream);

// Exhaustive block search.
constexpr int kMaxSeed = 16384;  // seed components are clamped to +-kMaxSeed
struct SearchArgs {
  FrameDesc frames[kMaxFrames];
  const int32_t* seeds;   // optional, table layout n x brows x bcols x 3 ({sx, sy, -})
  int32_t* out;           // table layout; centre rows untouched
  int w, h, bcols, brows, n, center;
  int aligned4;           // every luma plane pointer and stride is a multiple of 4
  int total_blocks;       // (n - 1) * brows * bcols
};
// Pyramid level (2x2 rounded mean of Y, U and V) and seed upsampling.
constexpr int kDownsampleFrames = 16;  // frames per downsample launch
struct Downsample2Frame {
  const uint8_t *src_y, *src_u, *src_v;
  uint8_t *dst_y, *dst_u, *dst_v;
  uint32_t src_y_stride, src_uv_stride, dst_y_stride, dst_uv_stride;
};
struct Downsample2Args {
  Downsample2Frame f[kDownsampleFrames];
  int w, h;  // source luma size (shared by the frames)
};
hipError_t launch_downsample2(const Downsample2Args& a, int frames, hipStream_t stream);
hipError_t launch_upsample_table(const int32_t* coarse, int cbw, int cbh, int32_t* fine, int fbw, int fbh, int n,
                                 hipStream_t stream);
// Exhaustive fallback of a seeded search (queue of (frame, block) items).
struct FallbackArgs {
  const int32_t* first;   // first-stage table (table layout)
  int32_t* out;           // output table (table layout); centre rows untouched
  uint32_t* queue;        // capacity >= total_blocks items
  uint32_t* count;        // queue length, cleared by the launcher
  uint32_t* queued;       // optional: receives the queue length
  int w, h, bcols, brows, n, center, total_blocks;
  int gate;               // 1: queue only items with a 4-neighbour under the cut-off
};
hipError_t launch_search_fallback(const SearchArgs& a, const FallbackArgs& f, int radius, int cus,
                                  hipStream_t stream);
bool block_search_supported(int radius);
hipError_t launch_refine_candidates(const SearchArgs& a, const int32_t* first, hipStream_t stream);
hipError_t launch_block_search(const SearchArgs& a, int radius, hipStream_t stream);

// Motion compensation / SAD scoring.
struct McPlanes {
  const uint8_t* y;
  const uint8_t* u;
  const uint8_t* v;
  uint32_t y_stride, uv_stride;
  int lw, lh;  // luma clamp bounds
  int cw, ch;  // chroma clamp bounds
};

struct McArgs {
  McPlanes ref;
  const uint32_t* mb_xy;   // n x {mb_col, mb_row}
  const int16_t* mv;       // n x {row, col} quarter-pel
  uint8_t* out;            // n x 384 (y256 | u64 | v64)
  int n;
  int full_pixel;
  int32_t filters[8][6];
  int aligned4;            // set by the launcher: planes and strides 4-byte aligned
  int dot;                 // set by the launcher: the table suits the dot-product form
  uint32_t dot_lo[8], dot_hi[8];  // set by the launcher: rows packed as signed bytes
  uint32_t ident_rows;     // set by the launcher: bit r = row r is {0, 0, 128, 0, 0, 0}
};

struct SadArgs {
  const uint8_t* plane;
  uint32_t stride;
  int w, h;                // clamp bounds (LumaRef width/height)
  const uint32_t* mb_xy;
  const uint8_t* src;      // n x 256
  const int16_t* mv;       // n x {row, col}
  uint32_t* sad;           // n
  int n;
  int aligned4;            // set by the launcher: plane and stride 4-byte aligned
};

struct HalfPlaneArgs {
  const uint8_t* plane;
  uint32_t stride;
  int w, h;
  const int32_t* xy;       // n x {x, y}: whole-pel position of the 16x16 block
  const uint8_t* src;      // n x 256
  uint32_t* sad;           // n x 8, (dy, dx) raster order without (0, 0)
  int n;
  int diagonals;
  int32_t taps[6];
};

hipError_t launch_half_plane_sads(const HalfPlaneArgs& a, hipStream_t stream);

struct RefineArgs {
  const uint8_t* plane;
  uint32_t stride;
  int w, h;
  const uint32_t* mb_xy;
  const uint8_t* src;      // n x 256
  const int16_t* center;   // n x {row, col}
  uint32_t* out;           // n x {row | col << 16, sad} (vp8g_search_result layout)
  int n;
  int step;                // 2: half-pel ladder, 1: quarter-pel ladder
  int aligned4;            // set by the launcher: plane and stride 4-byte aligned
};

// Encoder inter search (whole_pixel_search, sub_pixel_refine,
// small_diamond_search_luma).
struct WpsArgs {
  const uint8_t* plane;
  uint32_t stride;
  int w, h;                      // LumaRef clamp bounds
  const float* bits;             // MvCostTable [2][2047]
  const vp8g_mb_search* jobs;    // n
  const uint8_t* src;            // n x 256
  vp8g_found* out;               // n
  int n;
  uint32_t descent_rounds;
  int wide_pattern;
};

struct SprArgs {
  const uint8_t* plane;
  uint32_t stride;
  int w, h;
  const float* bits;
  const vp8g_mb_search* jobs;
  const uint8_t* src;
  const vp8g_found* start;
  vp8g_found* out;
  int n;
  int half_diagonals, quarter;
  int32_t taps[6];               // the caller's filters[4] (HalfPlanes)
};

struct DiamondArgs {
  const uint8_t* plane;
  uint32_t stride;
  int w, h;
  const uint32_t* mb_xy;         // n x {mb_col, mb_row}
  const uint8_t* src;            // n x 256
  const int16_t* center;         // n x {row, col}
  uint32_t* out;                 // n x vp8g_search_result
  int n;
  uint32_t max_iters;
};

hipError_t launch_whole_pixel_search(const WpsArgs& a, hipStream_t stream);
hipError_t launch_sub_pixel_refine(const SprArgs& a, hipStream_t stream);
hipError_t launch_small_diamond(const DiamondArgs& a, hipStream_t stream);

// SPLITMV prediction and the residual of the inter reconstructions.
struct SplitArgs {
  McPlanes ref;
  const uint32_t* mb_xy;         // n x {mb_col, mb_row}
  const int16_t* mvs;            // n x 16 x {row, col} quarter-pel
  uint8_t* out;                  // n x 384
  int n;
  int full_pixel;
  int32_t filters[8][6];
};

struct ResidualArgs {
  const int16_t* coeffs;         // n x (y2 ? 25 : 24) x 16, dequantised
  const uint8_t* skip;           // n (mb_skip_coeff)
  uint8_t* out;                  // n x 384, holding the prediction
  int n;
  int y2;                        // 1: a Y2 block comes first (non-SPLITMV)
};

hipError_t launch_split_predict(const SplitArgs& a, hipStream_t stream);
hipError_t launch_mb_residual(const ResidualArgs& a, hipStream_t stream);

// Luma SAD of frame pairs (the stream driver's scene-cut metric).
constexpr int kLumaSadPairsPerLaunch = 64;
struct LumaSadPair {
  const uint8_t* a;
  const uint8_t* b;
  uint32_t a_stride, b_stride;
};
struct LumaSadArgs {
  LumaSadPair pairs[kLumaSadPairsPerLaunch];
  uint64_t* sad;                 // one total per pair
  int w, h;
};
hipError_t launch_luma_sad_pairs(const LumaSadArgs& a, int pairs, hipStream_t stream);

hipError_t launch_predict_inter_mb(const McArgs& a, hipStream_t stream);
hipError_t launch_mb_luma_sad_at_mv(const SadArgs& a, hipStream_t stream);
hipError_t launch_refine_ladder(const RefineArgs& a, hipStream_t stream);

// Loop filter: one workgroup per instance (candidate x band); see
// lf_kernels.hip.
constexpr int kLfMaxCands = 256;
constexpr int kLfMaxWaves = 12;  // waves per instance workgroup (168 VGPRs each)
struct LfArgs {
  const uint8_t* y;              // reconstruction (read)
  const uint8_t* u;
  const uint8_t* v;
  uint8_t* oy;                   // frame mode: filtered output (may alias y/u/v)
  uint8_t* ou;
  uint8_t* ov;
  const uint8_t* sy;             // score mode: source picture
  const uint8_t* su;
  const uint8_t* sv;
  size_t y_stride, uv_stride, sy_stride, suv_stride;
  int mb_cols, mb_rows;
  int vis_w, vis_h, vis_cw, vis_ch;  // visible window of the source
  int src_aligned4;              // source planes and strides are multiples of 4
  const uint32_t* mbs;           // vp8g_lf_mb per MB, raster order
  int key_frame, segmentation_enabled, segment_abs, delta_enabled;
  int segment_lf_level[4], ref_delta[4], mode_delta[4];
  int n_bands, every_rows, band_rows;  // instance i: candidate i / n_bands, rows from (i % n_bands) * every_rows
  uint8_t cand_level[kLfMaxCands], cand_sharp[kLfMaxCands];
  uint8_t* ring;                 // row rings between waves, ring_instance_bytes per instance
  size_t ring_instance_bytes;
  unsigned long long* inst_ssd;  // score mode: one SSD per instance
  int* err;                      // set when a row wait gives up
  int spin_limit;                // polls before a row wait gives up
};
size_t lf_lds_bytes(int waves);
size_t lf_ring_bytes(int waves, int mb_cols);
hipError_t launch_loop_filter(const LfArgs& a, int instances, int waves, bool simple, bool score,
                              hipStream_t stream);
hipError_t launch_lf_sum_bands(const unsigned long long* inst, int n_cands, int n_bands, unsigned long long* out,
                               hipStream_t stream);

}  // namespace vp8g
