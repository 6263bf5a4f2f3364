// Learned motion proposer: a small BF16 CNN on MFMA (gfx950).
//
// Input: the 4x-downsampled (anchor, neighbour) luma pair, u8.
// Network (per 16x16 macroblock "cell" = 4x4 downsampled pixels):
//   x0  = features: the 81 mean absolute differences of the cell's 4x4
//         anchor pixels against the neighbour displaced by (ddx, ddy) in
//         [-4, 4]^2 downsampled pixels (a fixed correlation layer), then
//         the cell's anchor and neighbour pixels / 255, zero-padded to 128
//   a0  = lrelu(conv1x1(x0, 128 -> 32) + b0)
//   a1  = lrelu(conv3x3(a0, 32 -> 32) + b1)
//   a2  = lrelu(conv3x3(a1, 32 -> 64) + b2)
//   a3  = lrelu(conv3x3(a2, 64 -> 32) + b3)
//   a4  = lrelu(conv3x3(a3, 32 -> 32) + b4)
//   (lrelu(z) = z for z > 0, z / 16 otherwise)
//   mv  = 8 * (W5 a4 + b5)                      (1x1, 32 -> 2: mvx, mvy)
// Activations are NHWC bf16 [image][cell row][cell col][channel]; products
// run on v_mfma_f32_16x16x32_bf16 with fp32 accumulation.  Parameters are
// fp32 masters; the kernels read bf16 copies of the conv weights.
//
// MFMA operand layout (lane l, element i): A[l%16][8*(l/16)+i],
// B[8*(l/16)+i][l%16], D[4*(l/16)+i][l%16].
#include "proposer.h"
#include "vp8g_device.h"

namespace vp8g {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f32x4 mfma16(const bf16x8& a, const bf16x8& b, const f32x4& c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ bf16x8 as_bf16x8(const uint4& v) { return __builtin_bit_cast(bf16x8, v); }

__device__ __forceinline__ uint16_t bf16_bits(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }
// Two values rounded to bf16 (to nearest even, as bf16_bits) in one
// instruction, lo in bits 0..15.
__device__ __forceinline__ uint32_t bf16_pair(float lo, float hi) {
  typedef float f32x2 __attribute__((ext_vector_type(2)));
  typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
  return __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2{lo, hi}), bf16x2));
}
__device__ __forceinline__ float bf16_val(uint16_t b) { return (float)__builtin_bit_cast(__bf16, b); }

// ---------------------------------------------------------------- inputs
// One thread per 4 horizontally adjacent downsampled pixels of a plane: 4
// rows of 16 full-resolution bytes (dword loads when the row segment is
// inside the frame and aligned).
// Grid: x over the plane's quads, y over the planes: 2 * pairs (anchor and
// neighbour of each pair), or 1 + pairs with a shared anchor (so the plane,
// its source frame and stride are uniform per workgroup).
__global__ void prep_downsample_kernel(const PrepArgs a) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  const int pf = blockIdx.y;
  const int dw = 4 * a.cw, dh = 4 * a.ch;
  const uint32_t qw = (uint32_t)a.cw;  // quads per downsampled row (dw / 4)
  if (t >= qw * (uint32_t)dh) return;
  const int y = (int)(t / qw), xq = (int)(t - (uint32_t)y * qw);
  const int p = a.shared_anchor ? pf - 1 : pf >> 1;
  const bool anchor = a.shared_anchor ? pf == 0 : (pf & 1) == 0;
  const uint8_t* src = anchor ? a.anchor : a.neighbours[p];
  const uint32_t stride = anchor ? a.anchor_stride : a.neighbour_stride[p];
  const int x0 = 16 * xq;
  const bool fast = x0 + 16 <= a.w && ((((uintptr_t)src) | stride) & 3) == 0;
  uint32_t sum[4] = {0, 0, 0, 0};
#pragma unroll
  for (int yy = 0; yy < 4; yy++) {
    const uint8_t* row = src + (size_t)min(4 * y + yy, a.h - 1) * stride;
    if (fast) {
      const uint32_t* r32 = reinterpret_cast<const uint32_t*>(row + x0);
#pragma unroll
      for (int k = 0; k < 4; k++) sum[k] = sad4(r32[k], 0u, sum[k]);  // + the dword's 4 bytes
    } else {
#pragma unroll
      for (int k = 0; k < 4; k++)
#pragma unroll
        for (int xx = 0; xx < 4; xx++) sum[k] += row[min(x0 + 4 * k + xx, a.w - 1)];
    }
  }
  const uint32_t packed = ((sum[0] + 8) >> 4) | (((sum[1] + 8) >> 4) << 8) | (((sum[2] + 8) >> 4) << 16) |
                          (((sum[3] + 8) >> 4) << 24);
  *reinterpret_cast<uint32_t*>(a.out + (size_t)pf * dw * dh + (size_t)y * dw + 4 * xq) = packed;
}

// Features of one segment of 32 cells of one cell row of one image per
// workgroup, 10 threads per cell: thread k < 9 computes the 9 SADs of
// displacement row ddy = k - 4 (the cell's 4x4 anchor pixels against 4 rows
// of the clamped neighbour window, v_sad_u8); thread 9 writes both frames'
// pixels and the zero channels.  The segment's 4 anchor rows and its 12
// clamped neighbour rows (columns x0 - 4 .. x0 + 131) are staged in LDS with
// coalesced loads; the cells' 128-channel rows are assembled in LDS and
// leave as coalesced 16-byte stores.
constexpr int kFeatSeg = 32;

__global__ void __launch_bounds__(320) features_kernel(const FeatArgs a) {
  constexpr int AD = kFeatSeg;          // anchor dwords per row
  constexpr int BD = kFeatSeg + 2;      // neighbour dwords per row
  constexpr int BR = 4 + 2 * kSearch;   // neighbour rows
  constexpr int NA = 4 * AD, NB = BR * BD;
  constexpr int TPC = 2 * kSearch + 2;  // threads per cell
  __shared__ uint32_t sA[NA];
  __shared__ uint32_t sB[NB];
  // a cell's kFeat channels at a row pitch of kFeat + 8 (272 bytes, 68
  // dwords): the cells' rows start 4 banks apart, so the 2-byte writes of
  // one displacement from different cells do not meet in one bank
  constexpr int kRow = kFeat + 8;
  __shared__ __align__(16) uint16_t s_out[kFeatSeg * kRow];
  const int cx0 = blockIdx.x * kFeatSeg, cy = blockIdx.y, img = blockIdx.z;
  const int dw = 4 * a.cw, dh = 4 * a.ch;
  const int ncell = min(kFeatSeg, a.cw - cx0);
  const uint8_t* A = a.shared_anchor ? a.planes : a.planes + (size_t)img * 2 * dw * dh;
  const uint8_t* B = a.shared_anchor ? a.planes + (size_t)(1 + img) * dw * dh : A + (size_t)dw * dh;
  const int x0 = 4 * cx0, y0 = 4 * cy;
  const int tid = threadIdx.x;
  // staging dword i: anchor (row i / AD) or neighbour (row (i - NA) / BD,
  // clamped to the plane; columns clamped per byte at the edges)
  auto fetch = [&](int i) -> uint32_t {
    if (i < NA) {
      const int r = i / AD, q = i - r * AD;
      const int x = x0 + 4 * q;
      return x + 4 <= dw ? *reinterpret_cast<const uint32_t*>(A + (size_t)(y0 + r) * dw + x) : 0u;
    }
    const int t = i - NA, r = t / BD, q = t - r * BD;
    const uint8_t* row = B + (size_t)clampi(y0 - kSearch + r, 0, dh - 1) * dw;
    const int x = x0 - 4 + 4 * q;
    if (x >= 0 && x + 4 <= dw) return *reinterpret_cast<const uint32_t*>(row + x);
    uint32_t v = 0;
#pragma unroll
    for (int b = 0; b < 4; b++) v |= (uint32_t)row[clampi(x + b, 0, dw - 1)] << (8 * b);
    return v;
  };
  {
    const int i0 = tid, i1 = tid + 320;
    const uint32_t v0 = fetch(i0), v1 = i1 < NA + NB ? fetch(i1) : 0u;
    (i0 < NA ? sA[i0] : sB[i0 - NA]) = v0;
    if (i1 < NA + NB) sB[i1 - NA] = v1;
  }
  __syncthreads();
  const int jc = tid / TPC, k = tid - jc * TPC;
  if (jc < ncell) {
    uint16_t* out = s_out + jc * kRow;
    uint32_t arow[4];
#pragma unroll
    for (int r = 0; r < 4; r++) arow[r] = sA[r * AD + jc];
    if (k < 2 * kSearch + 1) {
      // window rows y0 + k - 4 + r, columns of the cell - 4 .. + 7
      uint32_t win[4][3];
#pragma unroll
      for (int r = 0; r < 4; r++)
#pragma unroll
        for (int q = 0; q < 3; q++) win[r][q] = sB[(k + r) * BD + jc + q];
#pragma unroll
      for (int ddx = 0; ddx < 2 * kSearch + 1; ddx++) {
        uint32_t sum = 0;
#pragma unroll
        for (int r = 0; r < 4; r++) {
          // bytes ddx .. ddx + 3 of the 12-byte window row (alignbyte uses
          // the shift mod 4, so the whole-dword offset 8 is taken directly)
          const uint32_t ref = ddx < 4 ? funnel(win[r][0], win[r][1], ddx)
                                       : (ddx < 8 ? funnel(win[r][1], win[r][2], ddx - 4) : win[r][2]);
          sum = sad4(arow[r], ref, sum);
        }
        out[(2 * kSearch + 1) * k + ddx] = bf16_bits((float)sum * (1.0f / 1024.0f));
      }
    } else {
      uint32_t brow[4];
#pragma unroll
      for (int r = 0; r < 4; r++) brow[r] = sB[(kSearch + r) * BD + jc + 1];
      auto px = [&](int p) -> uint32_t {  // pixel channel p: A then B, / 255
        const uint32_t v = p < 16 ? (arow[p >> 2] >> (8 * (p & 3))) & 255 : (brow[(p - 16) >> 2] >> (8 * (p & 3))) & 255;
        return bf16_bits((float)v * (1.0f / 255.0f));
      };
      // channel 81 alone (odd), then channel pairs 82/83 .. 126/127 as dwords
      out[kCostChannels] = (uint16_t)px(0);
      uint32_t* o32 = reinterpret_cast<uint32_t*>(out + kCostChannels + 1);
#pragma unroll
      for (int c = 0; c < (kFeat - kCostChannels - 1) / 2; c++) {
        const int p0 = 1 + 2 * c, p1 = 2 + 2 * c;  // pixel indices of the pair
        const uint32_t lo = p0 < 32 ? px(p0) : 0u, hi = p1 < 32 ? px(p1) : 0u;
        o32[c] = lo | (hi << 16);
      }
    }
  }
  __syncthreads();
  if (a.w0) {
    // The first layer straight from LDS: wave w < 4 computes output channels
    // 16 (w >> 1) .. +15 of cells 16 (w & 1) .. +15, with the operand layout,
    // summation order and epilogue of conv_kernel<1, kFeat, 32>, so the
    // activations are the ones that kernel writes.  Columns of cells past
    // the segment hold stale LDS; a column only reaches its own outputs,
    // which are not stored.
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), l = tid & 63;
    if (wave < 4) {
      const int m = wave >> 1, cell = 16 * (wave & 1) + (l & 15);
      const uint4* wrow = reinterpret_cast<const uint4*>(a.w0 + (size_t)(16 * m + (l & 15)) * kFeat);
      const uint4* xrow = reinterpret_cast<const uint4*>(s_out + cell * kRow);
      f32x4 acc = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int c = 0; c < kFeat / 32; c++) {
        const int ck = 4 * c + (l >> 4);
        acc = mfma16(as_bf16x8(wrow[ck]), as_bf16x8(xrow[ck]), acc);
      }
      if (cell < ncell) {
        const int co = 16 * m + 4 * (l >> 4);
        float v[4] = {acc[0], acc[1], acc[2], acc[3]};
#pragma unroll
        for (int i = 0; i < 4; i++) {
          const float z = v[i] + a.b0[co + i];
          v[i] = z > 0.f ? z : kLeak * z;
        }
        uint2 o;
        o.x = (uint32_t)bf16_bits(v[0]) | ((uint32_t)bf16_bits(v[1]) << 16);
        o.y = (uint32_t)bf16_bits(v[2]) | ((uint32_t)bf16_bits(v[3]) << 16);
        *reinterpret_cast<uint2*>(a.out0 + (((size_t)img * a.ch + cy) * a.cw + cx0 + cell) * 32 + co) = o;
      }
    }
    return;
  }
  // the segment's cells: consecutive rows of kFeat bf16 = 16 uint4 each
  uint4* dst = reinterpret_cast<uint4*>(a.out + (((size_t)img * a.ch + cy) * a.cw + cx0) * kFeat);
  const uint4* srcv = reinterpret_cast<const uint4*>(s_out);
  for (int i = tid; i < ncell * (kFeat / 8); i += 320) dst[i] = srcv[(i / (kFeat / 8)) * (kRow / 8) + i % (kFeat / 8)];
}

// ---------------------------------------------------------------- conv3x3
// LDS rows of 16-byte chunks are XOR-swizzled by the row index so that any
// 8 consecutive rows read at one chunk column (a 128-byte pass of a 16-byte
// read) hit distinct banks: with 64-byte rows (C = 32) two rows share a
// 128-byte line, so the key is the row pair's index.
template <int C>
__device__ __forceinline__ int swz(int row, int chunk) {
  return C >= 64 ? chunk ^ (row & (C / 8 - 1)) : chunk ^ ((row >> 1) & (C / 8 - 1));
}

// out = EPI_BIAS_RELU: lrelu(conv(in, w) + bias)
//       EPI_MASK:      conv(in, w) * (mask > 0 ? 1 : 1/16)  (backward through lrelu)
// w: [TAPS][COUT][CIN] bf16; TAPS 9: 3x3, tap = 3 * dy + dx reads offset
// (dy - 1, dx - 1), zero padded; TAPS 1: 1x1.
// One persistent workgroup of 4 waves per tile of 64 cells of one cell row;
// wave w computes cells 16w .. 16w + 15 for all COUT channels.
template <int TAPS, int CIN, int COUT, int EPI>
__global__ void __launch_bounds__(256) conv_kernel(const ConvArgs a) {
  constexpr int CK = CIN / 8;   // 16-byte chunks per cell
  constexpr int TW = 64;        // cells per tile
  constexpr int HALO = TAPS == 9 ? 1 : 0;
  constexpr int ROWS = 1 + 2 * HALO;
  constexpr int IW = TW + 2 * HALO;  // staged cells per row
  __shared__ __align__(16) uint4 s_w[TAPS * COUT * CK];
  __shared__ __align__(16) uint4 s_in[ROWS * IW * CK];
  const int tid = threadIdx.x, wave = __builtin_amdgcn_readfirstlane(tid >> 6), l = tid & 63;

  constexpr int WITERS = (TAPS * COUT * CK + 255) / 256;  // weight chunks per thread
  constexpr int IITERS = (ROWS * IW * CK + 255) / 256;    // input chunks per thread
  const uint4* wg = reinterpret_cast<const uint4*>(a.w);
  const int segs = (a.cw + TW - 1) / TW;
  const int total = a.images * a.ch * segs;
  const uint4* in = reinterpret_cast<const uint4*>(a.in);
  bool first = true;
  for (int tile = blockIdx.x; tile < total; tile += gridDim.x) {
    const int img = tile / (a.ch * segs);
    const int rem = tile - img * a.ch * segs;
    const int cy = rem / segs, cx0 = (rem - cy * segs) * TW;
    __syncthreads();  // previous tile's readers are done
    // Every staging load of the thread (the weights on the first tile, then
    // the input rows) is issued before the first LDS store.
    uint4 wv[WITERS], iv[IITERS];
    if (first) {
#pragma unroll
      for (int k = 0; k < WITERS; k++) wv[k] = wg[min(tid + 256 * k, TAPS * COUT * CK - 1)];
    }
#pragma unroll
    for (int k = 0; k < IITERS; k++) {
      const int i = min(tid + 256 * k, ROWS * IW * CK - 1);
      const int row = i / CK, ck = i - row * CK;  // row = r * IW + c
      const int r = row / IW, c = row - r * IW;
      const int y = cy + r - HALO, x = cx0 + c - HALO;
      iv[k] = make_uint4(0, 0, 0, 0);
      if (y >= 0 && y < a.ch && x >= 0 && x < a.cw) iv[k] = in[(((size_t)img * a.ch + y) * a.cw + x) * CK + ck];
    }
    if (first) {
#pragma unroll
      for (int k = 0; k < WITERS; k++) {
        const int i = tid + 256 * k;
        if (i < TAPS * COUT * CK) {
          const int row = i / CK, ck = i - row * CK;  // row = tap * COUT + co
          s_w[row * CK + swz<CIN>(row, ck)] = wv[k];
        }
      }
      first = false;
    }
#pragma unroll
    for (int k = 0; k < IITERS; k++) {
      const int i = tid + 256 * k;
      if (i < ROWS * IW * CK) {
        const int row = i / CK, ck = i - row * CK;
        s_in[row * CK + swz<CIN>(row, ck)] = iv[k];
      }
    }
    __syncthreads();
    f32x4 acc[COUT / 16];
#pragma unroll
    for (int m = 0; m < COUT / 16; m++) acc[m] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int t = 0; t < TAPS; t++) {
      const int dy = TAPS == 9 ? t / 3 : 0, dx = TAPS == 9 ? t - 3 * (t / 3) : 0;
      const int brow = dy * IW + 16 * wave + (l & 15) + dx;
#pragma unroll
      for (int c = 0; c < CIN / 32; c++) {
        const int ck = 4 * c + (l >> 4);
        const bf16x8 b = as_bf16x8(s_in[brow * CK + swz<CIN>(brow, ck)]);
#pragma unroll
        for (int m = 0; m < COUT / 16; m++) {
          const int arow = t * COUT + 16 * m + (l & 15);
          const bf16x8 av = as_bf16x8(s_w[arow * CK + swz<CIN>(arow, ck)]);
          acc[m] = mfma16(av, b, acc[m]);
        }
      }
    }
    const int cell_x = cx0 + 16 * wave + (l & 15);
    if (cell_x < a.cw) {
      const size_t cell = ((size_t)img * a.ch + cy) * a.cw + cell_x;
#pragma unroll
      for (int m = 0; m < COUT / 16; m++) {
        const int co = 16 * m + 4 * (l >> 4);
        float v[4] = {acc[m][0], acc[m][1], acc[m][2], acc[m][3]};
        if (EPI == EPI_BIAS_RELU) {
#pragma unroll
          for (int i = 0; i < 4; i++) {
            const float z = v[i] + a.bias[co + i];
            v[i] = z > 0.f ? z : kLeak * z;
          }
        } else {
          const uint2 mk = *reinterpret_cast<const uint2*>(a.mask + cell * COUT + co);
          const uint16_t mb[4] = {(uint16_t)mk.x, (uint16_t)(mk.x >> 16), (uint16_t)mk.y, (uint16_t)(mk.y >> 16)};
#pragma unroll
          for (int i = 0; i < 4; i++)
            if (!(bf16_val(mb[i]) > 0.f)) v[i] *= kLeak;
        }
        uint2 o;
        o.x = (uint32_t)bf16_bits(v[0]) | ((uint32_t)bf16_bits(v[1]) << 16);
        o.y = (uint32_t)bf16_bits(v[2]) | ((uint32_t)bf16_bits(v[3]) << 16);
        *reinterpret_cast<uint2*>(a.out + cell * COUT + co) = o;
      }
    }
  }
}

// ---------------------------------------------------------------- conv stack
// The four inference 3x3 layers and the head for one tile of 16 x 16 cells
// in one workgroup.  Each layer reaches one cell further, so the workgroup
// stages a0 with a 4-cell halo (a 24 x 24 region) and computes layer k on
// the region shrunk by k cells on each side: about 1.7x the products of the
// tile alone, far cheaper than a launch per layer (a layer of a 1080p group
// is under a microsecond of MFMA time, a launch several).  Cells of a layer
// outside the image are stored as zeros, the zero padding the next layer
// sees in the per-layer form.  The products of a cell, their order and its
// epilogue are those of conv_kernel<9, CIN, COUT, EPI_BIAS_RELU>, and the
// head's those of head_kernel, so the table equals forward() followed by
// head_kernel.
//
// Region cell (r, c) is at flat index 24 r + c.  Layer k (1..3) computes
// 16-cell blocks of consecutive flat indices from 25 k to past 25 (23 - k)
// (rows and columns k .. 23 - k are the cells it must get right); a block's
// products read flat index j + 24 (dy - 1) + dx - 1.  Cells of a block
// outside that square (the row ends, and the last block's run past the
// range) only feed cells of the next layer outside its own square, so their
// values do not matter.  The last layer takes one block per tile row
// (columns 4..19) and hands each cell's 32 activations to the head through
// LDS.
//
// LDS (150 KB, one workgroup per CU): S1 holds a0 (32 channels, 608 cells:
// the first layer's last block reads up to flat index 577), later a2 (64
// channels); S2 holds a1, later a3; one layer's weights, 16-byte chunks
// XOR-swizzled (swz).  A layer copies its weights into
// registers, then the next layer's are fetched into LDS by buffer loads
// while it runs.
__device__ __forceinline__ uint32_t stack_lds_addr(const void* p) {
  return (uint32_t)(uintptr_t)(const __attribute__((address_space(3))) void*)p;
}
typedef uint32_t u32x4v __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ bf16x8 stack_lds_read(uint32_t a) {
  return __builtin_bit_cast(bf16x8, *(const __attribute__((address_space(3))) u32x4v*)(uintptr_t)a);
}
__device__ __forceinline__ void stack_lds_write(uint32_t a, uint2 v) {
  *(__attribute__((address_space(3))) u32x2v*)(uintptr_t)a = u32x2v{v.x, v.y};
}

namespace stack {
// 8 waves, two per SIMD: one wave's reads and epilogue overlap the other's
// products (4 waves with two blocks per pass each were 15% slower).
constexpr int kWaves = 8;
constexpr int kThreads = 64 * kWaves;
constexpr int kP = 24;              // region pitch and rows, in cells
constexpr int kBufCells = 608;
constexpr int kWChunks = 9 * 64 * 4;  // the largest layer: 9 x 64 x 32 or 9 x 32 x 64 bf16
}  // namespace stack

template <int CIN, int COUT>
__device__ __forceinline__ void stack_fetch_weights(const uint16_t* w, uint4* s_w, int tid, int wave) {
  constexpr int CK = CIN / 8, NW = 9 * COUT * CK, IT = (NW + stack::kThreads - 1) / stack::kThreads;
  using LdsP = __attribute__((address_space(3))) void*;
  const __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc((void*)w, (short)0, NW * 16, 0x00020000);
#pragma unroll
  for (int k = 0; k < IT; k++) {
    const int slot = tid + stack::kThreads * k;
    if (slot < NW) {
      const int row = slot / CK, ck = swz<CIN>(row, slot - row * CK);
      __builtin_amdgcn_raw_ptr_buffer_load_lds(r, (LdsP)(s_w + stack::kThreads * k + 64 * wave), 16, (uint32_t)(row * CK + ck) * 16u,
                                               0, 0, 0);
    }
  }
}

// One layer: weights from LDS into registers, a barrier, the next layer's
// weights requested (next != null), then this wave's blocks.
template <int CIN, int COUT, int NCIN, int NCOUT, bool LAST>
__device__ __forceinline__ void stack_layer(const ConvStackArgs& a, int k, const uint4* src, uint4* dst, uint4* s_w,
                                            const uint16_t* next, int tid, int wave, int l, int img, int ty0,
                                            int tx0) {
  using namespace stack;
  constexpr int CK = CIN / 8, OK = COUT / 8, NC = CIN / 32, NM = COUT / 16;
  bf16x8 wr[9][NC][NM];
#pragma unroll
  for (int t = 0; t < 9; t++)
#pragma unroll
    for (int c = 0; c < NC; c++)
#pragma unroll
      for (int m = 0; m < NM; m++) {
        const int arow = t * COUT + 16 * m + (l & 15), ck = 4 * c + (l >> 4);
        wr[t][c][m] = as_bf16x8(s_w[arow * CK + swz<CIN>(arow, ck)]);
      }
  float bias[NM][4];
#pragma unroll
  for (int m = 0; m < NM; m++)
#pragma unroll
    for (int i = 0; i < 4; i++) bias[m][i] = a.bias[k - 1][16 * m + 4 * (l >> 4) + i];
  __syncthreads();  // every wave holds its weights
  if (next) stack_fetch_weights<NCIN, NCOUT>(next, s_w, tid, wave);
  const int first = 25 * k;
  const int nb = LAST ? 16 : (25 * (23 - k) - first) / 16 + 1;
  const int jf = LAST ? 4 * kP + 4 : first;  // block 0's first cell
  constexpr int kStep = LAST ? kP : 16;      // cells from one block to the next
  uint16_t* s_a4 = reinterpret_cast<uint16_t*>(dst) + wave * 16 * 32;
  // The lane's LDS byte addresses at block 0: its B chunk for tap row dy = 0
  // of each dx and 32-channel chunk, and its output chunk of each 16-channel
  // block.  The swizzle key of a row is the same 16 or 24 cells on and one
  // tap row (24 cells) down, so a block adds a uniform step and a tap row an
  // immediate offset.
  uint32_t boff[3][NC], woff[NM];
#pragma unroll
  for (int dx = 0; dx < 3; dx++)
#pragma unroll
    for (int c = 0; c < NC; c++) {
      const int row = jf + (l & 15) - kP + dx - 1;
      boff[dx][c] = stack_lds_addr(src) + (uint32_t)(row * CK + swz<CIN>(row, 4 * c + (l >> 4))) * 16u;
    }
#pragma unroll
  for (int m = 0; m < NM; m++) {
    const int co = 16 * m + 4 * (l >> 4), row = jf + (l & 15);
    woff[m] = stack_lds_addr(dst) + (uint32_t)(row * OK + swz<COUT>(row, co >> 3)) * 16u + (uint32_t)((co >> 2) & 1) * 8u;
  }
  // A block's epilogue: bias, leaky ReLU, bf16; zeros outside the image;
  // the last layer hands the cells to the head.
  const auto finish = [&](int b, const f32x4 (&acc)[NM]) {
    const int j = jf + kStep * b + (l & 15);
    const int rr = (int)((uint32_t)j / (uint32_t)kP), cc = j - rr * kP;
    const int y = ty0 - 4 + rr, x = tx0 - 4 + cc;
    const bool inside = y >= 0 && y < a.ch && x >= 0 && x < a.cw;
#pragma unroll
    for (int m = 0; m < NM; m++) {
      const int co = 16 * m + 4 * (l >> 4);
      float vv[4];
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const float z = acc[m][i] + bias[m][i];
        vv[i] = fmaxf(z, kLeak * z);  // z > 0 ? z : z / 16, signed zeros included
      }
      uint2 o = make_uint2(bf16_pair(vv[0], vv[1]), bf16_pair(vv[2], vv[3]));
      if (LAST) {
        *reinterpret_cast<uint2*>(s_a4 + (l & 15) * 32 + co) = o;
      } else {
        if (!inside) o = make_uint2(0u, 0u);
        stack_lds_write(woff[m] + (uint32_t)(b * kStep * OK * 16), o);
      }
    }
    if (LAST) {
      // a cell's 32 channels come from this wave's lanes only; lane l < 32
      // forms output l / 16 of cell l % 16
      wave_lds_sync();
      const int hc = l & 15, ho = (l >> 4) & 1;
      const int hy = ty0 + b, hx = tx0 + hc;
      if (l < 32 && hy < a.ch && hx < a.cw) {
        const uint16_t* s = s_a4 + hc * 32;
        float sum = a.head_params[64 + ho];
#pragma unroll
        for (int c = 0; c < 32; c++) sum += a.head_params[32 * ho + c] * bf16_val(s[c]);
        float mv = rintf(sum * kMvScale);
        mv = fminf(fmaxf(mv, (float)-a.clamp), (float)a.clamp);
        int32_t* tb = a.table + 3 * ((size_t)a.frame_of_image[img] * a.cw * a.ch + (size_t)hy * a.cw + hx);
        tb[ho] = (int32_t)mv;
        if (ho == 0) tb[2] = 0;
      }
      wave_lds_sync();  // the reads of s_a4 are done before the next block writes it
    }
  };
  for (int b = wave; b < nb; b += kWaves) {
    f32x4 acc[NM];
#pragma unroll
    for (int m = 0; m < NM; m++) acc[m] = f32x4{0.f, 0.f, 0.f, 0.f};
    // every B fragment of the block is read before the first product, so
    // one LDS latency covers the block
    bf16x8 bv[9][NC];
#pragma unroll
    for (int dx = 0; dx < 3; dx++)
#pragma unroll
      for (int c = 0; c < NC; c++) {
        uint32_t p = boff[dx][c] + (uint32_t)(b * kStep * CK * 16);
        asm volatile("" : "+v"(p));  // tap rows then fold into the read's offset
#pragma unroll
        for (int dy = 0; dy < 3; dy++) bv[3 * dy + dx][c] = stack_lds_read(p + (uint32_t)(dy * kP * CK * 16));
      }
#pragma unroll
    for (int t = 0; t < 9; t++)
#pragma unroll
      for (int c = 0; c < NC; c++)
#pragma unroll
        for (int m = 0; m < NM; m++) acc[m] = mfma16(wr[t][c][m], bv[t][c], acc[m]);
    // keep that order: the scheduler would otherwise place each read just
    // before its products and wait on it
    __builtin_amdgcn_sched_group_barrier(0x100, 9 * NC, 0);        // DS reads
    __builtin_amdgcn_sched_group_barrier(0x008, 9 * NC * NM, 0);   // MFMA
    finish(b, acc);
  }
  if (!LAST) {
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");  // the next layer's weights are in
    __syncthreads();                                  // and this layer's outputs
  }
}

__global__ void __launch_bounds__(stack::kThreads) conv_stack_kernel(const ConvStackArgs a) {
  using namespace stack;
  __shared__ __align__(16) uint4 s1[kBufCells * 8];
  __shared__ __align__(16) uint4 s2[kBufCells * 4];
  __shared__ __align__(16) uint4 s_w[kWChunks];
  const int tid = threadIdx.x, wave = __builtin_amdgcn_readfirstlane(tid >> 6), l = tid & 63;
  const int txs = (a.cw + 15) / 16, per = ((a.ch + 15) / 16) * txs;
  const int t = xcd_band(blockIdx.x, gridDim.x);
  const int img = t / per, rem = t - img * per;
  const int ty0 = rem / txs * 16, tx0 = (rem - rem / txs * txs) * 16;
  // a0's region (zeros outside the image and in the 32 cells past it) and
  // the first layer's weights, every load issued before the one wait
  stack_fetch_weights<32, 32>(a.w[0], s_w, tid, wave);
  {
    using LdsP = __attribute__((address_space(3))) void*;
    const uint32_t in_bytes = (uint32_t)a.images * a.ch * a.cw * 64;
    const __amdgpu_buffer_rsrc_t ir = __builtin_amdgcn_make_buffer_rsrc((void*)a.in, (short)0, (int)in_bytes, 0x00020000);
    constexpr int NIN = kBufCells * 4, IT = (NIN + kThreads - 1) / kThreads;
#pragma unroll
    for (int k = 0; k < IT; k++) {
      const int slot = tid + kThreads * k;
      if (slot < NIN) {
        const int cell = slot >> 2, ck = swz<32>(cell, slot & 3);
        const int rr = (int)((uint32_t)cell / (uint32_t)kP), cc = cell - rr * kP;
        const int y = ty0 - 4 + rr, x = tx0 - 4 + cc;
        if (cell < kP * kP && y >= 0 && y < a.ch && x >= 0 && x < a.cw) {
          const uint32_t off = (uint32_t)((((img * a.ch + y) * a.cw + x) * 4 + ck) * 16);
          __builtin_amdgcn_raw_ptr_buffer_load_lds(ir, (LdsP)(s1 + kThreads * k + 64 * wave), 16, off, 0, 0, 0);
        } else {
          s1[slot] = make_uint4(0, 0, 0, 0);
        }
      }
    }
  }
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  __syncthreads();
  stack_layer<32, 32, 32, 64, false>(a, 1, s1, s2, s_w, a.w[1], tid, wave, l, img, ty0, tx0);
  stack_layer<32, 64, 64, 32, false>(a, 2, s2, s1, s_w, a.w[2], tid, wave, l, img, ty0, tx0);
  stack_layer<64, 32, 32, 32, false>(a, 3, s1, s2, s_w, a.w[3], tid, wave, l, img, ty0, tx0);
  stack_layer<32, 32, 32, 32, true>(a, 4, s2, s1, s_w, nullptr, tid, wave, l, img, ty0, tx0);
}

// ---------------------------------------------------------------- wgrad
// Partial weight and bias gradients of one conv over a slice of 32-cell
// chunks: dW[t][co][ci] = sum_p dy[p][co] * x[p + d_t][ci], db[co] =
// sum_p dy[p][co].  dy and x are staged transposed (channel-major, 32 cells
// per row) so both MFMA operands read 16 contiguous bytes; x is staged once
// per tap (the dx shift is an element offset).  Each workgroup writes its
// own partial slot; a reduce kernel sums the slots in order.
template <int TAPS, int CIN, int COUT>
__global__ void __launch_bounds__(256) wgrad_kernel(const WgradArgs a) {
  constexpr int KC = 32;                         // cells per chunk (MFMA K)
  constexpr int HALO = TAPS == 9 ? 1 : 0;
  constexpr int TILES = TAPS * (COUT / 16) * (CIN / 16);
  constexpr int TPW = TILES / 4;                 // tiles per wave
  static_assert(TILES % 4 == 0, "tiles");
  __shared__ __align__(16) uint16_t s_dy[COUT * KC];
  __shared__ __align__(16) uint16_t s_x[TAPS * CIN * KC];
  const int tid = threadIdx.x, wave = __builtin_amdgcn_readfirstlane(tid >> 6), l = tid & 63;
  const int segs = (a.cw + KC - 1) / KC;
  const int total = a.images * a.ch * segs;
  f32x4 acc[TPW];
#pragma unroll
  for (int i = 0; i < TPW; i++) acc[i] = f32x4{0.f, 0.f, 0.f, 0.f};
  float bsum = 0.f;
  for (int chunk = blockIdx.x; chunk < total; chunk += gridDim.x) {
    const int img = chunk / (a.ch * segs);
    const int rem = chunk - img * a.ch * segs;
    const int cy = rem / segs, cx0 = (rem - cy * segs) * KC;
    __syncthreads();
    // dy^T: s_dy[co][p]
    for (int i = tid; i < KC * (COUT / 8); i += 256) {
      const int p = i / (COUT / 8), ck = i - p * (COUT / 8);
      uint4 v = make_uint4(0, 0, 0, 0);
      if (cx0 + p < a.cw) v = *reinterpret_cast<const uint4*>(a.dy + (((size_t)img * a.ch + cy) * a.cw + cx0 + p) * COUT + 8 * ck);
      const uint16_t e[8] = {(uint16_t)v.x, (uint16_t)(v.x >> 16), (uint16_t)v.y, (uint16_t)(v.y >> 16),
                             (uint16_t)v.z, (uint16_t)(v.z >> 16), (uint16_t)v.w, (uint16_t)(v.w >> 16)};
#pragma unroll
      for (int j = 0; j < 8; j++) s_dy[(8 * ck + j) * KC + p] = e[j];
    }
    // x^T per tap: s_x[t][ci][p] = x[cy + dy - 1][cx0 + p + dx - 1][ci]
    for (int i = tid; i < (1 + 2 * HALO) * (KC + 2 * HALO) * (CIN / 8); i += 256) {
      const int rc = i / (CIN / 8), ck = i - rc * (CIN / 8);
      const int r = rc / (KC + 2 * HALO), c = rc - r * (KC + 2 * HALO);
      const int y = cy + r - HALO, x = cx0 + c - HALO;
      uint4 v = make_uint4(0, 0, 0, 0);
      if (y >= 0 && y < a.ch && x >= 0 && x < a.cw)
        v = *reinterpret_cast<const uint4*>(a.x + (((size_t)img * a.ch + y) * a.cw + x) * CIN + 8 * ck);
      const uint16_t e[8] = {(uint16_t)v.x, (uint16_t)(v.x >> 16), (uint16_t)v.y, (uint16_t)(v.y >> 16),
                             (uint16_t)v.z, (uint16_t)(v.z >> 16), (uint16_t)v.w, (uint16_t)(v.w >> 16)};
#pragma unroll
      for (int dx = 0; dx < 1 + 2 * HALO; dx++) {
        const int p = c - dx;
        if (p >= 0 && p < KC) {
#pragma unroll
          for (int j = 0; j < 8; j++) s_x[(((1 + 2 * HALO) * r + dx) * CIN + 8 * ck + j) * KC + p] = e[j];
        }
      }
    }
    __syncthreads();
    if (tid < COUT) {
      float s = 0.f;
      for (int p = 0; p < KC; p++) s += bf16_val(s_dy[tid * KC + p]);
      bsum += s;
    }
#pragma unroll
    for (int i = 0; i < TPW; i++) {
      const int tile = wave * TPW + i;
      const int t = tile / ((COUT / 16) * (CIN / 16));
      const int mt = (tile / (CIN / 16)) % (COUT / 16);
      const int nt = tile % (CIN / 16);
      const bf16x8 av = as_bf16x8(*reinterpret_cast<const uint4*>(s_dy + (16 * mt + (l & 15)) * KC + 8 * (l >> 4)));
      const bf16x8 bv = as_bf16x8(*reinterpret_cast<const uint4*>(s_x + (t * CIN + 16 * nt + (l & 15)) * KC + 8 * (l >> 4)));
      acc[i] = mfma16(av, bv, acc[i]);
    }
  }
  float* part = a.partials + (size_t)blockIdx.x * (TAPS * COUT * CIN + COUT);
#pragma unroll
  for (int i = 0; i < TPW; i++) {
    const int tile = wave * TPW + i;
    const int t = tile / ((COUT / 16) * (CIN / 16));
    const int mt = (tile / (CIN / 16)) % (COUT / 16);
    const int nt = tile % (CIN / 16);
#pragma unroll
    for (int e = 0; e < 4; e++) {
      const int co = 16 * mt + 4 * (l >> 4) + e, ci = 16 * nt + (l & 15);
      part[(t * COUT + co) * CIN + ci] = acc[i][e];
    }
  }
  if (tid < COUT) part[TAPS * COUT * CIN + tid] = bsum;
}

// Sum `slots` partial vectors of length n (slot-major) into out (+= when
// accumulate), in slot order.
__global__ void reduce_partials_kernel(const float* partials, int slots, int n, float* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float s = 0.f;
  for (int k = 0; k < slots; k++) s += partials[(size_t)k * n + i];
  out[i] = s;
}

// ---------------------------------------------------------------- head
// Per cell: pred = W5 a4 + b5 (fp32 from bf16 activations).  Inference
// writes the rounded, clamped vector into the alignment table; training
// adds the Huber loss against label / 8, the gradient of a4's
// pre-activation (bf16) and per-workgroup partial sums
// {dW5[2][32], db5[2], loss}.
__device__ __forceinline__ float wave_sum_f(float v) {
#pragma unroll
  for (int m = 1; m < 64; m <<= 1) v += __shfl_xor(v, m, 64);
  return v;
}

__global__ void __launch_bounds__(256) head_kernel(const HeadArgs a) {
  __shared__ float s_red[4 * kHeadPartials];
  const int tid = threadIdx.x;
  const long long cell = (long long)blockIdx.x * 256 + tid;
  const long long cells = (long long)a.images * a.cells;
  const bool live = cell < cells;
  float gw[2][32];
  float g[2] = {0.f, 0.f}, loss = 0.f;
#pragma unroll
  for (int o = 0; o < 2; o++)
#pragma unroll
    for (int c = 0; c < 32; c++) gw[o][c] = 0.f;
  if (live) {
    float x[32];
    const uint4* src = reinterpret_cast<const uint4*>(a.a4 + cell * 32);
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const uint4 v = src[q];
      const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
      for (int j = 0; j < 4; j++) {
        x[8 * q + 2 * j] = bf16_val((uint16_t)w[j]);
        x[8 * q + 2 * j + 1] = bf16_val((uint16_t)(w[j] >> 16));
      }
    }
    float pred[2];
#pragma unroll
    for (int o = 0; o < 2; o++) {
      float s = a.params[64 + o];
#pragma unroll
      for (int c = 0; c < 32; c++) s += a.params[32 * o + c] * x[c];
      pred[o] = s;
    }
    if (a.table) {
      const int img = (int)(cell / a.cells), rest = (int)(cell - (long long)img * a.cells);
      int32_t* t = a.table + 3 * ((size_t)a.frame_of_image[img] * a.cells + rest);
#pragma unroll
      for (int o = 0; o < 2; o++) {
        float mv = rintf(pred[o] * kMvScale);
        mv = fminf(fmaxf(mv, (float)-a.clamp), (float)a.clamp);
        t[o] = (int32_t)mv;
      }
      t[2] = 0;
    }
    // A label whose mvx is kIgnoreLabel contributes no loss and no gradient.
    if (a.labels && a.labels[2 * cell] != kIgnoreLabel) {
#pragma unroll
      for (int o = 0; o < 2; o++) {
        const float d = pred[o] - (float)a.labels[2 * cell + o] * (1.0f / kMvScale);
        const float ad = fabsf(d);
        loss += (ad < kHuberDelta ? 0.5f * d * d / kHuberDelta : ad - 0.5f * kHuberDelta) * a.inv_count;
        g[o] = (ad < kHuberDelta ? d / kHuberDelta : (d > 0.f ? 1.f : -1.f)) * a.inv_count;
      }
      uint4 out[4];
      uint32_t* ow = reinterpret_cast<uint32_t*>(out);
#pragma unroll
      for (int c = 0; c < 32; c += 2) {
        float v0 = a.params[c] * g[0] + a.params[32 + c] * g[1];
        float v1 = a.params[c + 1] * g[0] + a.params[32 + c + 1] * g[1];
        if (!(x[c] > 0.f)) v0 *= kLeak;
        if (!(x[c + 1] > 0.f)) v1 *= kLeak;
        ow[c / 2] = (uint32_t)bf16_bits(v0) | ((uint32_t)bf16_bits(v1) << 16);
      }
      uint4* dst = reinterpret_cast<uint4*>(a.g4 + cell * 32);
#pragma unroll
      for (int q = 0; q < 4; q++) dst[q] = out[q];
#pragma unroll
      for (int o = 0; o < 2; o++)
#pragma unroll
        for (int c = 0; c < 32; c++) gw[o][c] = g[o] * x[c];
    } else if (a.labels) {
      uint4* dst = reinterpret_cast<uint4*>(a.g4 + cell * 32);
#pragma unroll
      for (int q = 0; q < 4; q++) dst[q] = make_uint4(0, 0, 0, 0);
    }
  }
  if (!a.labels) return;
  // Workgroup sums of the 67 training statistics: wave sums, then the 4
  // waves in order.
  const int wave = tid >> 6, lane = tid & 63;
#pragma unroll
  for (int k = 0; k < kHeadPartials; k++) {
    const float v = wave_sum_f(k < 64 ? gw[k >> 5][k & 31] : (k < 66 ? g[k - 64] : loss));
    if (lane == 0) s_red[wave * kHeadPartials + k] = v;
  }
  __syncthreads();
  if (tid < kHeadPartials) {
    float* part = a.partials + (size_t)blockIdx.x * kHeadPartials;
    part[tid] = ((s_red[tid] + s_red[kHeadPartials + tid]) + s_red[2 * kHeadPartials + tid]) +
                s_red[3 * kHeadPartials + tid];
  }
}

// ---------------------------------------------------------------- Adam
// params, grads, m, v: fp32 [n]; bf16 conv weights are rewritten from the
// new masters: forward [9][COUT][CIN] and backward-data [9][CIN][COUT] with
// the tap flipped (t -> 8 - t).
__global__ void adam_kernel(const AdamArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const float g = a.grads[i];
  const float m = a.b1 * a.m[i] + (1.f - a.b1) * g;
  const float v = a.b2 * a.v[i] + (1.f - a.b2) * g * g;
  a.m[i] = m;
  a.v[i] = v;
  const float p = a.params[i] - a.lr * (m * a.c1) / (sqrtf(v * a.c2) + a.eps);
  a.params[i] = p;
}

__global__ void pack_weights_kernel(const PackArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = a.taps * a.cout * a.cin;
  if (i >= n) return;
  const int t = i / (a.cout * a.cin), r = i - t * a.cout * a.cin;
  const int co = r / a.cin, ci = r - co * a.cin;
  const uint16_t b = bf16_bits(a.master[i]);
  a.fwd[i] = b;
  if (a.bwd) a.bwd[((a.taps - 1 - t) * a.cin + ci) * a.cout + co] = b;
}

// ---------------------------------------------------------------- launchers
static int persistent_grid(int total, int per_cu, int cus) {
  const int g = per_cu * cus;
  return total < g ? total : g;
}

hipError_t launch_prep_downsample(const PrepArgs& a, hipStream_t s) {
  const long long quads = (long long)a.cw * (4 * a.ch);  // per plane
  if (quads == 0 || a.pairs == 0) return hipSuccess;
  const unsigned planes = a.shared_anchor ? (unsigned)(1 + a.pairs) : (unsigned)(2 * a.pairs);
  hipLaunchKernelGGL(prep_downsample_kernel, dim3((unsigned)((quads + 255) / 256), planes), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_features(const FeatArgs& a, hipStream_t s) {
  if (a.images == 0 || a.cw == 0 || a.ch == 0) return hipSuccess;
  hipLaunchKernelGGL(features_kernel, dim3((unsigned)((a.cw + kFeatSeg - 1) / kFeatSeg), (unsigned)a.ch,
                                           (unsigned)a.images),
                     dim3(320), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_conv(const ConvArgs& a, int taps, int cin, int cout, int epi, int cus, hipStream_t s) {
  const int total = a.images * a.ch * ((a.cw + 63) / 64);
  if (total == 0) return hipSuccess;
  // one tile per workgroup up to 8 workgroups per CU, then persistent
  const dim3 grid(persistent_grid(total, 8, cus)), block(256);
#define VP8G_CONV(T, CI, CO, E)                                                              \
  if (taps == T && cin == CI && cout == CO && epi == E) {                                    \
    hipLaunchKernelGGL((conv_kernel<T, CI, CO, E>), grid, block, 0, s, a);                   \
    return hipGetLastError();                                                                \
  }
  VP8G_CONV(1, kFeat, 32, EPI_BIAS_RELU)
  VP8G_CONV(9, 32, 32, EPI_BIAS_RELU)
  VP8G_CONV(9, 32, 64, EPI_BIAS_RELU)
  VP8G_CONV(9, 64, 32, EPI_BIAS_RELU)
  VP8G_CONV(9, 32, 32, EPI_MASK)
  VP8G_CONV(9, 64, 32, EPI_MASK)
  VP8G_CONV(9, 32, 64, EPI_MASK)
#undef VP8G_CONV
  return hipErrorInvalidValue;
}

hipError_t launch_conv_stack(const ConvStackArgs& a, hipStream_t s) {
  const long long tiles = (long long)a.images * ((a.ch + 15) / 16) * ((a.cw + 15) / 16);
  if (tiles == 0) return hipSuccess;
  // a0's bytes must fit the 32-bit buffer offsets; at most 64 images
  if ((long long)a.images * a.ch * a.cw * 64 > 0x7FFFFFFFLL || a.images > 64 || tiles > 0x7FFFFFFF)
    return hipErrorInvalidValue;
  hipLaunchKernelGGL(conv_stack_kernel, dim3((unsigned)tiles), dim3(stack::kThreads), 0, s, a);
  return hipGetLastError();
}

int wgrad_slots(const WgradArgs& a, int cus) {
  const int total = a.images * a.ch * ((a.cw + 31) / 32);
  return persistent_grid(total, 2, cus);
}

hipError_t launch_wgrad(const WgradArgs& a, int taps, int cin, int cout, int slots, hipStream_t s) {
  if (slots <= 0) return hipSuccess;
  const dim3 grid(slots), block(256);
  if (taps == 1 && cin == kFeat && cout == 32) hipLaunchKernelGGL((wgrad_kernel<1, kFeat, 32>), grid, block, 0, s, a);
  else if (taps == 9 && cin == 32 && cout == 32) hipLaunchKernelGGL((wgrad_kernel<9, 32, 32>), grid, block, 0, s, a);
  else if (taps == 9 && cin == 32 && cout == 64) hipLaunchKernelGGL((wgrad_kernel<9, 32, 64>), grid, block, 0, s, a);
  else if (taps == 9 && cin == 64 && cout == 32) hipLaunchKernelGGL((wgrad_kernel<9, 64, 32>), grid, block, 0, s, a);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_reduce_partials(const float* partials, int slots, int n, float* out, hipStream_t s) {
  hipLaunchKernelGGL(reduce_partials_kernel, dim3((n + 255) / 256), dim3(256), 0, s, partials, slots, n, out);
  return hipGetLastError();
}

int head_blocks(long long cells) { return (int)((cells + 255) / 256); }

hipError_t launch_head(const HeadArgs& a, hipStream_t s) {
  const long long cells = (long long)a.images * a.cells;
  if (cells == 0) return hipSuccess;
  hipLaunchKernelGGL(head_kernel, dim3(head_blocks(cells)), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_adam(const AdamArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(adam_kernel, dim3((a.n + 255) / 256), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_pack_weights(const PackArgs& a, hipStream_t s) {
  const int n = a.taps * a.cout * a.cin;
  hipLaunchKernelGGL(pack_weights_kernel, dim3((n + 255) / 256), dim3(256), 0, s, a);
  return hipGetLastError();
}

}  // namespace vp8g
