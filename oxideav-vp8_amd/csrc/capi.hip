// C ABI (include/vp8g.h): argument validation in the reference's order,
// device staging for the host-pointer entry points, kernel dispatch.
#include <dlfcn.h>
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <deque>

#include <mutex>
#include <string>
#include <vector>

#include "../../include/vp8g.h"
#include "tuning.h"
#include "vp8g_ctx.h"
#include "vp8g_internal.h"

using namespace vp8g;

// Scratch of vp8g_search_fallback_noalloc_dev: the queue counter in the
// first 256 bytes, then up to kNoAllocQueueItems queued (frame, block)
// items.  One copy per GPU, held until the process ends and shared by every
// call on that GPU.  Host code never passes the array's name to a kernel:
// it uses the device address hipGetSymbolAddress returns for the current
// device (noalloc_scratch_address).
namespace {
constexpr size_t kNoAllocHeaderWords = 64;
constexpr size_t kNoAllocQueueItems = size_t(1) << 20;
constexpr int kNoAllocMaxDevices = 64;
}  // namespace
static __device__ __align__(256) uint32_t scratch[kNoAllocHeaderWords + kNoAllocQueueItems];

namespace {
std::mutex g_scratch_mutex;
void* g_scratch_by_device[kNoAllocMaxDevices] = {};

// The device address of `scratch` on `device` (which must be current),
// looked up once per device; nullptr if the lookup fails.
void* noalloc_scratch_address(int device) {
  if (device < 0 || device >= kNoAllocMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_by_device[device]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess) return nullptr;
    g_scratch_by_device[device] = p;
  }
  return g_scratch_by_device[device];
}
}  // namespace

struct vp8g_arnr_batch {
  vp8g_ctx* ctx = nullptr;
  uint8_t* dev = nullptr;  // jobs | frames | unit_job | tickets
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int mbs = 0, waves = 0, workgroups = 0;
  size_t units = 0;
};

namespace {

bool hip_ok(hipError_t e) { return e == hipSuccess; }

// Host-side fault injection (VP8G_OPT_FAULT / $VP8G_FAULT): the chosen step
// reports a HIP failure without running, so the error path is exercised
// with no faulty device work.
template <class F>
int launch_status(const vp8g_ctx* c, F&& launch) {
  if (c->fault == VP8G_FAULT_LAUNCH) return VP8G_E_DEVICE;
  return launch() == hipSuccess ? VP8G_OK : VP8G_E_DEVICE;
}

int fault_from_env() {
  const char* e = getenv("VP8G_FAULT");
  if (!e) return VP8G_FAULT_NONE;
  if (!strcmp(e, "alloc")) return VP8G_FAULT_ALLOC;
  if (!strcmp(e, "copy")) return VP8G_FAULT_COPY;
  if (!strcmp(e, "launch")) return VP8G_FAULT_LAUNCH;
  return VP8G_FAULT_NONE;
}

int ensure_arena(vp8g_ctx* c, size_t bytes) {
  if (c->fault == VP8G_FAULT_ALLOC) return VP8G_E_DEVICE;
  if (bytes <= c->arena_bytes) return VP8G_OK;
  if (c->arena) {
    if (!hip_ok(hipStreamSynchronize(c->stream))) return VP8G_E_DEVICE;
    (void)hipFree(c->arena);
    (void)hipHostFree(c->pinned);
    c->arena = nullptr;
    c->pinned = nullptr;
    c->arena_bytes = 0;
  }
  size_t want = bytes + (bytes >> 2);
  if (!hip_ok(hipMalloc(&c->arena, want))) return VP8G_E_DEVICE;
  if (!hip_ok(hipHostMalloc(reinterpret_cast<void**>(&c->pinned), want, hipHostMallocDefault))) {
    (void)hipFree(c->arena);
    c->arena = nullptr;
    return VP8G_E_DEVICE;
  }
  c->arena_bytes = want;
  return VP8G_OK;
}

size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

int ensure_work(vp8g_ctx* c, size_t bytes) {
  if (c->fault == VP8G_FAULT_ALLOC) return VP8G_E_DEVICE;
  if (bytes <= c->work_bytes) return VP8G_OK;
  if (c->work) {
    if (!hip_ok(hipStreamSynchronize(c->stream))) return VP8G_E_DEVICE;
    (void)hipFree(c->work);
    c->work = nullptr;
    c->work_bytes = 0;
  }
  const size_t want = bytes + (bytes >> 2);
  if (!hip_ok(hipMalloc(&c->work, want))) return VP8G_E_DEVICE;
  c->work_bytes = want;
  return VP8G_OK;
}

// Search arguments of a validated group (frames, geometry, seeds, output).
SearchArgs make_search_args(const vp8g_i420* frames, size_t n, size_t center, const int32_t* seeds, int32_t* out) {
  SearchArgs a;
  memset(&a, 0, sizeof(a));
  bool aligned4 = true;
  for (size_t i = 0; i < n; i++) {
    const vp8g_i420& f = frames[i];
    a.frames[i] = FrameDesc{f.y, f.u, f.v, (uint32_t)f.y_stride, (uint32_t)f.uv_stride};
    if (((uintptr_t)f.y & 3) || (f.y_stride & 3)) aligned4 = false;
  }
  a.seeds = seeds;
  a.out = out;
  a.w = (int)frames[center].width;
  a.h = (int)frames[center].height;
  a.bcols = (a.w + 15) / 16;
  a.brows = (a.h + 15) / 16;
  a.n = (int)n;
  a.center = (int)center;
  a.aligned4 = aligned4 ? 1 : 0;
  a.total_blocks = (int)(n - 1) * a.bcols * a.brows;
  return a;
}

// Rows of `w` bytes from a strided host plane into a packed run of pitch `p`.
void pack_plane(uint8_t* dst, size_t p, const uint8_t* src, size_t stride, size_t w, size_t h) {
  if (stride == p && w == p) {
    memcpy(dst, src, p * h);
    return;
  }
  for (size_t r = 0; r < h; r++) memcpy(dst + r * p, src + r * stride, w);
}

// Reference check order, src/arnr.rs:494-515.
int validate_arnr(const vp8g_i420* frames, size_t n, size_t center, vp8g_err* err) {
  if (!frames && n) return VP8G_E_ARG;
  if (center >= n) return VP8G_E_CENTER_OUT_OF_RANGE;
  const vp8g_i420& c = frames[center];
  if (c.width == 0 || c.height == 0 || c.width > 0x3FFF || c.height > 0x3FFF) {
    if (err) *err = vp8g_err{c.width, c.height, 0, 0};
    return VP8G_E_INVALID_DIMENSIONS;
  }
  for (size_t i = 0; i < n; i++) {
    if (frames[i].width != c.width || frames[i].height != c.height) {
      if (err) *err = vp8g_err{c.width, c.height, frames[i].width, frames[i].height};
      return VP8G_E_REF_DIMS_MISMATCH;
    }
  }
  if (n > (size_t)VP8G_MAX_FRAMES) return VP8G_E_ARG;
  const size_t cw = (c.width + 1) / 2;
  for (size_t i = 0; i < n; i++) {
    const vp8g_i420& f = frames[i];
    if (!f.y || !f.u || !f.v || f.y_stride < c.width || f.uv_stride < cw) return VP8G_E_ARG;
    if (f.y_stride > 0xFFFFFFFFu || f.uv_stride > 0xFFFFFFFFu) return VP8G_E_ARG;
  }
  return VP8G_OK;
}

// Validation shared by the fallback entry points (reference order first).
int check_fallback(const vp8g_i420* frames, size_t n, size_t center, const int32_t* first, int radius, int flags,
                   const int32_t* out, const uint32_t* queued, vp8g_err* err) {
  const int st = validate_arnr(frames, n, center, err);
  if (st != VP8G_OK) return st;
  if (!first || !out || first == out || (radius != 16 && radius != 64)) return VP8G_E_ARG;
  if ((flags & ~VP8G_FALLBACK_GATE_NEIGHBOUR) != 0) return VP8G_E_ARG;
  if (((uintptr_t)first | (uintptr_t)out | (uintptr_t)queued) & 3) return VP8G_E_ARG;
  return VP8G_OK;
}

// Enqueue the fallback with its queue and counter at `scratch` (device,
// 256-byte aligned, 256 + 4 * total_blocks bytes).
int enqueue_fallback(vp8g_ctx* c, const vp8g_i420* frames, size_t n, size_t center, const int32_t* first, int radius,
                     int flags, int32_t* out, uint32_t* queued, uint8_t* scratch) {
  const SearchArgs a = make_search_args(frames, n, center, nullptr, nullptr);
  FallbackArgs f;
  memset(&f, 0, sizeof(f));
  f.first = first;
  f.out = out;
  f.count = reinterpret_cast<uint32_t*>(scratch);
  f.queue = reinterpret_cast<uint32_t*>(scratch + 256);
  f.w = a.w;
  f.h = a.h;
  f.bcols = a.bcols;
  f.brows = a.brows;
  f.n = a.n;
  f.center = a.center;
  f.total_blocks = a.total_blocks;
  f.gate = (flags & VP8G_FALLBACK_GATE_NEIGHBOUR) ? 1 : 0;
  f.queued = queued;
  return launch_status(c, [&] { return launch_search_fallback(a, f, radius, c->cus, c->stream); });
}

// Shrink (M, K) until M * K <= 16 waves and the workgroup's LDS fits:
// M first, then K.
void fit_shape(int* k, int* m, bool table) {
  while (*m > 1 && (*m * *k > 16 || arnr_lds_bytes(*k, *m, table) > kMaxLdsPerWorkgroup))
    *m = arnr_supported_mbs(*m - 1);
  while (*k > 1 && (*m * *k > 16 || arnr_lds_bytes(*k, *m, table) > kMaxLdsPerWorkgroup)) (*k)--;
}

// Launch shape of one group: explicit options first, then the tuning
// table's entry for (resolution class, group length), then the defaults:
// 4 MBs per workgroup, and K waves per MB (frame splits) such that the
// launch holds about 4096 waves, at most one per neighbour.
void pick_arnr_shape(const vp8g_ctx* c, int w, int h, int n, int nnb, int* k_out, int* m_out) {
  const TuneEntry* e = lookup_tuning(c->tuning, resolution_class(w, h), n);
  const int mbs = ((w + 15) / 16) * ((h + 15) / 16);
  int k = c->arnr_waves > 0 ? c->arnr_waves : (e && e->waves > 0 ? e->waves : (4096 + mbs - 1) / mbs);
  if (k > nnb) k = nnb;
  if (k < 1) k = 1;
  if (k > 16) k = 16;
  int m = arnr_supported_mbs(c->arnr_mbs > 0 ? c->arnr_mbs : (e && e->mbs > 0 ? e->mbs : 4));
  fit_shape(&k, &m, false);
  *k_out = k;
  *m_out = m;
}

// Batch launch shape: options, then the table's "batch" entry, then one
// wave per MB and 4 MBs per workgroup; same limits.
void pick_batch_shape(const vp8g_ctx* c, int* k_out, int* m_out, int* wgs_per_cu) {
  const TuneEntry* e = lookup_tuning(c->tuning, "batch", 0);
  int k = c->arnr_waves > 0 ? c->arnr_waves : (e && e->waves > 0 ? e->waves : 1);
  int m = arnr_supported_mbs(c->arnr_mbs > 0 ? c->arnr_mbs : (e && e->mbs > 0 ? e->mbs : 4));
  fit_shape(&k, &m, false);
  *k_out = k;
  *m_out = m;
  *wgs_per_cu = c->batch_wgs_per_cu > 0 ? c->batch_wgs_per_cu : (e ? e->wgs_per_cu : 0);
}

// The table next to the library (tuning/gfx950.json), or $VP8G_TUNING
// (empty: none).
std::string default_tuning_path() {
  const char* env = getenv("VP8G_TUNING");
  if (env) return env;
  Dl_info info;
  if (!dladdr(reinterpret_cast<void*>(&default_tuning_path), &info) || !info.dli_fname) return "";
  std::string p = info.dli_fname;
  const size_t slash = p.rfind('/');
  p = (slash == std::string::npos ? std::string(".") : p.substr(0, slash)) + "/tuning/gfx950.json";
  return p;
}

void batch_free(vp8g_arnr_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->dev) (void)hipFree(b->dev);
  delete b;
}

// Fill the kernel descriptor of one validated group: frame descriptors
// into fd[0..n), the group geometry, the weight LUT.
void fill_group(const vp8g_i420* frames, size_t n, size_t center, uint8_t strength, uint8_t* oy,
                uint8_t* ou, uint8_t* ov, int32_t* mv_out, FrameDesc* fd, ArnrGroup* g,
                uint32_t wt4[64]) {
  memset(g, 0, sizeof(*g));
  const vp8g_i420& ctr = frames[center];
  bool aligned4 = true, aligned16 = true, aligned4c = true;
  const size_t rows = ctr.height, crows = (rows + 1) / 2;
  bool uniform_strides = rows * frames[0].y_stride < (1u << 31) && crows * frames[0].uv_stride < (1u << 31);
  for (size_t i = 0; i < n; i++) {
    const vp8g_i420& f = frames[i];
    fd[i] = FrameDesc{f.y, f.u, f.v, (uint32_t)f.y_stride, (uint32_t)f.uv_stride};
    if (((uintptr_t)f.y & 3) || (f.y_stride & 3)) aligned4 = false;
    if (((uintptr_t)f.y & 15) || (f.y_stride & 15)) aligned16 = false;
    if (((uintptr_t)f.u & 3) || ((uintptr_t)f.v & 3) || (f.uv_stride & 3)) aligned4c = false;
    if (f.y_stride != frames[0].y_stride || f.uv_stride != frames[0].uv_stride) uniform_strides = false;
  }
  g->out_y = oy;
  g->out_u = ou;
  g->out_v = ov;
  g->mv_out = mv_out;
  g->w = (int)ctr.width;
  g->h = (int)ctr.height;
  g->cw = (g->w + 1) / 2;
  g->ch = (g->h + 1) / 2;
  g->bcols = (g->w + 15) / 16;
  g->brows = (g->h + 15) / 16;
  g->n = (int)n;
  g->center = (int)center;
  g->filter_on = (strength != 0 && n >= 2) ? 1 : 0;
  g->aligned4 = aligned4 ? 1 : 0;
  g->aligned16 = aligned16 ? 1 : 0;
  g->aligned4c = aligned4c ? 1 : 0;
  g->uniform_strides = uniform_strides ? 1 : 0;
  // weight_table(8 << min(strength, 6)), src/arnr.rs:307-323, 533-536.
  const uint32_t s = 8u << (strength < kMaxStrength ? strength : kMaxStrength);
  uint8_t wt[256];
  for (uint32_t d = 0; d < 256; d++) wt[d] = (uint8_t)((kWMax * s) / (s + d * d));
  memcpy(wt4, wt, sizeof(wt));
}

// ceil(2^32 / d) for d >= 2, 0 for d == 1.  umulhi(n, ceil(2^32 / d)) is
// n / d + n * e / 2^32 rounded down, e < 1, so it equals n / d whenever
// n * d < 2^32: every strip index (< 2^20) of every row width (<= 1024
// strips) qualifies.
uint32_t strip_divider(int d) {
  return d <= 1 ? 0u : (uint32_t)((((uint64_t)1 << 32) + (uint64_t)d - 1) / (uint64_t)d);
}

// Enqueue the fused ARNR kernel over device-resident frames.
// MB rows [row0, row1) of the group (all of them by default).
int enqueue_arnr(vp8g_ctx* c, const vp8g_i420* frames, size_t n, size_t center, uint8_t strength,
                 uint8_t* oy, uint8_t* ou, uint8_t* ov, int32_t* mv_out, int row0 = 0, int row1 = -1,
                 hipStream_t stream = nullptr) {
  ArnrArgs a;
  memset(&a, 0, sizeof(a));
  fill_group(frames, n, center, strength, oy, ou, ov, mv_out, a.frames, &a.g, a.wt4);
  a.xcd_remap = c->xcd_remap;
  int k, m;
  pick_arnr_shape(c, a.g.w, a.g.h, (int)n, a.g.filter_on ? (int)n - 1 : 1, &k, &m);
  a.mbs_per_wg = m;
  a.g.strips_per_row = (a.g.bcols + m - 1) / m;
  a.spr_mul = strip_divider(a.g.strips_per_row);
  if (row1 < 0) row1 = a.g.brows;
  a.strip_base = row0 * a.g.strips_per_row;
  a.strip_count = (row1 - row0) * a.g.strips_per_row;
  return launch_status(c, [&] { return launch_arnr_fused(a, k, stream ? stream : c->stream); });
}

// Caller tap tables whose taps keep every six-tap sum inside int32.
bool taps_valid(const int32_t filters[8][6]) {
  for (int r = 0; r < 8; r++)
    for (int k = 0; k < 6; k++)
      if (filters[r][k] < -(1 << 20) || filters[r][k] > (1 << 20)) return false;
  return true;
}

int check_ref_planes(const vp8g_ref_planes* r) {
  if (!r || !r->y || !r->u || !r->v || r->mb_cols == 0 || r->mb_rows == 0) return VP8G_E_ARG;
  if (r->y_stride < r->mb_cols * 16 || r->uv_stride < r->mb_cols * 8) return VP8G_E_ARG;
  if (r->mb_cols * 16 > 0x7FFFFFF || r->mb_rows * 16 > 0x7FFFFFF) return VP8G_E_ARG;
  if (r->y_stride > 0xFFFFFFFFu || r->uv_stride > 0xFFFFFFFFu) return VP8G_E_ARG;
  return VP8G_OK;
}

int check_luma_ref(const vp8g_luma_ref* r) {
  if (!r || !r->plane || r->width == 0 || r->height == 0 || r->stride < r->width) return VP8G_E_ARG;
  if (r->width > 0x7FFFFFF || r->height > 0x7FFFFFF || r->stride > 0xFFFFFFFFu) return VP8G_E_ARG;
  return VP8G_OK;
}

McArgs make_mc_args(const vp8g_ref_planes* r, const uint32_t* mb_xy, const vp8g_mv* mv, size_t n,
                    int full_pixel, const int32_t filters[8][6], uint8_t* out) {
  McArgs a;
  memset(&a, 0, sizeof(a));
  a.ref = McPlanes{r->y, r->u, r->v, (uint32_t)r->y_stride, (uint32_t)r->uv_stride,
                   (int)(r->mb_cols * 16), (int)(r->mb_rows * 16), (int)(r->mb_cols * 8),
                   (int)(r->mb_rows * 8)};
  a.mb_xy = mb_xy;
  a.mv = reinterpret_cast<const int16_t*>(mv);
  a.out = out;
  a.n = (int)n;
  a.full_pixel = full_pixel ? 1 : 0;
  memcpy(a.filters, filters, sizeof(a.filters));
  return a;
}

SadArgs make_sad_args(const vp8g_luma_ref* r, const uint32_t* mb_xy, const uint8_t* src,
                      const vp8g_mv* mv, size_t n, uint32_t* sad) {
  SadArgs a;
  a.plane = r->plane;
  a.stride = (uint32_t)r->stride;
  a.w = (int)r->width;
  a.h = (int)r->height;
  a.mb_xy = mb_xy;
  a.src = src;
  a.mv = reinterpret_cast<const int16_t*>(mv);
  a.sad = sad;
  a.n = (int)n;
  return a;
}

// Macroblock positions past the plane are legal (every fetch clamps, as in
// the reference); the kernels compute pixel origins in 32-bit, so positions
// are limited to 2^20 macroblocks per axis.
bool mb_xy_in_range(const uint32_t* mb_xy, size_t n) {
  for (size_t i = 0; i < 2 * n; i++)
    if (mb_xy[i] > 0xFFFFFu) return false;
  return true;
}

// Bytes of a strided plane that the kernels may read.
size_t plane_span(size_t stride, size_t width, size_t rows) { return rows ? (rows - 1) * stride + width : 0; }

// Host <-> arena copies through the pinned mirror (same offset as the
// device pointer in the arena).  Outputs reach the caller in finish().
int copy_out(vp8g_ctx* c, void* host, const void* dev, size_t bytes) {
  if (c->fault == VP8G_FAULT_COPY) return VP8G_E_DEVICE;
  if (!bytes) return VP8G_OK;
  const size_t off = static_cast<const uint8_t*>(dev) - c->arena;
  if (!hip_ok(hipMemcpyAsync(c->pinned + off, dev, bytes, hipMemcpyDeviceToHost, c->stream))) return VP8G_E_DEVICE;
  c->pending.push_back({host, off, bytes});
  return VP8G_OK;
}

int copy_in(vp8g_ctx* c, void* dev, const void* host, size_t bytes) {
  if (c->fault == VP8G_FAULT_COPY) return VP8G_E_DEVICE;
  if (!bytes) return VP8G_OK;
  const size_t off = static_cast<uint8_t*>(dev) - c->arena;
  memcpy(c->pinned + off, host, bytes);
  return hip_ok(hipMemcpyAsync(dev, c->pinned + off, bytes, hipMemcpyHostToDevice, c->stream)) ? VP8G_OK
                                                                                              : VP8G_E_DEVICE;
}

int finish(vp8g_ctx* c, int st) {
  const bool synced = hip_ok(hipStreamSynchronize(c->stream));
  if (st == VP8G_OK && !synced) st = VP8G_E_DEVICE;
  if (st == VP8G_OK)
    for (const auto& p : c->pending) memcpy(p.host, c->pinned + p.off, p.bytes);
  c->pending.clear();
  return st;
}

// finish() after work on the link stream: it is drained first, so no copy
// still reads the staging on return.
int finish_io(vp8g_ctx* c, int st) {
  if (c->link && !hip_ok(hipStreamSynchronize(c->link))) st = VP8G_E_DEVICE;
  return finish(c, st);
}

// The host threads, link stream and events of the host-pointer ARNR entry
// point, and `bytes` of write-combined pinned staging (grown, never shrunk).
int ensure_host_io(vp8g_ctx* c, size_t bytes) {
  if (c->fault == VP8G_FAULT_ALLOC) return VP8G_E_DEVICE;
  if (!c->pool) {
    const unsigned hw = std::thread::hardware_concurrency();
    const int def = (int)std::min(8u, std::max(1u, hw / 2));
    const int threads = c->host_threads > 0 ? c->host_threads : def;
    c->pool = new vp8g::HostPool(threads - 1);  // the calling thread works too
  }
  if (!c->link && !hip_ok(hipStreamCreateWithFlags(&c->link, hipStreamNonBlocking))) return VP8G_E_DEVICE;
  for (int b = 0; b < vp8g_ctx::kMaxBands; b++) {
    if (!c->band_in[b] && !hip_ok(hipEventCreateWithFlags(&c->band_in[b], hipEventDisableTiming)))
      return VP8G_E_DEVICE;
    if (!c->band_done[b] && !hip_ok(hipEventCreateWithFlags(&c->band_done[b], hipEventDisableTiming)))
      return VP8G_E_DEVICE;
  }
  if (bytes <= c->staging_bytes) return VP8G_OK;
  if (c->staging) {
    (void)hipHostFree(c->staging);
    c->staging = nullptr;
    c->staging_bytes = 0;
  }
  const size_t want = bytes + (bytes >> 2);
  if (!hip_ok(hipHostMalloc(reinterpret_cast<void**>(&c->staging), want, hipHostMallocWriteCombined)))
    return VP8G_E_DEVICE;
  c->staging_bytes = want;
  return VP8G_OK;
}

// Rows [r0, r1) of a plane of w-byte rows from stride `ss` to pitch `dp`.
void copy_rows(uint8_t* dst, size_t dp, const uint8_t* src, size_t ss, size_t w, size_t r0, size_t r1) {
  if (ss == dp && w == dp) {
    memcpy(dst + r0 * dp, src + r0 * ss, (r1 - r0) * dp);
    return;
  }
  for (size_t r = r0; r < r1; r++) memcpy(dst + r * dp, src + r * ss, w);
}

// Copies of up to 3 byte ranges between pinned host memory and device
// memory by a kernel, which reads or writes the host memory over PCIe.  The
// source and destination of a range share their alignment mod 16: the
// bytes before the first and after the last 16-byte boundary go one a
// lane, the rest 16 at a time.
struct LinkSeg {
  const uint8_t* src;
  uint8_t* dst;
  size_t bytes;
};
struct LinkArgs {
  LinkSeg seg[3];
  int segs;
};

__global__ void __launch_bounds__(256) link_copy_kernel(const LinkArgs a) {
  const size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x, lanes = (size_t)gridDim.x * blockDim.x;
  for (int k = 0; k < a.segs; k++) {
    const uint8_t* src = a.seg[k].src;
    uint8_t* dst = a.seg[k].dst;
    const size_t bytes = a.seg[k].bytes;
    const size_t head = min(bytes, (size_t)((16 - ((uintptr_t)src & 15)) & 15));
    const size_t body = (bytes - head) & ~(size_t)15, tail = bytes - head - body;
    if (tid < head) dst[tid] = src[tid];
    if (tid < tail) dst[head + body + tid] = src[head + body + tid];
    const uint4* s4 = reinterpret_cast<const uint4*>(src + head);
    uint4* d4 = reinterpret_cast<uint4*>(dst + head);
    for (size_t i = tid; i < body / 16; i += lanes) d4[i] = s4[i];
  }
}

// 256 workgroups keep enough transfers in flight to fill the link (about
// 52-56 GB/s measured either way, as the copy engines).
hipError_t launch_link_copy(const LinkArgs& a, hipStream_t s) {
  for (int k = 0; k < a.segs; k++)
    if ((((uintptr_t)a.seg[k].src ^ (uintptr_t)a.seg[k].dst) & 15) != 0) return hipErrorInvalidValue;
  hipLaunchKernelGGL(link_copy_kernel, dim3(256), dim3(256), 0, s, a);
  return hipGetLastError();
}

// The same ranges by the copy engines: they do not take CU time from the
// kernels that run meanwhile, and their completion releases a waiting
// stream within microseconds (a stream waiting on a copy kernel of another
// stream was measured to start only when some kernel ended).
hipError_t copy_engine_pull(const LinkArgs& a, hipStream_t s) {
  for (int k = 0; k < a.segs; k++) {
    if (!a.seg[k].bytes) continue;
    const hipError_t e = hipMemcpyAsync(a.seg[k].dst, a.seg[k].src, a.seg[k].bytes, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

// A plane copy cut into row bands of about 256 KiB, for the host threads.
struct PlaneJob {
  uint8_t* dst;
  size_t dp;
  const uint8_t* src;
  size_t ss, w, h, rows_per_band;
};

void add_plane(std::vector<std::pair<const PlaneJob*, size_t>>& tasks, const PlaneJob& j) {
  for (size_t r = 0; r < j.h; r += j.rows_per_band) tasks.push_back({&j, r});
}

PlaneJob plane_job(uint8_t* dst, size_t dp, const uint8_t* src, size_t ss, size_t w, size_t h) {
  const size_t band = std::max<size_t>(1, (256u << 10) / std::max<size_t>(1, w));
  return PlaneJob{dst, dp, src, ss, w, h, band};
}

void run_plane_tasks(vp8g_ctx* c, const std::vector<std::pair<const PlaneJob*, size_t>>& tasks) {
  c->pool->run(tasks.size(), [&](size_t i) {
    const PlaneJob& j = *tasks[i].first;
    const size_t r0 = tasks[i].second, r1 = std::min(j.h, r0 + j.rows_per_band);
    copy_rows(j.dst, j.dp, j.src, j.ss, j.w, r0, r1);
  });
}

}  // namespace

int vp8g_validate_group(const vp8g_i420* frames, size_t n, size_t center, vp8g_err* err) {
  return validate_arnr(frames, n, center, err);
}

extern "C" {

const char* vp8g_status_str(int s) {
  switch (s) {
    case VP8G_OK: return "ok";
    case VP8G_E_INVALID_DIMENSIONS: return "invalid dimensions";
    case VP8G_E_REF_DIMS_MISMATCH: return "reference dimensions mismatch";
    case VP8G_E_CENTER_OUT_OF_RANGE: return "center index out of range";
    case VP8G_E_ARG: return "invalid argument";
    case VP8G_E_DEVICE: return "HIP device error";
    case VP8G_E_NO_DEVICE: return "no usable device";
    default: return "unknown status";
  }
}

int vp8g_ctx_create(int device, vp8g_ctx** out) {
  if (!out) return VP8G_E_ARG;
  *out = nullptr;
  int count = 0;
  if (!hip_ok(hipGetDeviceCount(&count)) || count <= 0) return VP8G_E_NO_DEVICE;
  if (device < 0 || device >= count) return VP8G_E_NO_DEVICE;
  hipDeviceProp_t prop;
  if (!hip_ok(hipGetDeviceProperties(&prop, device))) return VP8G_E_NO_DEVICE;
  if (strncmp(prop.gcnArchName, "gfx950", 6) != 0) return VP8G_E_NO_DEVICE;
  if (!hip_ok(hipSetDevice(device))) return VP8G_E_DEVICE;
  vp8g_ctx* c = new vp8g_ctx();
  c->device = device;
  c->cus = prop.multiProcessorCount;
  // The device is current: look up this device's copy of the no-alloc
  // scratch (nullptr makes only vp8g_search_fallback_noalloc_dev fail).
  c->scratch = static_cast<uint8_t*>(noalloc_scratch_address(device));
  if (!hip_ok(hipStreamCreateWithFlags(&c->own, hipStreamNonBlocking))) {
    delete c;
    return VP8G_E_DEVICE;
  }
  c->stream = c->own;
  c->fault = fault_from_env();
  // A missing or unreadable table leaves the built-in defaults.
  const std::string tp = default_tuning_path();
  if (!tp.empty()) (void)load_tuning_file(tp.c_str(), &c->tuning);
  *out = c;
  return VP8G_OK;
}

int vp8g_ctx_load_tuning(vp8g_ctx* c, const char* path) {
  if (!c) return VP8G_E_ARG;
  if (!path) {
    c->tuning.clear();
    return VP8G_OK;
  }
  std::vector<TuneEntry> t;
  if (!load_tuning_file(path, &t)) return VP8G_E_ARG;
  c->tuning = t;
  return VP8G_OK;
}

int vp8g_tuning_lookup(const char* path, uint32_t width, uint32_t height, size_t n_frames,
                       int batch, int* mbs, int* waves, int* wgs_per_cu) {
  if (!path) return VP8G_E_ARG;
  std::vector<TuneEntry> t;
  if (!load_tuning_file(path, &t)) return VP8G_E_ARG;
  const TuneEntry* e = batch ? lookup_tuning(t, "batch", 0)
                             : lookup_tuning(t, resolution_class((int)width, (int)height), (int)n_frames);
  if (mbs) *mbs = e ? e->mbs : 0;
  if (waves) *waves = e ? e->waves : 0;
  if (wgs_per_cu) *wgs_per_cu = e ? e->wgs_per_cu : 0;
  return VP8G_OK;
}

int vp8g_arnr_shape(vp8g_ctx* c, uint32_t width, uint32_t height, size_t n_frames, uint8_t strength,
                    int* mbs, int* waves) {
  if (!c || width == 0 || height == 0 || n_frames == 0 || n_frames > VP8G_MAX_FRAMES) return VP8G_E_ARG;
  const int n = (int)n_frames;
  const bool on = strength != 0 && n >= 2;
  int k, m;
  pick_arnr_shape(c, (int)width, (int)height, n, on ? n - 1 : 1, &k, &m);
  if (mbs) *mbs = m;
  if (waves) *waves = k;
  return VP8G_OK;
}

void vp8g_ctx_destroy(vp8g_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->link) (void)hipStreamSynchronize(c->link);
  if (c->arena) (void)hipFree(c->arena);
  if (c->pinned) (void)hipHostFree(c->pinned);  // allocated with the arena
  if (c->work) (void)hipFree(c->work);
  if (c->link) (void)hipStreamDestroy(c->link);
  for (int b = 0; b < vp8g_ctx::kMaxBands; b++) {
    if (c->band_in[b]) (void)hipEventDestroy(c->band_in[b]);
    if (c->band_done[b]) (void)hipEventDestroy(c->band_done[b]);
  }
  if (c->staging) (void)hipHostFree(c->staging);
  if (c->lf_work) (void)hipFree(c->lf_work);
  if (c->lf_io) (void)hipFree(c->lf_io);
  if (c->lf_pinned) (void)hipHostFree(c->lf_pinned);
  if (c->lf_ssd) (void)hipFree(c->lf_ssd);
  if (c->lf_err) (void)hipFree(c->lf_err);
  delete c->pool;
  if (c->own) (void)hipStreamDestroy(c->own);
  delete c;
}

int vp8g_ctx_set_stream(vp8g_ctx* c, void* stream) {
  if (!c) return VP8G_E_ARG;
  c->stream = stream ? (hipStream_t)stream : c->own;
  return VP8G_OK;
}

void* vp8g_ctx_stream(vp8g_ctx* c) { return c ? (void*)c->stream : nullptr; }

int vp8g_ctx_synchronize(vp8g_ctx* c) {
  if (!c) return VP8G_E_ARG;
  return hip_ok(hipStreamSynchronize(c->stream)) ? VP8G_OK : VP8G_E_DEVICE;
}

int vp8g_ctx_set_option(vp8g_ctx* c, int option, int value) {
  if (!c) return VP8G_E_ARG;
  switch (option) {
    case VP8G_OPT_ARNR_WAVES:
      if (value < 0 || value > 16) return VP8G_E_ARG;
      c->arnr_waves = value;
      return VP8G_OK;
    case VP8G_OPT_XCD_REMAP:
      if (value < 0 || value > 1) return VP8G_E_ARG;
      c->xcd_remap = value;
      return VP8G_OK;
    case VP8G_OPT_ARNR_MBS:
      if (value < 0 || value > 8) return VP8G_E_ARG;
      c->arnr_mbs = value;
      return VP8G_OK;
    case VP8G_OPT_BATCH_WGS_PER_CU:
      if (value < 0 || value > 32) return VP8G_E_ARG;
      c->batch_wgs_per_cu = value;
      return VP8G_OK;
    case VP8G_OPT_FAULT:
      if (value < VP8G_FAULT_NONE || value > VP8G_FAULT_LAUNCH) return VP8G_E_ARG;
      c->fault = value;
      return VP8G_OK;
    case VP8G_OPT_HOST_THREADS:
      if (value < 0 || value > 64) return VP8G_E_ARG;
      if (value != c->host_threads) {
        delete c->pool;  // rebuilt with the new count on next use
        c->pool = nullptr;
      }
      c->host_threads = value;
      return VP8G_OK;
    case VP8G_OPT_COPY_STREAMS:
      if (value < 0 || value > vp8g_ctx::kMaxCopyStreams) return VP8G_E_ARG;
      c->copy_streams = value;
      return VP8G_OK;
    case VP8G_OPT_HOST_BANDS:
      if (value < 0 || value > vp8g_ctx::kMaxBands) return VP8G_E_ARG;
      c->host_bands = value;
      return VP8G_OK;
    default:
      return VP8G_E_ARG;
  }
}

int vp8g_build_arnr_altref_dev(vp8g_ctx* c, const vp8g_i420* frames, size_t n, size_t center,
                               uint8_t strength, uint8_t* oy, uint8_t* ou, uint8_t* ov,
                               vp8g_err* err, int32_t* mv_out) {
  if (!c) return VP8G_E_ARG;
  int st = validate_arnr(frames, n, center, err);
  if (st != VP8G_OK) return st;
  if (!oy || !ou || !ov) return VP8G_E_ARG;
  if (!hip_ok(hipSetDevice(c->device))) return VP8G_E_DEVICE;
  return enqueue_arnr(c, frames, n, center, strength, oy, ou, ov, mv_out);
}

int vp8g_build_arnr_altref(vp8g_ctx* c, const vp8g_i420* frames, size_t n, size_t center,
                           uint8_t strength, uint8_t* oy, uint8_t* ou, uint8_t* ov,
                           vp8g_err* err) {
  if (!c) return VP8G_E_ARG;
  int st = validate_arnr(frames, n, center, err);
  if (st != VP8G_OK) return st;
  if (!oy || !ou || !ov) return VP8G_E_ARG;
  if (!hip_ok(hipSetDevice(c->device))) return VP8G_E_DEVICE;
  const size_t w = frames[center].width, h = frames[center].height;
  const size_t cw = (w + 1) / 2, ch = (h + 1) / 2;
  // Device layout: the frames' rows interleaved, one plane of Y and one of
  // chroma for the whole group: row r of frame i's luma at (r n + i) yp, so
  // frame i's plane starts at i yp with stride n yp (pitches rounded to 64
  // bytes); chroma row r holds every frame's U row r, then every frame's V
  // row r (stride 2 n cp).  Rows a..b-1 of every frame are then one
  // contiguous range of each.  Then the packed outputs.
  const size_t yp = align_up(w, 64), cp = align_up(cw, 64);
  const size_t yrow = yp * n, crow = cp * n;  // bytes of one row of every frame
  const size_t ysz = align_up(yrow * h, 256), csz = align_up(2 * crow * ch, 256);
  const size_t in_bytes = ysz + csz;
  const size_t oyb = align_up(w * h, 256), ocb = align_up(cw * ch, 256);
  if (yrow > 0x7FFFFFFFu || ysz > 0x7FFFFFFFu || csz > 0x7FFFFFFFu) return VP8G_E_ARG;
  st = ensure_arena(c, in_bytes + oyb + 2 * ocb);
  if (st == VP8G_OK) st = ensure_host_io(c, in_bytes);
  if (st != VP8G_OK) return st;
  if (c->fault == VP8G_FAULT_COPY) return VP8G_E_DEVICE;
  // The link copies need not wait for the context's stream: the arena and
  // the staging are only used by the host-pointer entry points, which all
  // return synchronised, so no queued work reads them.
  uint8_t* const dY = c->arena;
  uint8_t* const dU = dY + ysz;
  uint8_t* const dV = dU + crow;
  uint8_t* const pY = c->staging;
  uint8_t* const pU = pY + ysz;
  uint8_t* const pV = pU + crow;
  std::vector<vp8g_i420> dev(n);
  for (size_t i = 0; i < n; i++)
    dev[i] = vp8g_i420{frames[i].width, frames[i].height, dY + i * yp, dU + i * cp, dV + i * cp, yrow, 2 * crow};
  uint8_t* dy = c->arena + in_bytes;
  uint8_t* du = dy + oyb;
  uint8_t* dv = du + ocb;
  // the pinned mirror of the outputs (same offsets as in the arena, so a
  // range and its mirror share their alignment)
  const size_t off = (size_t)(dy - c->arena);
  uint8_t* const hy = c->pinned + off;
  uint8_t* const hu = hy + oyb;
  uint8_t* const hv = hu + ocb;
  // Row bands.  The search of an MB row reaches 15 luma and 7 chroma rows
  // below it (and as many above), clamped to the frame, so band b (MB rows
  // r0..r1-1) can run once every frame's rows up to 16 r1 + 15 and 8 r1 + 7
  // are in.  Each band sends only the rows the previous ones did not, as
  // two pieces (its new luma rows, then its new chroma rows): host threads
  // pack a piece into the pinned staging, then the copy engines pull it to
  // the device planes over PCIe on the link stream.  The pulls run back to
  // back; band b's kernel runs on the context's stream once band_in[b] has
  // fired, writes the device outputs, and a copy kernel pushes its altref
  // rows into the pinned mirror (the link carries both directions at once).
  // The host threads pack piece p + 1 while the calling thread launches
  // piece p's work (the pieces' staging rows are disjoint), then it joins
  // them; after the last launch it copies each band's rows out as soon as
  // band_done[b] has fired.  The last band is made half as tall as the
  // others, since nothing overlaps its kernel, push and copy-out.
  const int brows = (int)((h + 15) / 16);
  const int B = std::max(1, std::min(brows, c->host_bands > 0 ? c->host_bands : 3));
  std::vector<int> edge(B + 1);  // band b is MB rows edge[b] .. edge[b + 1] - 1
  for (int b = 0; b <= B; b++) edge[b] = (int)((int64_t)brows * 2 * b / (2 * B - 1));
  edge[B] = brows;
  struct Piece {
    size_t t0, t1;  // its packing tasks
    LinkArgs pull;
    int band;       // the band whose kernel this piece completes, else -1
  };
  std::deque<PlaneJob> jobs;  // stable addresses for the tasks
  std::vector<std::pair<const PlaneJob*, size_t>> tasks;
  std::vector<Piece> pieces;
  // Rows [r0, r1) of every frame's luma or chroma as one piece (each copy
  // costs the copy engines about 10 us beyond its bytes, more than a finer
  // split saves); it completes `band` (-1: none).
  const auto add_rows = [&](bool luma, size_t r0, size_t r1, int band) {
    const size_t row = luma ? yrow : 2 * crow;
    Piece pc{tasks.size(), 0, LinkArgs{}, band};
    if (r1 > r0) {
      for (size_t i = 0; i < n; i++) {
        const vp8g_i420& f = frames[i];
        if (luma) {
          jobs.push_back(plane_job(pY + r0 * yrow + i * yp, yrow, f.y + r0 * f.y_stride, f.y_stride, w, r1 - r0));
          add_plane(tasks, jobs.back());
        } else {
          jobs.push_back(plane_job(pU + r0 * row + i * cp, row, f.u + r0 * f.uv_stride, f.uv_stride, cw, r1 - r0));
          add_plane(tasks, jobs.back());
          jobs.push_back(plane_job(pV + r0 * row + i * cp, row, f.v + r0 * f.uv_stride, f.uv_stride, cw, r1 - r0));
          add_plane(tasks, jobs.back());
        }
      }
      pc.pull.seg[0] = {(luma ? pY : pU) + r0 * row, (luma ? dY : dU) + r0 * row, (r1 - r0) * row};
      pc.pull.segs = 1;
    }
    pc.t1 = tasks.size();
    if (pc.pull.segs || band >= 0) pieces.push_back(pc);
  };
  size_t sent_y = 0, sent_c = 0;  // rows of every frame already packed
  for (int b = 0; b < B; b++) {
    const int r1 = edge[b + 1];
    if (r1 <= edge[b]) continue;
    const bool last = b == B - 1;
    const size_t need_y = last ? h : std::min(h, (size_t)16 * r1 + 15);
    const size_t need_c = last ? ch : std::min(ch, (size_t)8 * r1 + 7);
    if (need_y > sent_y) add_rows(true, sent_y, need_y, -1);
    add_rows(false, sent_c, need_c, b);
    sent_y = need_y, sent_c = need_c;
  }
  size_t base = 0;
  const std::function<void(size_t)> pack = [&](size_t i) {
    const PlaneJob& j = *tasks[base + i].first;
    const size_t q0 = tasks[base + i].second, q1 = std::min(j.h, q0 + j.rows_per_band);
    copy_rows(j.dst, j.dp, j.src, j.ss, j.w, q0, q1);
  };
  base = pieces[0].t0;
  c->pool->start(pieces[0].t1 - pieces[0].t0, pack);
  for (size_t k = 0; k < pieces.size(); k++) {
    c->pool->finish();  // piece k is packed
    const Piece& pk = pieces[k];
    if (k + 1 < pieces.size()) {
      base = pieces[k + 1].t0;
      c->pool->start(pieces[k + 1].t1 - pieces[k + 1].t0, pack);
    }
    if (pk.pull.segs) st = launch_status(c, [&] { return copy_engine_pull(pk.pull, c->link); });
    if (st == VP8G_OK && pk.band >= 0) {
      const int b = pk.band, r0 = edge[b], r1 = edge[b + 1];
      const size_t y0 = (size_t)16 * r0, y1 = std::min(h, (size_t)16 * r1);
      const size_t c0 = (size_t)8 * r0, c1 = std::min(ch, (size_t)8 * r1);
      LinkArgs push{};
      push.seg[0] = {dy + y0 * w, hy + y0 * w, (y1 - y0) * w};
      push.seg[1] = {du + c0 * cw, hu + c0 * cw, (c1 - c0) * cw};
      push.seg[2] = {dv + c0 * cw, hv + c0 * cw, (c1 - c0) * cw};
      push.segs = 3;
      // The last band's work follows its pull on the link stream itself:
      // nothing is left to overlap, and the wait on another stream's copy
      // costs more than the kernel.
      const bool tail = k + 1 == pieces.size();
      hipStream_t ks = tail ? c->link : c->stream;
      if (!tail && (!hip_ok(hipEventRecord(c->band_in[b], c->link)) ||
                    !hip_ok(hipStreamWaitEvent(c->stream, c->band_in[b], 0))))
        st = VP8G_E_DEVICE;
      if (st == VP8G_OK) st = enqueue_arnr(c, dev.data(), n, center, strength, dy, du, dv, nullptr, r0, r1, ks);
      if (st == VP8G_OK) st = launch_status(c, [&] { return launch_link_copy(push, ks); });
      if (st == VP8G_OK && !hip_ok(hipEventRecord(c->band_done[b], ks))) st = VP8G_E_DEVICE;
    }
    if (st != VP8G_OK) {
      c->pool->finish();  // no packing task may outlive this call
      return finish_io(c, st);
    }
  }
  // Each band's rows go to the caller as soon as they are back; if a later
  // band fails on the device, the rows of the bands before it are written.
  for (int b = 0; b < B; b++) {
    const int r0 = edge[b], r1 = edge[b + 1];
    if (r1 <= r0) continue;
    if (!hip_ok(hipEventSynchronize(c->band_done[b]))) return finish_io(c, VP8G_E_DEVICE);
    const size_t y0 = (size_t)16 * r0, y1 = std::min(h, (size_t)16 * r1);
    const size_t c0 = (size_t)8 * r0, c1 = std::min(ch, (size_t)8 * r1);
    const PlaneJob outs[3] = {plane_job(oy + y0 * w, w, hy + y0 * w, w, w, y1 - y0),
                              plane_job(ou + c0 * cw, cw, hu + c0 * cw, cw, cw, c1 - c0),
                              plane_job(ov + c0 * cw, cw, hv + c0 * cw, cw, cw, c1 - c0)};
    tasks.clear();
    for (const PlaneJob& j : outs) add_plane(tasks, j);
    run_plane_tasks(c, tasks);
  }
  // the link's last work was the last band's, waited for above
  return finish(c, VP8G_OK);
}

int vp8g_block_search_dev(vp8g_ctx* c, const vp8g_i420* frames, size_t n, size_t center,
                          int radius, const int32_t* seeds, int32_t* out, vp8g_err* err) {
  if (!c) return VP8G_E_ARG;
  int st = validate_arnr(frames, n, center, err);
  if (st != VP8G_OK) return st;
  if (!out || !block_search_supported(radius)) return VP8G_E_ARG;
  if (((uintptr_t)out | (uintptr_t)seeds) & 3) return VP8G_E_ARG;
  if (!hip_ok(hipSetDevice(c->device))) return VP8G_E_DEVICE;
  SearchArgs a;
  memset(&a, 0, sizeof(a));
  bool aligned4 = true;
  for (size_t i = 0; i < n; i++) {
    const vp8g_i420& f = frames[i];
    a.frames[i] = FrameDesc{f.y, f.u, f.v, (uint32_t)f.y_stride, (uint32_t)f.uv_stride};
    if (((uintptr_t)f.y & 3) || (f.y_stride & 3)) aligned4 = false;
  }
  a.seeds = seeds;
  a.out = out;
  a.w = (int)frames[center].width;
  a.h = (int)frames[center].height;
  a.bcols = (a.w + 15) / 16;
  a.brows = (a.h + 15) / 16;
  a.n = (int)n;
  a.center = (int)center;
  a.aligned4 = aligned4 ? 1 : 0;
  a.total_blocks = (int)(n - 1) * a.bcols * a.brows;
  return launch_status(c, [&] { return launch_block_search(a, radius, c->stream); });
}

int vp8g_search_fallback_dev(vp8g_ctx* c, const vp8g_i420* frames, size_t n, size_t center, const int32_t* first,
                             int radius, int flags, int32_t* out, uint32_t* queued, vp8g_err* err) {
  if (!c) return VP8G_E_ARG;
  int st = check_fallback(frames, n, center, first, radius, flags, out, queued, err);
  if (st != VP8G_OK) return st;
  if (!hip_ok(hipSetDevice(c->device))) return VP8G_E_DEVICE;
  const size_t per_frame = (size_t)((frames[center].width + 15) / 16) * ((frames[center].height + 15) / 16);
  st = ensure_work(c, 256 + 4 * (n - 1) * per_frame);
  if (st != VP8G_OK) return st;
  return enqueue_fallback(c, frames, n, center, first, radius, flags, out, queued, c->work);
}

int vp8g_search_fallback_noalloc_dev(vp8g_ctx* c, const vp8g_i420* frames, size_t n, size_t center,
                                     const int32_t* first, int radius, int flags, int32_t* out, uint32_t* queued,
                                     vp8g_err* err) {
  if (!c) return VP8G_E_ARG;
  int st = check_fallback(frames, n, center, first, radius, flags, out, queued, err);
  if (st != VP8G_OK) return st;
  const size_t per_frame = (size_t)((frames[center].width + 15) / 16) * ((frames[center].height + 15) / 16);
  if ((n - 1) * per_frame > kNoAllocQueueItems) return VP8G_E_ARG;
  if (!c->scratch) return VP8G_E_DEVICE;
  if (!hip_ok(hipSetDevice(c->device))) return VP8G_E_DEVICE;
  return enqueue_fallback(c, frames, n, center, first, radius, flags, out, queued, c->scratch);
}

int vp8g_refine_candidates_dev(vp8g_ctx* c, const vp8g_i420* frames, size_t n, size_t center, const int32_t* seeds,
                               const int32_t* first, int32_t* out, vp8g_err* err) {
  if (!c) return VP8G_E_ARG;
  int st = validate_arnr(frames, n, center, err);
  if (st != VP8G_OK) return st;
  if (!out || !first || out == first) return VP8G_E_ARG;
  if (((uintptr_t)out | (uintptr_t)seeds | (uintptr_t)first) & 3) return VP8G_E_ARG;
  if (!hip_ok(hipSetDevice(c->device))) return VP8G_E_DEVICE;
  SearchArgs a;
  memset(&a, 0, sizeof(a));
  bool aligned4 = true;
  for (size_t i = 0; i < n; i++) {
    const vp8g_i420& f = frames[i];
    a.frames[i] = FrameDesc{f.y, f.u, f.v, (uint32_t)f.y_stride, (uint32_t)f.uv_stride};
    if (((uintptr_t)f.y & 3) || (f.y_stride & 3)) aligned4 = false;
  }
  a.seeds = seeds;
  a.out = out;
  a.w = (int)frames[center].width;
  a.h = (int)frames[center].height;
  a.bcols = (a.w + 15) / 16;
  a.brows = (a.h + 15) / 16;
  a.n = (int)n;
  a.center = (int)center;
  a.aligned4 = aligned4 ? 1 : 0;
  a.total_blocks = (int)(n - 1) * a.bcols * a.brows;
  return launch_status(c, [&] { return launch_refine_candidates(a, first, c->stream); });
}

int vp8g_downsample2_i420_dev(vp8g_ctx* c, const vp8g_i420* src, const vp8g_i420* dst, size_t n) {
  if (!c || !src || !dst || n == 0 || n > (size_t)VP8G_MAX_FRAMES) return VP8G_E_ARG;
  const size_t w = src[0].width, h = src[0].height;
  if (w == 0 || h == 0 || w > 0x3FFF || h > 0x3FFF) return VP8G_E_INVALID_DIMENSIONS;
  for (size_t i = 0; i < n; i++) {
    if (src[i].width != w || src[i].height != h) return VP8G_E_REF_DIMS_MISMATCH;
    if (dst[i].width != (w + 1) / 2 || dst[i].height != (h + 1) / 2) return VP8G_E_REF_DIMS_MISMATCH;
    const vp8g_i420* fs[2] = {src + i, dst + i};
    for (const vp8g_i420* f : fs) {
      if (!f->y || !f->u || !f->v || f->y_stride < f->width || f->uv_stride < (f->width + 1) / 2) return VP8G_E_ARG;
      if (f->y_stride > 0xFFFFFFFFu || f->uv_stride > 0xFFFFFFFFu) return VP8G_E_ARG;
    }
  }
  if (!hip_ok(hipSetDevice(c->device))) return VP8G_E_DEVICE;
  for (size_t i0 = 0; i0 < n; i0 += kDownsampleFrames) {
    const int cnt = (int)std::min<size_t>(kDownsampleFrames, n - i0);
    Downsample2Args a;
    memset(&a, 0, sizeof(a));
    a.w = (int)w;
    a.h = (int)h;
    for (int k = 0; k < cnt; k++) {
      const vp8g_i420 &s = src[i0 + k], &d = dst[i0 + k];
      // dst's planes are the outputs (the descriptor type carries const pointers)
      a.f[k] = Downsample2Frame{s.y, s.u, s.v, const_cast<uint8_t*>(d.y), const_cast<uint8_t*>(d.u),
                                const_cast<uint8_t*>(d.v), (uint32_t)s.y_stride, (uint32_t)s.uv_stride,
                                (uint32_t)d.y_stride, (uint32_t)d.uv_stride};
    }
    if (launch_status(c, [&] { return launch_downsample2(a, cnt, c->stream); }) != VP8G_OK) return VP8G_E_DEVICE;
  }
  return VP8G_OK;
}

int vp8g_upsample_table_dev(vp8g_ctx* c, const int32_t* coarse, size_t n, int cw, int ch, int32_t* fine, int fw,
                            int fh) {
  if (!c || !coarse || !fine || n == 0 || n > (size_t)VP8G_MAX_FRAMES) return VP8G_E_ARG;
  // block grids of frames up to 0x3FFF pixels a side: at most 1024 blocks,
  // so n * fine_w * fine_h stays below 2^26
  constexpr int kMaxBlocks = (0x3FFF + 15) / 16;
  if (cw <= 0 || ch <= 0 || fw <= 0 || fh <= 0) return VP8G_E_ARG;
  if (cw > kMaxBlocks || ch > kMaxBlocks || fw > kMaxBlocks || fh > kMaxBlocks) return VP8G_E_ARG;
  if (((uintptr_t)coarse | (uintptr_t)fine) & 3) return VP8G_E_ARG;
  if (!hip_ok(hipSetDevice(c->device))) return VP8G_E_DEVICE;
  return launch_status(c, [&] { return launch_upsample_table(coarse, cw, ch, fine, fw, fh, (int)n, c->stream); });
}

int vp8g_build_arnr_altref_table_dev(vp8g_ctx* c, const vp8g_i420* frames, size_t n, size_t center,
                                     uint8_t strength, const int32_t* table, uint8_t* oy,
                                     uint8_t* ou, uint8_t* ov, vp8g_err* err) {
  if (!c) return VP8G_E_ARG;
  int st = validate_arnr(frames, n, center, err);
  if (st != VP8G_OK) return st;
  if (!oy || !ou || !ov || !table || ((uintptr_t)table & 3)) return VP8G_E_ARG;
  if (!hip_ok(hipSetDevice(c->device))) return VP8G_E_DEVICE;
  ArnrArgs a;
  memset(&a, 0, sizeof(a));
  fill_group(frames, n, center, strength, oy, ou, ov, nullptr, a.frames, &a.g, a.wt4);
  a.xcd_remap = c->xcd_remap;
  int k, m;
  pick_arnr_shape(c, a.g.w, a.g.h, (int)n, a.g.filter_on ? (int)n - 1 : 1, &k, &m);
  a.mbs_per_wg = m;
  a.g.strips_per_row = (a.g.bcols + m - 1) / m;
  a.spr_mul = strip_divider(a.g.strips_per_row);
  return launch_status(c, [&] { return launch_arnr_table(a, table, k, c->stream); });
}

int vp8g_arnr_batch_create(vp8g_ctx* c, const vp8g_arnr_job* jobs, size_t nj,
                           vp8g_arnr_batch** out, size_t* failed_job, vp8g_err* err) {
  if (!c || !out || (!jobs && nj)) return VP8G_E_ARG;
  *out = nullptr;
  if (nj == 0 || nj > 0xFFFF) return VP8G_E_ARG;
  size_t total_frames = 0;
  for (size_t j = 0; j < nj; j++) {
    const vp8g_arnr_job& jb = jobs[j];
    int st = validate_arnr(jb.frames, jb.n_frames, jb.center, err);
    if (st == VP8G_OK && (!jb.out_y || !jb.out_u || !jb.out_v)) st = VP8G_E_ARG;
    if (st != VP8G_OK) {
      if (failed_job) *failed_job = j;
      return st;
    }
    total_frames += jb.n_frames;
  }
  if (!hip_ok(hipSetDevice(c->device))) return VP8G_E_DEVICE;
  int k, m, want_per_cu;
  pick_batch_shape(c, &k, &m, &want_per_cu);
  std::vector<ArnrBatchJob> hj(nj);
  std::vector<FrameDesc> hf(total_frames);
  std::vector<uint16_t> hu;
  size_t fb = 0, units = 0;
  for (size_t j = 0; j < nj; j++) {
    const vp8g_arnr_job& jb = jobs[j];
    ArnrBatchJob& d = hj[j];
    memset(&d, 0, sizeof(d));
    fill_group(jb.frames, jb.n_frames, jb.center, jb.strength, jb.out_y, jb.out_u, jb.out_v,
               jb.mv_out, hf.data() + fb, &d.g, d.wt4);
    d.g.strips_per_row = (d.g.bcols + m - 1) / m;
    d.frame_base = (int)fb;
    d.unit_base = (int)units;
    const size_t ju = (size_t)d.g.strips_per_row * d.g.brows;
    hu.insert(hu.end(), ju, (uint16_t)j);
    units += ju;
    fb += jb.n_frames;
  }
  if (units > 0x7FFFFFFF) return VP8G_E_ARG;
  int per_cu = arnr_batch_workgroups_per_cu(k, m);
  if (per_cu < 1) return VP8G_E_DEVICE;
  if (want_per_cu > 0 && want_per_cu < per_cu) per_cu = want_per_cu;
  size_t wgs = (size_t)per_cu * (size_t)c->cus;
  if (wgs > units) wgs = units;

  vp8g_arnr_batch* b = new vp8g_arnr_batch();
  b->ctx = c;
  b->mbs = m;
  b->waves = k;
  b->workgroups = (int)wgs;
  b->units = units;
  const size_t o_jobs = 0, o_frames = align_up(sizeof(ArnrBatchJob) * nj, 256);
  const size_t o_units = o_frames + align_up(sizeof(FrameDesc) * total_frames, 256);
  const size_t o_ticket = o_units + align_up(sizeof(uint16_t) * units, 256);
  const size_t bytes = o_ticket + 256;
  hipStream_t cs = nullptr;
  bool ok = c->fault != VP8G_FAULT_ALLOC && hip_ok(hipMalloc(&b->dev, bytes)) && c->fault != VP8G_FAULT_COPY &&
            hip_ok(hipMemcpyAsync(b->dev + o_jobs, hj.data(), sizeof(ArnrBatchJob) * nj,
                                  hipMemcpyHostToDevice, c->stream)) &&
            hip_ok(hipMemcpyAsync(b->dev + o_frames, hf.data(), sizeof(FrameDesc) * total_frames,
                                  hipMemcpyHostToDevice, c->stream)) &&
            hip_ok(hipMemcpyAsync(b->dev + o_units, hu.data(), sizeof(uint16_t) * units,
                                  hipMemcpyHostToDevice, c->stream)) &&
            hip_ok(hipStreamSynchronize(c->stream));
  if (ok) {
    ArnrBatchArgs a;
    a.jobs = reinterpret_cast<const ArnrBatchJob*>(b->dev + o_jobs);
    a.frames = reinterpret_cast<const FrameDesc*>(b->dev + o_frames);
    a.unit_job = reinterpret_cast<const uint16_t*>(b->dev + o_units);
    a.total_units = (int)units;
    a.ticket = reinterpret_cast<uint32_t*>(b->dev + o_ticket);
    a.mbs_per_wg = m;
    // Captured on a private stream: the tickets are cleared by a kernel
    // (not a memset node), then the persistent kernel runs.
    ok = hip_ok(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking)) &&
         hip_ok(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
    if (ok) {
      const bool launched = launch_status(c, [&] { return launch_zero_u32(a.ticket, kBatchChunks, cs); }) == VP8G_OK &&
                            hip_ok(launch_arnr_batch(a, k, (int)wgs, cs));
      ok = hip_ok(hipStreamEndCapture(cs, &b->graph)) && launched;
    }
    ok = ok && hip_ok(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  }
  if (cs) (void)hipStreamDestroy(cs);
  if (!ok) {
    batch_free(b);
    return VP8G_E_DEVICE;
  }
  *out = b;
  return VP8G_OK;
}

int vp8g_arnr_batch_run(vp8g_arnr_batch* b) {
  if (!b) return VP8G_E_ARG;
  if (!hip_ok(hipSetDevice(b->ctx->device))) return VP8G_E_DEVICE;
  return launch_status(b->ctx, [&] { return hipGraphLaunch(b->exec, b->ctx->stream); });
}

int vp8g_arnr_batch_shape(const vp8g_arnr_batch* b, int* mbs, int* waves, int* workgroups,
                          size_t* units) {
  if (!b) return VP8G_E_ARG;
  if (mbs) *mbs = b->mbs;
  if (waves) *waves = b->waves;
  if (workgroups) *workgroups = b->workgroups;
  if (units) *units = b->units;
  return VP8G_OK;
}

void vp8g_arnr_batch_destroy(vp8g_arnr_batch* b) {
  if (!b) return;
  (void)hipSetDevice(b->ctx->device);
  (void)hipStreamSynchronize(b->ctx->stream);
  batch_free(b);
}

int vp8g_predict_inter_mb_batch_dev(vp8g_ctx* c, const vp8g_ref_planes* ref,
                                    const uint32_t* mb_xy, const vp8g_mv* mv, size_t n,
                                    int full_pixel, const int32_t filters[8][6], uint8_t* out) {
  if (!c || !filters) return VP8G_E_ARG;
  int st = check_ref_planes(ref);
  if (st != VP8G_OK) return st;
  if (n == 0) return VP8G_OK;
  if (!mb_xy || !mv || !out || n > 0x7FFFFFFF) return VP8G_E_ARG;
  if (((uintptr_t)mb_xy | (uintptr_t)mv | (uintptr_t)out) & 3) return VP8G_E_ARG;
  if (!hip_ok(hipSetDevice(c->device))) return VP8G_E_DEVICE;
  McArgs a = make_mc_args(ref, mb_xy, mv, n, full_pixel, filters, out);
  return launch_status(c, [&] { return launch_predict_inter_mb(a, c->stream); });
}

int vp8g_predict_inter_mb_batch(vp8g_ctx* c, const vp8g_ref_planes* ref, const uint32_t* mb_xy,
                                const vp8g_mv* mv, size_t n, int full_pixel,
                                const int32_t filters[8][6], uint8_t* out) {
  if (!c || !filters) return VP8G_E_ARG;
  int st = check_ref_planes(ref);
  if (st != VP8G_OK) return st;
  if (n == 0) return VP8G_OK;
  if (!mb_xy || !mv || !out || n > 0x7FFFFFFF) return VP8G_E_ARG;
  if (!mb_xy_in_range(mb_xy, n)) return VP8G_E_ARG;
  if (!hip_ok(hipSetDevice(c->device))) return VP8G_E_DEVICE;
  const size_t ys = plane_span(ref->y_stride, ref->mb_cols * 16, ref->mb_rows * 16);
  const size_t cs = plane_span(ref->uv_stride, ref->mb_cols * 8, ref->mb_rows * 8);
  const size_t o_y = 0, o_u = align_up(ys, 256), o_v = o_u + align_up(cs, 256);
  const size_t o_xy = o_v + align_up(cs, 256), o_mv = o_xy + align_up(8 * n, 256);
  const size_t o_out = o_mv + align_up(4 * n, 256), total = o_out + 384 * n;
  st = ensure_arena(c, total);
  if (st != VP8G_OK) return st;
  uint8_t* b = c->arena;
  st = copy_in(c, b + o_y, ref->y, ys);
  if (st == VP8G_OK) st = copy_in(c, b + o_u, ref->u, cs);
  if (st == VP8G_OK) st = copy_in(c, b + o_v, ref->v, cs);
  if (st == VP8G_OK) st = copy_in(c, b + o_xy, mb_xy, 8 * n);
  if (st == VP8G_OK) st = copy_in(c, b + o_mv, mv, 4 * n);
  if (st != VP8G_OK) return finish(c, st);
  vp8g_ref_planes d = *ref;
  d.y = b + o_y;
  d.u = b + o_u;
  d.v = b + o_v;
  McArgs a = make_mc_args(&d, reinterpret_cast<const uint32_t*>(b + o_xy),
                          reinterpret_cast<const vp8g_mv*>(b + o_mv), n, full_pixel, filters,
                          b + o_out);
  st = launch_status(c, [&] { return launch_predict_inter_mb(a, c->stream); });
  if (st == VP8G_OK) st = copy_out(c, out, b + o_out, 384 * n);
  return finish(c, st);
}

int vp8g_mb_luma_sad_at_mv_batch_dev(vp8g_ctx* c, const vp8g_luma_ref* ref,
                                     const uint32_t* mb_xy, const uint8_t* src,
                                     const vp8g_mv* mv, size_t n, uint32_t* sad) {
  if (!c) return VP8G_E_ARG;
  int st = check_luma_ref(ref);
  if (st != VP8G_OK) return st;
  if (n == 0) return VP8G_OK;
  if (!mb_xy || !src || !mv || !sad || n > 0x7FFFFFFF) return VP8G_E_ARG;
  if (((uintptr_t)mb_xy | (uintptr_t)src | (uintptr_t)mv | (uintptr_t)sad) & 3) return VP8G_E_ARG;
  if (!hip_ok(hipSetDevice(c->device))) return VP8G_E_DEVICE;
  SadArgs a = make_sad_args(ref, mb_xy, src, mv, n, sad);
  return launch_status(c, [&] { return launch_mb_luma_sad_at_mv(a, c->stream); });
}

// Host staging shared by the SAD-family entry points: plane, mb_xy, src,
// per-MB vectors; returns device pointers through the out-params.
static int stage_luma_batch(vp8g_ctx* c, const vp8g_luma_ref* ref, const uint32_t* mb_xy,
                            const uint8_t* src, const vp8g_mv* mv, size_t n, size_t out_bytes,
                            vp8g_luma_ref* dref, uint32_t** dxy, uint8_t** dsrc, vp8g_mv** dmv,
                            uint8_t** dout) {
  const size_t ps = plane_span(ref->stride, ref->width, ref->height);
  const size_t o_xy = align_up(ps, 256), o_src = o_xy + align_up(8 * n, 256);
  const size_t o_mv = o_src + align_up(256 * n, 256), o_out = o_mv + align_up(4 * n, 256);
  int st = ensure_arena(c, o_out + out_bytes);
  if (st != VP8G_OK) return st;
  uint8_t* b = c->arena;
  st = copy_in(c, b, ref->plane, ps);
  if (st == VP8G_OK) st = copy_in(c, b + o_xy, mb_xy, 8 * n);
  if (st == VP8G_OK) st = copy_in(c, b + o_src, src, 256 * n);
  if (st == VP8G_OK) st = copy_in(c, b + o_mv, mv, 4 * n);
  *dref = *ref;
  dref->plane = b;
  *dxy = reinterpret_cast<uint32_t*>(b + o_xy);
  *dsrc = b + o_src;
  *dmv = reinterpret_cast<vp8g_mv*>(b + o_mv);
  *dout = b + o_out;
  return st;
}

int vp8g_mb_luma_sad_at_mv_batch(vp8g_ctx* c, const vp8g_luma_ref* ref, const uint32_t* mb_xy,
                                 const uint8_t* src, const vp8g_mv* mv, size_t n, uint32_t* sad) {
  if (!c) return VP8G_E_ARG;
  int st = check_luma_ref(ref);
  if (st != VP8G_OK) return st;
  if (n == 0) return VP8G_OK;
  if (!mb_xy || !src || !mv || !sad || n > 0x7FFFFFFF) return VP8G_E_ARG;
  if (!mb_xy_in_range(mb_xy, n)) return VP8G_E_ARG;
  if (!hip_ok(hipSetDevice(c->device))) return VP8G_E_DEVICE;
  vp8g_luma_ref d;
  uint32_t* dxy;
  uint8_t *dsrc, *dout;
  vp8g_mv* dmv;
  st = stage_luma_batch(c, ref, mb_xy, src, mv, n, 4 * n, &d, &dxy, &dsrc, &dmv, &dout);
  if (st != VP8G_OK) return finish(c, st);
  SadArgs a = make_sad_args(&d, dxy, dsrc, dmv, n, reinterpret_cast<uint32_t*>(dout));
  st = launch_status(c, [&] { return launch_mb_luma_sad_at_mv(a, c->stream); });
  if (st == VP8G_OK) st = copy_out(c, sad, dout, 4 * n);
  return finish(c, st);
}

static int refine_batch(vp8g_ctx* c, const vp8g_luma_ref* ref, const uint32_t* mb_xy,
                        const uint8_t* src, const vp8g_mv* centers, size_t n,
                        vp8g_search_result* out, int step) {
  if (!c) return VP8G_E_ARG;
  int st = check_luma_ref(ref);
  if (st != VP8G_OK) return st;
  if (n == 0) return VP8G_OK;
  if (!mb_xy || !src || !centers || !out || n > 0x7FFFFFFF) return VP8G_E_ARG;
  if (!mb_xy_in_range(mb_xy, n)) return VP8G_E_ARG;
  for (size_t i = 0; i < n; i++)
    if (centers[i].row < -1023 || centers[i].row > 1023 || centers[i].col < -1023 ||
        centers[i].col > 1023)
      return VP8G_E_ARG;
  if (!hip_ok(hipSetDevice(c->device))) return VP8G_E_DEVICE;
  static_assert(sizeof(vp8g_search_result) == 8, "search result layout");
  vp8g_luma_ref d;
  uint32_t* dxy;
  uint8_t *dsrc, *dout;
  vp8g_mv* dmv;
  st = stage_luma_batch(c, ref, mb_xy, src, centers, n, 8 * n, &d, &dxy, &dsrc, &dmv, &dout);
  if (st != VP8G_OK) return finish(c, st);
  RefineArgs a;
  a.plane = d.plane;
  a.stride = (uint32_t)d.stride;
  a.w = (int)d.width;
  a.h = (int)d.height;
  a.mb_xy = dxy;
  a.src = dsrc;
  a.center = reinterpret_cast<const int16_t*>(dmv);
  a.out = reinterpret_cast<uint32_t*>(dout);
  a.n = (int)n;
  a.step = step;
  st = launch_status(c, [&] { return launch_refine_ladder(a, c->stream); });
  if (st == VP8G_OK) st = copy_out(c, out, dout, 8 * n);
  return finish(c, st);
}

int vp8g_half_plane_sads_batch(vp8g_ctx* c, const vp8g_luma_ref* ref, const int32_t* xy,
                               const uint8_t* src, size_t n, const int32_t taps[6], int diagonals,
                               uint32_t* sad) {
  if (!c || !taps) return VP8G_E_ARG;
  int st = check_luma_ref(ref);
  if (st != VP8G_OK) return st;
  if (n == 0) return VP8G_OK;
  if (!xy || !src || !sad || n > 0x7FFFFFF) return VP8G_E_ARG;
  for (size_t i = 0; i < 2 * n; i++)
    if (xy[i] < -(1 << 24) || xy[i] > (1 << 24)) return VP8G_E_ARG;
  if (!hip_ok(hipSetDevice(c->device))) return VP8G_E_DEVICE;
  const size_t ps = plane_span(ref->stride, ref->width, ref->height);
  const size_t o_xy = align_up(ps, 256), o_src = o_xy + align_up(8 * n, 256);
  const size_t o_out = o_src + align_up(256 * n, 256);
  st = ensure_arena(c, o_out + 32 * n);
  if (st != VP8G_OK) return st;
  uint8_t* b = c->arena;
  st = copy_in(c, b, ref->plane, ps);
  if (st == VP8G_OK) st = copy_in(c, b + o_xy, xy, 8 * n);
  if (st == VP8G_OK) st = copy_in(c, b + o_src, src, 256 * n);
  if (st != VP8G_OK) return finish(c, st);
  HalfPlaneArgs a;
  a.plane = b;
  a.stride = (uint32_t)ref->stride;
  a.w = (int)ref->width;
  a.h = (int)ref->height;
  a.xy = reinterpret_cast<const int32_t*>(b + o_xy);
  a.src = b + o_src;
  a.sad = reinterpret_cast<uint32_t*>(b + o_out);
  a.n = (int)n;
  a.diagonals = diagonals ? 1 : 0;
  memcpy(a.taps, taps, sizeof(a.taps));
  st = launch_status(c, [&] { return launch_half_plane_sads(a, c->stream); });
  if (st == VP8G_OK) st = copy_out(c, sad, b + o_out, 32 * n);
  return finish(c, st);
}

static int refine_batch_dev(vp8g_ctx* c, const vp8g_luma_ref* ref, const uint32_t* mb_xy,
                            const uint8_t* src, const vp8g_mv* centers, size_t n,
                            vp8g_search_result* out, int step) {
  if (!c) return VP8G_E_ARG;
  int st = check_luma_ref(ref);
  if (st != VP8G_OK) return st;
  if (n == 0) return VP8G_OK;
  if (!mb_xy || !src || !centers || !out || n > 0x7FFFFFFF) return VP8G_E_ARG;
  if (((uintptr_t)mb_xy | (uintptr_t)src | (uintptr_t)centers | (uintptr_t)out) & 3) return VP8G_E_ARG;
  if (!hip_ok(hipSetDevice(c->device))) return VP8G_E_DEVICE;
  RefineArgs a;
  a.plane = ref->plane;
  a.stride = (uint32_t)ref->stride;
  a.w = (int)ref->width;
  a.h = (int)ref->height;
  a.mb_xy = mb_xy;
  a.src = src;
  a.center = reinterpret_cast<const int16_t*>(centers);
  a.out = reinterpret_cast<uint32_t*>(out);
  a.n = (int)n;
  a.step = step;
  return launch_status(c, [&] { return launch_refine_ladder(a, c->stream); });
}

int vp8g_half_pixel_refine_luma_batch_dev(vp8g_ctx* c, const vp8g_luma_ref* ref, const uint32_t* mb_xy,
                                          const uint8_t* src, const vp8g_mv* centers, size_t n,
                                          vp8g_search_result* out) {
  return refine_batch_dev(c, ref, mb_xy, src, centers, n, out, 2);
}

int vp8g_quarter_pixel_refine_luma_batch_dev(vp8g_ctx* c, const vp8g_luma_ref* ref, const uint32_t* mb_xy,
                                             const uint8_t* src, const vp8g_mv* centers, size_t n,
                                             vp8g_search_result* out) {
  return refine_batch_dev(c, ref, mb_xy, src, centers, n, out, 1);
}

int vp8g_half_pixel_refine_luma_batch(vp8g_ctx* c, const vp8g_luma_ref* ref,
                                      const uint32_t* mb_xy, const uint8_t* src,
                                      const vp8g_mv* centers, size_t n, vp8g_search_result* out) {
  return refine_batch(c, ref, mb_xy, src, centers, n, out, 2);
}

int vp8g_quarter_pixel_refine_luma_batch(vp8g_ctx* c, const vp8g_luma_ref* ref,
                                         const uint32_t* mb_xy, const uint8_t* src,
                                         const vp8g_mv* centers, size_t n,
                                         vp8g_search_result* out) {
  return refine_batch(c, ref, mb_xy, src, centers, n, out, 1);
}

// ---- stream driver scene-cut metric
}  // extern "C"

namespace {

int check_mad_pairs(const vp8g_i420* prev, const vp8g_i420* cur, size_t n) {
  if (!prev || !cur) return VP8G_E_ARG;
  const uint32_t w = prev[0].width, h = prev[0].height;
  if (w == 0 || h == 0 || w > 0x3FFF || h > 0x3FFF) return VP8G_E_INVALID_DIMENSIONS;
  for (size_t i = 0; i < n; i++) {
    const vp8g_i420* fs[2] = {prev + i, cur + i};
    for (const vp8g_i420* f : fs) {
      if (f->width != w || f->height != h) return VP8G_E_REF_DIMS_MISMATCH;
      if (!f->y || f->y_stride < w || f->y_stride > 0xFFFFFFFFu) return VP8G_E_ARG;
    }
  }
  return VP8G_OK;
}

int enqueue_luma_sad(vp8g_ctx* c, const vp8g_i420* prev, const vp8g_i420* cur, size_t n, uint64_t* sad) {
  for (size_t i0 = 0; i0 < n; i0 += kLumaSadPairsPerLaunch) {
    const int cnt = (int)std::min<size_t>(kLumaSadPairsPerLaunch, n - i0);
    LumaSadArgs a;
    memset(&a, 0, sizeof(a));
    for (int k = 0; k < cnt; k++)
      a.pairs[k] = LumaSadPair{prev[i0 + k].y, cur[i0 + k].y, (uint32_t)prev[i0 + k].y_stride,
                               (uint32_t)cur[i0 + k].y_stride};
    a.sad = sad + i0;
    a.w = (int)prev[0].width;
    a.h = (int)prev[0].height;
    const int st = launch_status(c, [&] { return launch_luma_sad_pairs(a, cnt, c->stream); });
    if (st != VP8G_OK) return st;
  }
  return VP8G_OK;
}

}  // namespace

extern "C" {

int vp8g_luma_sad_batch_dev(vp8g_ctx* c, const vp8g_i420* prev, const vp8g_i420* cur, size_t n, uint64_t* sad) {
  if (!c) return VP8G_E_ARG;
  if (n == 0) return VP8G_OK;
  int st = check_mad_pairs(prev, cur, n);
  if (st != VP8G_OK) return st;
  if (!sad || ((uintptr_t)sad & 7)) return VP8G_E_ARG;
  if (!hip_ok(hipSetDevice(c->device))) return VP8G_E_DEVICE;
  return enqueue_luma_sad(c, prev, cur, n, sad);
}

int vp8g_luma_mad_batch(vp8g_ctx* c, const vp8g_i420* prev, const vp8g_i420* cur, size_t n, double* mad) {
  if (!c) return VP8G_E_ARG;
  if (n == 0) return VP8G_OK;
  int st = check_mad_pairs(prev, cur, n);
  if (st != VP8G_OK) return st;
  if (!mad) return VP8G_E_ARG;
  if (!hip_ok(hipSetDevice(c->device))) return VP8G_E_DEVICE;
  const size_t w = prev[0].width, h = prev[0].height, per = align_up(w * h, 256);
  const size_t o_sad = 2 * n * per;
  st = ensure_arena(c, o_sad + 8 * n);
  if (st != VP8G_OK) return st;
  std::vector<vp8g_i420> dp(n), dc(n);
  for (size_t i = 0; i < n && st == VP8G_OK; i++) {
    const vp8g_i420* fs[2] = {prev + i, cur + i};
    for (int k = 0; k < 2 && st == VP8G_OK; k++) {
      uint8_t* dst = c->arena + (2 * i + k) * per;
      if (c->fault == VP8G_FAULT_COPY) {
        st = VP8G_E_DEVICE;
        break;
      }
      pack_plane(c->pinned + (dst - c->arena), w, fs[k]->y, fs[k]->y_stride, w, h);
      if (!hip_ok(hipMemcpyAsync(dst, c->pinned + (dst - c->arena), w * h, hipMemcpyHostToDevice, c->stream)))
        st = VP8G_E_DEVICE;
      vp8g_i420 d = *fs[k];
      d.y = dst;
      d.y_stride = w;
      (k ? dc : dp)[i] = d;
    }
  }
  if (st != VP8G_OK) return finish(c, st);
  uint64_t* dsad = reinterpret_cast<uint64_t*>(c->arena + o_sad);
  st = enqueue_luma_sad(c, dp.data(), dc.data(), n, dsad);
  std::vector<uint64_t> sums(n);
  if (st == VP8G_OK) st = copy_out(c, sums.data(), dsad, 8 * n);
  st = finish(c, st);
  if (st == VP8G_OK)
    for (size_t i = 0; i < n; i++) mad[i] = (double)sums[i] / (double)(w * h);
  return st;
}

// ---- SPLITMV and reconstruction
}  // extern "C"

namespace {

SplitArgs make_split(const vp8g_ref_planes* r, const uint32_t* mb_xy, const vp8g_mv* mvs, size_t n, int full_pixel,
                     const int32_t filters[8][6], uint8_t* out) {
  SplitArgs a;
  memset(&a, 0, sizeof(a));
  a.ref = McPlanes{r->y, r->u, r->v, (uint32_t)r->y_stride, (uint32_t)r->uv_stride, (int)(r->mb_cols * 16),
                   (int)(r->mb_rows * 16), (int)(r->mb_cols * 8), (int)(r->mb_rows * 8)};
  a.mb_xy = mb_xy;
  a.mvs = reinterpret_cast<const int16_t*>(mvs);
  a.out = out;
  a.n = (int)n;
  a.full_pixel = full_pixel ? 1 : 0;
  memcpy(a.filters, filters, sizeof(a.filters));
  return a;
}

// Enqueue prediction (split or whole-MB) and, with coefficients, the
// residual, on device arrays.  mvs_per_mb: 16 (SPLITMV) or 1.
int enqueue_recon(vp8g_ctx* c, const vp8g_ref_planes* ref, const uint32_t* mb_xy, const vp8g_mv* mvs, size_t n,
                  int full_pixel, const int32_t filters[8][6], const uint8_t* skip, const int16_t* coeffs,
                  uint8_t* out, int mvs_per_mb) {
  int st;
  if (mvs_per_mb == 16) {
    const SplitArgs a = make_split(ref, mb_xy, mvs, n, full_pixel, filters, out);
    st = launch_status(c, [&] { return launch_split_predict(a, c->stream); });
  } else {
    const McArgs a = make_mc_args(ref, mb_xy, mvs, n, full_pixel, filters, out);
    st = launch_status(c, [&] { return launch_predict_inter_mb(a, c->stream); });
  }
  if (st != VP8G_OK || !coeffs) return st;
  ResidualArgs r;
  r.coeffs = coeffs;
  r.skip = skip;
  r.out = out;
  r.n = (int)n;
  r.y2 = mvs_per_mb == 16 ? 0 : 1;
  return launch_status(c, [&] { return launch_mb_residual(r, c->stream); });
}

int recon_dev(vp8g_ctx* c, const vp8g_ref_planes* ref, const uint32_t* mb_xy, const vp8g_mv* mvs, size_t n,
              int full_pixel, const int32_t filters[8][6], const uint8_t* skip, const int16_t* coeffs, uint8_t* out,
              int mvs_per_mb, bool with_residual) {
  if (!c || !filters || !taps_valid(filters)) return VP8G_E_ARG;
  int st = check_ref_planes(ref);
  if (st != VP8G_OK) return st;
  if (n == 0) return VP8G_OK;
  if (!mb_xy || !mvs || !out || n > 0x7FFFFFF || (with_residual && (!skip || !coeffs))) return VP8G_E_ARG;
  if (((uintptr_t)mb_xy | (uintptr_t)mvs | (uintptr_t)out | (uintptr_t)coeffs) & 3) return VP8G_E_ARG;
  if (!hip_ok(hipSetDevice(c->device))) return VP8G_E_DEVICE;
  return enqueue_recon(c, ref, mb_xy, mvs, n, full_pixel, filters, with_residual ? skip : nullptr,
                       with_residual ? coeffs : nullptr, out, mvs_per_mb);
}

int recon_host(vp8g_ctx* c, const vp8g_ref_planes* ref, const uint32_t* mb_xy, const vp8g_mv* mvs, size_t n,
               int full_pixel, const int32_t filters[8][6], const uint8_t* skip, const int16_t* coeffs, uint8_t* out,
               int mvs_per_mb, bool with_residual) {
  if (!c || !filters || !taps_valid(filters)) return VP8G_E_ARG;
  int st = check_ref_planes(ref);
  if (st != VP8G_OK) return st;
  if (n == 0) return VP8G_OK;
  if (!mb_xy || !mvs || !out || n > 0x7FFFFFF || (with_residual && (!skip || !coeffs))) return VP8G_E_ARG;
  if (!mb_xy_in_range(mb_xy, n)) return VP8G_E_ARG;
  if (!hip_ok(hipSetDevice(c->device))) return VP8G_E_DEVICE;
  const size_t ys = plane_span(ref->y_stride, ref->mb_cols * 16, ref->mb_rows * 16);
  const size_t cs = plane_span(ref->uv_stride, ref->mb_cols * 8, ref->mb_rows * 8);
  const size_t cf_bytes = with_residual ? (size_t)(mvs_per_mb == 16 ? 384 : 400) * 2 * n : 0;
  const size_t o_y = 0, o_u = align_up(ys, 256), o_v = o_u + align_up(cs, 256);
  const size_t o_xy = o_v + align_up(cs, 256), o_mv = o_xy + align_up(8 * n, 256);
  const size_t o_skip = o_mv + align_up(4 * mvs_per_mb * n, 256);
  const size_t o_cf = o_skip + align_up(with_residual ? n : 0, 256);
  const size_t o_out = o_cf + align_up(cf_bytes, 256), total = o_out + 384 * n;
  st = ensure_arena(c, total);
  if (st != VP8G_OK) return st;
  uint8_t* b = c->arena;
  st = copy_in(c, b + o_y, ref->y, ys);
  if (st == VP8G_OK) st = copy_in(c, b + o_u, ref->u, cs);
  if (st == VP8G_OK) st = copy_in(c, b + o_v, ref->v, cs);
  if (st == VP8G_OK) st = copy_in(c, b + o_xy, mb_xy, 8 * n);
  if (st == VP8G_OK) st = copy_in(c, b + o_mv, mvs, 4 * mvs_per_mb * n);
  if (st == VP8G_OK && with_residual) st = copy_in(c, b + o_skip, skip, n);
  if (st == VP8G_OK && with_residual) st = copy_in(c, b + o_cf, coeffs, cf_bytes);
  if (st != VP8G_OK) return finish(c, st);
  vp8g_ref_planes d = *ref;
  d.y = b + o_y;
  d.u = b + o_u;
  d.v = b + o_v;
  st = enqueue_recon(c, &d, reinterpret_cast<const uint32_t*>(b + o_xy), reinterpret_cast<const vp8g_mv*>(b + o_mv),
                     n, full_pixel, filters, with_residual ? b + o_skip : nullptr,
                     with_residual ? reinterpret_cast<const int16_t*>(b + o_cf) : nullptr, b + o_out, mvs_per_mb);
  if (st == VP8G_OK) st = copy_out(c, out, b + o_out, 384 * n);
  return finish(c, st);
}

}  // namespace

extern "C" {

int vp8g_predict_split_mv_batch(vp8g_ctx* c, const vp8g_ref_planes* ref, const uint32_t* mb_xy, const vp8g_mv* mvs,
                                size_t n, int full_pixel, const int32_t filters[8][6], uint8_t* out) {
  return recon_host(c, ref, mb_xy, mvs, n, full_pixel, filters, nullptr, nullptr, out, 16, false);
}

int vp8g_predict_split_mv_batch_dev(vp8g_ctx* c, const vp8g_ref_planes* ref, const uint32_t* mb_xy,
                                    const vp8g_mv* mvs, size_t n, int full_pixel, const int32_t filters[8][6],
                                    uint8_t* out) {
  return recon_dev(c, ref, mb_xy, mvs, n, full_pixel, filters, nullptr, nullptr, out, 16, false);
}

int vp8g_reconstruct_inter_mb_batch(vp8g_ctx* c, const vp8g_ref_planes* ref, const uint32_t* mb_xy,
                                    const vp8g_mv* mv, size_t n, int full_pixel, const int32_t filters[8][6],
                                    const uint8_t* skip, const int16_t* coeffs, uint8_t* out) {
  return recon_host(c, ref, mb_xy, mv, n, full_pixel, filters, skip, coeffs, out, 1, true);
}

int vp8g_reconstruct_inter_mb_batch_dev(vp8g_ctx* c, const vp8g_ref_planes* ref, const uint32_t* mb_xy,
                                        const vp8g_mv* mv, size_t n, int full_pixel, const int32_t filters[8][6],
                                        const uint8_t* skip, const int16_t* coeffs, uint8_t* out) {
  return recon_dev(c, ref, mb_xy, mv, n, full_pixel, filters, skip, coeffs, out, 1, true);
}

int vp8g_reconstruct_split_mv_mb_batch(vp8g_ctx* c, const vp8g_ref_planes* ref, const uint32_t* mb_xy,
                                       const vp8g_mv* mvs, size_t n, int full_pixel, const int32_t filters[8][6],
                                       const uint8_t* skip, const int16_t* coeffs, uint8_t* out) {
  return recon_host(c, ref, mb_xy, mvs, n, full_pixel, filters, skip, coeffs, out, 16, true);
}

int vp8g_reconstruct_split_mv_mb_batch_dev(vp8g_ctx* c, const vp8g_ref_planes* ref, const uint32_t* mb_xy,
                                           const vp8g_mv* mvs, size_t n, int full_pixel,
                                           const int32_t filters[8][6], const uint8_t* skip, const int16_t* coeffs,
                                           uint8_t* out) {
  return recon_dev(c, ref, mb_xy, mvs, n, full_pixel, filters, skip, coeffs, out, 16, true);
}

// ---- encoder inter search
}  // extern "C"

namespace {

bool jobs_valid(const vp8g_mb_search* jobs, size_t n) {
  for (size_t i = 0; i < n; i++) {
    const vp8g_mb_search& j = jobs[i];
    if (j.mb_col > 0xFFFFFu || j.mb_row > 0xFFFFFu || j.n_seeds > 8) return false;
    if (j.range_x[0] > j.range_x[1] || j.range_y[0] > j.range_y[1]) return false;
    if (j.range_x[0] < -8191 || j.range_x[1] > 8191 || j.range_y[0] < -8191 || j.range_y[1] > 8191) return false;
    if (!isfinite(j.lambda_sad) || j.lambda_sad < 0.0 || j.good_enough >= (1u << 30)) return false;
  }
  return true;
}

WpsArgs make_wps(const vp8g_luma_ref* r, const vp8g_mv_cost_table* costs, const vp8g_mb_search* jobs,
                 const uint8_t* src, size_t n, const vp8g_search_shape* shape, vp8g_found* out) {
  WpsArgs a;
  memset(&a, 0, sizeof(a));
  a.plane = r->plane;
  a.stride = (uint32_t)r->stride;
  a.w = (int)r->width;
  a.h = (int)r->height;
  a.bits = &costs->bits[0][0];
  a.jobs = jobs;
  a.src = src;
  a.out = out;
  a.n = (int)n;
  a.descent_rounds = shape->descent_rounds;
  a.wide_pattern = shape->wide_pattern ? 1 : 0;
  return a;
}

SprArgs make_spr(const vp8g_luma_ref* r, const vp8g_mv_cost_table* costs, const vp8g_mb_search* jobs,
                 const uint8_t* src, const vp8g_found* start, size_t n, const int32_t filters[8][6],
                 const vp8g_search_shape* shape, vp8g_found* out) {
  SprArgs a;
  memset(&a, 0, sizeof(a));
  a.plane = r->plane;
  a.stride = (uint32_t)r->stride;
  a.w = (int)r->width;
  a.h = (int)r->height;
  a.bits = &costs->bits[0][0];
  a.jobs = jobs;
  a.src = src;
  a.start = start;
  a.out = out;
  a.n = (int)n;
  a.half_diagonals = shape->half_diagonals ? 1 : 0;
  a.quarter = shape->quarter ? 1 : 0;
  memcpy(a.taps, filters[4], sizeof(a.taps));
  return a;
}

DiamondArgs make_diamond(const vp8g_luma_ref* r, const uint32_t* mb_xy, const uint8_t* src, const vp8g_mv* centers,
                         size_t n, uint32_t max_iters, vp8g_search_result* out) {
  DiamondArgs a;
  memset(&a, 0, sizeof(a));
  a.plane = r->plane;
  a.stride = (uint32_t)r->stride;
  a.w = (int)r->width;
  a.h = (int)r->height;
  a.mb_xy = mb_xy;
  a.src = src;
  a.center = reinterpret_cast<const int16_t*>(centers);
  a.out = reinterpret_cast<uint32_t*>(out);
  a.n = (int)n;
  a.max_iters = max_iters;
  return a;
}

// Device offsets of a staged inter-search batch in the arena.
struct InterStage {
  size_t plane, costs, jobs, src, start, out, total;
};

InterStage inter_stage_layout(const vp8g_luma_ref* ref, size_t n, bool with_start, size_t out_bytes) {
  InterStage L;
  L.plane = 0;
  L.costs = align_up(plane_span(ref->stride, ref->width, ref->height), 256);
  L.jobs = L.costs + align_up(sizeof(vp8g_mv_cost_table), 256);
  L.src = L.jobs + align_up(sizeof(vp8g_mb_search) * n, 256);
  L.start = L.src + align_up(256 * n, 256);
  L.out = L.start + (with_start ? align_up(sizeof(vp8g_found) * n, 256) : 0);
  L.total = L.out + out_bytes;
  return L;
}

}  // namespace

extern "C" {

int vp8g_whole_pixel_search_batch_dev(vp8g_ctx* c, const vp8g_luma_ref* ref, const vp8g_mv_cost_table* costs,
                                      const vp8g_mb_search* jobs, const uint8_t* src, size_t n,
                                      const vp8g_search_shape* shape, vp8g_found* out) {
  if (!c || !shape) return VP8G_E_ARG;
  int st = check_luma_ref(ref);
  if (st != VP8G_OK) return st;
  if (n == 0) return VP8G_OK;
  if (!costs || !jobs || !src || !out || n > 0x7FFFFFF) return VP8G_E_ARG;
  if (((uintptr_t)costs & 3) || ((uintptr_t)src & 3) || (((uintptr_t)jobs | (uintptr_t)out) & 7)) return VP8G_E_ARG;
  if (!hip_ok(hipSetDevice(c->device))) return VP8G_E_DEVICE;
  const WpsArgs a = make_wps(ref, costs, jobs, src, n, shape, out);
  return launch_status(c, [&] { return launch_whole_pixel_search(a, c->stream); });
}

int vp8g_whole_pixel_search_batch(vp8g_ctx* c, const vp8g_luma_ref* ref, const vp8g_mv_cost_table* costs,
                                  const vp8g_mb_search* jobs, const uint8_t* src, size_t n,
                                  const vp8g_search_shape* shape, vp8g_found* out) {
  if (!c || !shape) return VP8G_E_ARG;
  int st = check_luma_ref(ref);
  if (st != VP8G_OK) return st;
  if (n == 0) return VP8G_OK;
  if (!costs || !jobs || !src || !out || n > 0x7FFFFFF || !jobs_valid(jobs, n)) return VP8G_E_ARG;
  if (!hip_ok(hipSetDevice(c->device))) return VP8G_E_DEVICE;
  const InterStage L = inter_stage_layout(ref, n, false, sizeof(vp8g_found) * n);
  st = ensure_arena(c, L.total);
  if (st != VP8G_OK) return st;
  uint8_t* b = c->arena;
  st = copy_in(c, b + L.plane, ref->plane, plane_span(ref->stride, ref->width, ref->height));
  if (st == VP8G_OK) st = copy_in(c, b + L.costs, costs, sizeof(vp8g_mv_cost_table));
  if (st == VP8G_OK) st = copy_in(c, b + L.jobs, jobs, sizeof(vp8g_mb_search) * n);
  if (st == VP8G_OK) st = copy_in(c, b + L.src, src, 256 * n);
  if (st != VP8G_OK) return finish(c, st);
  vp8g_luma_ref d = *ref;
  d.plane = b + L.plane;
  const WpsArgs a = make_wps(&d, reinterpret_cast<const vp8g_mv_cost_table*>(b + L.costs),
                             reinterpret_cast<const vp8g_mb_search*>(b + L.jobs), b + L.src, n, shape,
                             reinterpret_cast<vp8g_found*>(b + L.out));
  st = launch_status(c, [&] { return launch_whole_pixel_search(a, c->stream); });
  if (st == VP8G_OK) st = copy_out(c, out, b + L.out, sizeof(vp8g_found) * n);
  return finish(c, st);
}

int vp8g_sub_pixel_refine_batch_dev(vp8g_ctx* c, const vp8g_luma_ref* ref, const vp8g_mv_cost_table* costs,
                                    const vp8g_mb_search* jobs, const uint8_t* src, const vp8g_found* start,
                                    size_t n, const int32_t filters[8][6], const vp8g_search_shape* shape,
                                    vp8g_found* out) {
  if (!c || !shape || !filters || !taps_valid(filters)) return VP8G_E_ARG;
  int st = check_luma_ref(ref);
  if (st != VP8G_OK) return st;
  if (n == 0) return VP8G_OK;
  if (!costs || !jobs || !src || !start || !out || n > 0x7FFFFFF) return VP8G_E_ARG;
  if (((uintptr_t)costs & 3) || ((uintptr_t)src & 3) || (((uintptr_t)jobs | (uintptr_t)out | (uintptr_t)start) & 7))
    return VP8G_E_ARG;
  if (!hip_ok(hipSetDevice(c->device))) return VP8G_E_DEVICE;
  const SprArgs a = make_spr(ref, costs, jobs, src, start, n, filters, shape, out);
  return launch_status(c, [&] { return launch_sub_pixel_refine(a, c->stream); });
}

int vp8g_sub_pixel_refine_batch(vp8g_ctx* c, const vp8g_luma_ref* ref, const vp8g_mv_cost_table* costs,
                                const vp8g_mb_search* jobs, const uint8_t* src, const vp8g_found* start, size_t n,
                                const int32_t filters[8][6], const vp8g_search_shape* shape, vp8g_found* out) {
  if (!c || !shape || !filters || !taps_valid(filters)) return VP8G_E_ARG;
  int st = check_luma_ref(ref);
  if (st != VP8G_OK) return st;
  if (n == 0) return VP8G_OK;
  if (!costs || !jobs || !src || !start || !out || n > 0x7FFFFFF || !jobs_valid(jobs, n)) return VP8G_E_ARG;
  if (!hip_ok(hipSetDevice(c->device))) return VP8G_E_DEVICE;
  const InterStage L = inter_stage_layout(ref, n, true, sizeof(vp8g_found) * n);
  st = ensure_arena(c, L.total);
  if (st != VP8G_OK) return st;
  uint8_t* b = c->arena;
  st = copy_in(c, b + L.plane, ref->plane, plane_span(ref->stride, ref->width, ref->height));
  if (st == VP8G_OK) st = copy_in(c, b + L.costs, costs, sizeof(vp8g_mv_cost_table));
  if (st == VP8G_OK) st = copy_in(c, b + L.jobs, jobs, sizeof(vp8g_mb_search) * n);
  if (st == VP8G_OK) st = copy_in(c, b + L.src, src, 256 * n);
  if (st == VP8G_OK) st = copy_in(c, b + L.start, start, sizeof(vp8g_found) * n);
  if (st != VP8G_OK) return finish(c, st);
  vp8g_luma_ref d = *ref;
  d.plane = b + L.plane;
  const SprArgs a = make_spr(&d, reinterpret_cast<const vp8g_mv_cost_table*>(b + L.costs),
                             reinterpret_cast<const vp8g_mb_search*>(b + L.jobs), b + L.src,
                             reinterpret_cast<const vp8g_found*>(b + L.start), n, filters, shape,
                             reinterpret_cast<vp8g_found*>(b + L.out));
  st = launch_status(c, [&] { return launch_sub_pixel_refine(a, c->stream); });
  if (st == VP8G_OK) st = copy_out(c, out, b + L.out, sizeof(vp8g_found) * n);
  return finish(c, st);
}

int vp8g_small_diamond_search_luma_batch_dev(vp8g_ctx* c, const vp8g_luma_ref* ref, const uint32_t* mb_xy,
                                             const uint8_t* src, const vp8g_mv* centers, size_t n,
                                             uint32_t max_iters, vp8g_search_result* out) {
  if (!c) return VP8G_E_ARG;
  int st = check_luma_ref(ref);
  if (st != VP8G_OK) return st;
  if (n == 0) return VP8G_OK;
  if (!mb_xy || !src || !centers || !out || n > 0x7FFFFFF) return VP8G_E_ARG;
  if (((uintptr_t)mb_xy | (uintptr_t)src | (uintptr_t)centers | (uintptr_t)out) & 3) return VP8G_E_ARG;
  if (!hip_ok(hipSetDevice(c->device))) return VP8G_E_DEVICE;
  const DiamondArgs a = make_diamond(ref, mb_xy, src, centers, n, max_iters, out);
  return launch_status(c, [&] { return launch_small_diamond(a, c->stream); });
}

int vp8g_small_diamond_search_luma_batch(vp8g_ctx* c, const vp8g_luma_ref* ref, const uint32_t* mb_xy,
                                         const uint8_t* src, const vp8g_mv* centers, size_t n, uint32_t max_iters,
                                         vp8g_search_result* out) {
  if (!c) return VP8G_E_ARG;
  int st = check_luma_ref(ref);
  if (st != VP8G_OK) return st;
  if (n == 0) return VP8G_OK;
  if (!mb_xy || !src || !centers || !out || n > 0x7FFFFFF) return VP8G_E_ARG;
  if (!mb_xy_in_range(mb_xy, n)) return VP8G_E_ARG;
  if (!hip_ok(hipSetDevice(c->device))) return VP8G_E_DEVICE;
  vp8g_luma_ref d;
  uint32_t* dxy;
  uint8_t *dsrc, *dout;
  vp8g_mv* dmv;
  st = stage_luma_batch(c, ref, mb_xy, src, centers, n, 8 * n, &d, &dxy, &dsrc, &dmv, &dout);
  if (st != VP8G_OK) return finish(c, st);
  const DiamondArgs a = make_diamond(&d, dxy, dsrc, dmv, n, max_iters, reinterpret_cast<vp8g_search_result*>(dout));
  st = launch_status(c, [&] { return launch_small_diamond(a, c->stream); });
  if (st == VP8G_OK) st = copy_out(c, out, dout, 8 * n);
  return finish(c, st);
}

}  // extern "C"
