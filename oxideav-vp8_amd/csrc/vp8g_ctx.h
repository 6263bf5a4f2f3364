// The context object behind the C ABI's vp8g_ctx handle (host only).
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/vp8g.h"
#include "host_pool.h"
#include "tuning.h"

struct vp8g_ctx {
  int device = 0;
  hipStream_t own = nullptr;
  hipStream_t stream = nullptr;
  int arnr_waves = 0;
  int xcd_remap = 1;
  int arnr_mbs = 0;
  int batch_wgs_per_cu = 0;
  int fault = VP8G_FAULT_NONE;  // injected host-side failure (VP8G_OPT_FAULT)
  int cus = 0;
  std::vector<vp8g::TuneEntry> tuning;  // launch shapes by (resolution class, group length)
  // Device arena for the host-pointer entry points (grown, never shrunk).
  uint8_t* arena = nullptr;
  size_t arena_bytes = 0;
  // Pinned host mirror of the arena (hipHostMalloc, same size and offsets):
  // host inputs are packed here and cross PCIe from pinned memory; outputs
  // land here and are copied to the caller once the stream has finished.
  uint8_t* pinned = nullptr;
  struct PendingOut {
    void* host;
    size_t off, bytes;
  };
  std::vector<PendingOut> pending;
  // Device workspace of the `_dev` entry points that need scratch (the
  // fallback queue); grown, never shrunk.
  uint8_t* work = nullptr;
  size_t work_bytes = 0;
  // Device address of this device's static no-alloc scratch (capi.hip).
  uint8_t* scratch = nullptr;
  // Host-pointer ARNR entry point: host threads that pack the caller's
  // planes into the write-combined pinned staging, and the link stream
  // whose copy kernels pull the packed rows over PCIe (both created on
  // first use).  Band b's rows are on the device when band_in[b] has fired
  // on the link stream, and its altref rows are back in pinned memory when
  // band_done[b] has fired on the context's stream.
  static constexpr int kMaxCopyStreams = 4;
  static constexpr int kMaxBands = 8;
  int host_threads = 0;  // VP8G_OPT_HOST_THREADS (0 = default)
  int copy_streams = 0;  // VP8G_OPT_COPY_STREAMS (range-checked, no effect)
  int host_bands = 0;    // VP8G_OPT_HOST_BANDS (0 = default)
  vp8g::HostPool* pool = nullptr;
  hipStream_t link = nullptr;
  hipEvent_t band_in[kMaxBands] = {};
  hipEvent_t band_done[kMaxBands] = {};
  uint8_t* staging = nullptr;
  size_t staging_bytes = 0;
  // Loop filter (lf_host.hip): row rings and per-instance SSDs, the device
  // copy of host-pointer inputs with its pinned mirror, the candidate SSDs
  // of the level search, and the flag a launch sets when a row wait gives up.
  uint8_t* lf_work = nullptr;
  size_t lf_work_bytes = 0;
  uint8_t* lf_io = nullptr;
  uint8_t* lf_pinned = nullptr;
  size_t lf_io_bytes = 0;
  unsigned long long* lf_ssd = nullptr;  // kLfSsdSlots entries
  int* lf_err = nullptr;
};

// Group validation in the reference's order (src/arnr.rs:494-515) plus the
// library's own limits; shared by every group entry point.
int vp8g_validate_group(const vp8g_i420* frames, size_t n, size_t center, vp8g_err* err);
