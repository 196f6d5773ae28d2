// Shared host-side helpers of the C-ABI layer (not part of the ABI).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <string>

#include "tfcehip.h"

namespace tfcehip {

struct Workspace {
  void* ptr = nullptr;
  size_t bytes = 0;
  int32_t* queue = nullptr;  // 256 bytes: work counters
  bool captured = false;     // a graph captured on this stream holds ptr
};

// Device memory a caller lends to a TFCE launch instead of the per-stream
// scratch (the no-allocation entry point).
struct ExtScratch {
  void* base;
  size_t bytes;
};

int set_error(int code, const std::string& msg);
int current_device();
int cu_count(int device);
int workspace_get(hipStream_t stream, size_t bytes, Workspace** out);
// True while stream is capturing a graph.
bool stream_capturing(hipStream_t stream);
// Give up a scratch buffer of `stream`: freed at once (after the stream has
// drained) unless a captured graph may still use it, in which case it stays
// allocated until the process ends.
void scratch_release(void* ptr, bool captured, hipStream_t stream);
int tfce_grid(int64_t n_items);
size_t device_memory(int device);
// tfce_batch_max, on ext when given (nothing allocated) or on the stream's scratch.
int batch_max_in(tfce_adj_t* adj, const float* d_maps, int P, int sign_mask, const float* d_vdensity,
                 float* d_max, const ExtScratch* ext, hipStream_t stream);

}  // namespace tfcehip
