// Internal declarations of the adjacency-builder kernels (not part of the ABI).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tfcehip {

// int64 words of scratch launch_exclusive_scan needs for n inputs
int64_t scan_scratch_words(int64_t n);
// out[i] = in[0] + ... + in[i-1] (int32, non-negative inputs), out[n] = total
hipError_t launch_exclusive_scan(const int32_t* in, int64_t n, int32_t* out, int64_t* sums, hipStream_t s);
hipError_t launch_mask_flags(const uint8_t* mask, int64_t n, int32_t* flags, hipStream_t s);
// col == null: cnt[label] = row length; otherwise the rows into col at row_ptr
hipError_t launch_voxel_rows(const uint8_t* mask, const int32_t* label, int X, int Y, int Z, int dirtype,
                             int32_t* cnt, const int32_t* row_ptr, int32_t* col, hipStream_t s);
hipError_t launch_face_count(const int32_t* faces, int64_t F, int32_t nv, int32_t* cnt, int32_t* bad, hipStream_t s);
hipError_t launch_face_scatter(const int32_t* faces, int64_t F, int32_t nv, const int32_t* off, int32_t* cur,
                               int32_t* tmp, hipStream_t s);
hipError_t launch_vertex_unique(const int32_t* off, int32_t nv, int32_t* tmp, int32_t* ulen, hipStream_t s);
hipError_t launch_vertex_copy(const int32_t* off, const int32_t* tmp, const int32_t* row_ptr, int32_t nv, int32_t* col,
                              hipStream_t s);

}  // namespace tfcehip
