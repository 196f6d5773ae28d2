// Adjacency builders on the GPU: the CSR of create_adjac_voxel
// (pyfunc.py:48-76) and of create_adjac_vertex (pyfunc.py:37-46), built once
// per analysis from a device-resident mask or face list.  Integer work only;
// every output is fixed by the input (no atomic order leaks into it).
#include "adj_internal.h"

namespace tfcehip {

namespace {

constexpr int kScanThreads = 1024;
constexpr int kScanPer = 4;
constexpr int64_t kScanChunk = (int64_t)kScanThreads * kScanPer;

__device__ __forceinline__ int64_t block_incl_scan(int64_t v, int64_t* s_wave) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int64_t t = __shfl_up(v, o, 64);
    if (lane >= o) v += t;
  }
  if (lane == 63) s_wave[wave] = v;
  __syncthreads();
  if (threadIdx.x < 64) {
    int64_t w = threadIdx.x < (kScanThreads / 64) ? s_wave[threadIdx.x] : 0;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int64_t t = __shfl_up(w, o, 64);
      if ((int)threadIdx.x >= o) w += t;
    }
    if (threadIdx.x < (kScanThreads / 64)) s_wave[threadIdx.x] = w;
  }
  __syncthreads();
  const int64_t base = wave > 0 ? s_wave[wave - 1] : 0;
  __syncthreads();
  return v + base;
}

// sums[b] = sum of in[b * chunk, (b + 1) * chunk)
__global__ __launch_bounds__(kScanThreads) void chunk_sum_kernel(const int32_t* in, int64_t n, int64_t* sums) {
  __shared__ int64_t s_wave[kScanThreads / 64];
  const int64_t i0 = (int64_t)blockIdx.x * kScanChunk + (int64_t)threadIdx.x * kScanPer;
  int64_t v = 0;
#pragma unroll
  for (int j = 0; j < kScanPer; ++j)
    if (i0 + j < n) v += in[i0 + j];
  v = block_incl_scan(v, s_wave);
  if (threadIdx.x == kScanThreads - 1) sums[blockIdx.x] = v;
}

// exclusive scan of sums[0, nb) in place, total in sums[nb]: one workgroup
__global__ __launch_bounds__(kScanThreads) void sums_scan_kernel(int64_t* sums, int64_t nb) {
  __shared__ int64_t s_wave[kScanThreads / 64];
  __shared__ int64_t s_carry;
  if (threadIdx.x == 0) s_carry = 0;
  __syncthreads();
  for (int64_t b0 = 0; b0 < nb; b0 += kScanThreads) {
    const int64_t b = b0 + threadIdx.x;
    const int64_t v = b < nb ? sums[b] : 0;
    const int64_t incl = block_incl_scan(v, s_wave);
    const int64_t carry = s_carry;
    if (b < nb) sums[b] = carry + incl - v;
    __syncthreads();
    if (threadIdx.x == kScanThreads - 1) s_carry = carry + incl;
    __syncthreads();
  }
  if (threadIdx.x == 0) sums[nb] = s_carry;
}

// out[i] = base of the chunk + exclusive prefix within it; out[n] = total
__global__ __launch_bounds__(kScanThreads) void chunk_scan_kernel(const int32_t* in, int64_t n, const int64_t* sums,
                                                                   int64_t nb, int32_t* out) {
  __shared__ int64_t s_wave[kScanThreads / 64];
  const int64_t i0 = (int64_t)blockIdx.x * kScanChunk + (int64_t)threadIdx.x * kScanPer;
  int32_t x[kScanPer];
  int64_t v = 0;
#pragma unroll
  for (int j = 0; j < kScanPer; ++j) {
    x[j] = (i0 + j < n) ? in[i0 + j] : 0;
    v += x[j];
  }
  int64_t run = block_incl_scan(v, s_wave) - v + sums[blockIdx.x];
#pragma unroll
  for (int j = 0; j < kScanPer; ++j) {
    if (i0 + j < n) out[i0 + j] = (int32_t)run;
    run += x[j];
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) out[n] = (int32_t)sums[nb];
}

__global__ void mask_flags_kernel(const uint8_t* mask, int64_t n, int32_t* flags) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) flags[i] = mask[i] != 0 ? 1 : 0;
}

// The neighbours of voxel (x, y, z) in the reference's order of labels:
// offsets in lexicographic (dx, dy, dz) order visit increasing linear
// indices, and labels increase with the linear index (np.where order), so
// the list comes out sorted as sorted(set) gives it.  Label 0 (the first
// in-mask voxel, and every out-of-mask voxel in the reference's label
// volume) is never a neighbour and row 0 stays empty; with 6-connectivity a
// voxel whose 3x3x3 window is clipped by the volume edge has no neighbours.
template <bool FILL>
__device__ __forceinline__ int voxel_row(const uint8_t* mask, const int32_t* label, int X, int Y, int Z, int dirtype,
                                         int x, int y, int z, int32_t* out) {
  if (dirtype == 6 && (x == 0 || y == 0 || z == 0 || x == X - 1 || y == Y - 1 || z == Z - 1)) return 0;
  int k = 0;
  for (int dx = -1; dx <= 1; ++dx)
    for (int dy = -1; dy <= 1; ++dy)
      for (int dz = -1; dz <= 1; ++dz) {
        const int nz = (dx != 0) + (dy != 0) + (dz != 0);
        if (nz == 0 || (dirtype == 6 && nz != 1)) continue;
        const int xx = x + dx, yy = y + dy, zz = z + dz;
        if (xx < 0 || yy < 0 || zz < 0 || xx >= X || yy >= Y || zz >= Z) continue;
        const int64_t j = ((int64_t)xx * Y + yy) * Z + zz;
        if (!mask[j] || label[j] == 0) continue;
        if (FILL) out[k] = label[j];
        ++k;
      }
  return k;
}

template <bool FILL>
__global__ void voxel_rows_kernel(const uint8_t* mask, const int32_t* label, int X, int Y, int Z, int dirtype,
                                  int32_t* cnt, const int32_t* row_ptr, int32_t* col) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)X * Y * Z || !mask[i]) return;
  const int32_t lab = label[i];
  const int z = (int)(i % Z), y = (int)((i / Z) % Y), x = (int)(i / ((int64_t)Y * Z));
  if (!FILL) {
    cnt[lab] = lab == 0 ? 0 : voxel_row<false>(mask, label, X, Y, Z, dirtype, x, y, z, nullptr);
  } else if (lab != 0) {
    voxel_row<true>(mask, label, X, Y, Z, dirtype, x, y, z, col + row_ptr[lab]);
  }
}

// every face adds its other two vertices to each of its three vertices' lists
__global__ void face_count_kernel(const int32_t* faces, int64_t F, int32_t nv, int32_t* cnt, int32_t* bad) {
  const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= F) return;
  const int a = faces[3 * f], b = faces[3 * f + 1], c = faces[3 * f + 2];
  if (a < 0 || b < 0 || c < 0 || a >= nv || b >= nv || c >= nv) {
    atomicOr(bad, 1);
    return;
  }
  atomicAdd(&cnt[a], 2);
  atomicAdd(&cnt[b], 2);
  atomicAdd(&cnt[c], 2);
}

__global__ void face_scatter_kernel(const int32_t* faces, int64_t F, int32_t nv, const int32_t* off, int32_t* cur,
                                    int32_t* tmp) {
  const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= F) return;
  const int a = faces[3 * f], b = faces[3 * f + 1], c = faces[3 * f + 2];
  if (a < 0 || b < 0 || c < 0 || a >= nv || b >= nv || c >= nv) return;
  int s = atomicAdd(&cur[a], 2);
  tmp[off[a] + s] = b;
  tmp[off[a] + s + 1] = c;
  s = atomicAdd(&cur[b], 2);
  tmp[off[b] + s] = a;
  tmp[off[b] + s + 1] = c;
  s = atomicAdd(&cur[c], 2);
  tmp[off[c] + s] = a;
  tmp[off[c] + s + 1] = b;
}

// sort a vertex's entries and keep each value once (a set, then sorted, as
// the reference's set per vertex is read in ascending order)
__global__ void vertex_unique_kernel(const int32_t* off, int32_t nv, int32_t* tmp, int32_t* ulen) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nv) return;
  int32_t* seg = tmp + off[v];
  const int n = off[v + 1] - off[v];
  for (int i = 1; i < n; ++i) {
    const int32_t key = seg[i];
    int j = i - 1;
    while (j >= 0 && seg[j] > key) {
      seg[j + 1] = seg[j];
      --j;
    }
    seg[j + 1] = key;
  }
  int u = 0;
  for (int i = 0; i < n; ++i)
    if (u == 0 || seg[i] != seg[u - 1]) seg[u++] = seg[i];
  ulen[v] = u;
}

__global__ void vertex_copy_kernel(const int32_t* off, const int32_t* tmp, const int32_t* row_ptr, int32_t nv,
                                   int32_t* col) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nv) return;
  const int n = row_ptr[v + 1] - row_ptr[v];
  for (int i = 0; i < n; ++i) col[row_ptr[v] + i] = tmp[off[v] + i];
}

inline unsigned blocks_of(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

}  // namespace

int64_t scan_scratch_words(int64_t n) { return (n + kScanChunk - 1) / kScanChunk + 1; }

hipError_t launch_exclusive_scan(const int32_t* in, int64_t n, int32_t* out, int64_t* sums, hipStream_t s) {
  const int64_t nb = (n + kScanChunk - 1) / kScanChunk;
  if (nb > 0) hipLaunchKernelGGL(chunk_sum_kernel, dim3((unsigned)nb), dim3(kScanThreads), 0, s, in, n, sums);
  hipLaunchKernelGGL(sums_scan_kernel, dim3(1), dim3(kScanThreads), 0, s, sums, nb);
  hipLaunchKernelGGL(chunk_scan_kernel, dim3((unsigned)(nb > 0 ? nb : 1)), dim3(kScanThreads), 0, s, in, n, sums, nb,
                     out);
  return hipGetLastError();
}

hipError_t launch_mask_flags(const uint8_t* mask, int64_t n, int32_t* flags, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(mask_flags_kernel, dim3(blocks_of(n, 256)), dim3(256), 0, s, mask, n, flags);
  return hipGetLastError();
}

hipError_t launch_voxel_rows(const uint8_t* mask, const int32_t* label, int X, int Y, int Z, int dirtype,
                             int32_t* cnt, const int32_t* row_ptr, int32_t* col, hipStream_t s) {
  const int64_t n = (int64_t)X * Y * Z;
  if (n == 0) return hipSuccess;
  if (col)
    hipLaunchKernelGGL(voxel_rows_kernel<true>, dim3(blocks_of(n, 256)), dim3(256), 0, s, mask, label, X, Y, Z,
                       dirtype, cnt, row_ptr, col);
  else
    hipLaunchKernelGGL(voxel_rows_kernel<false>, dim3(blocks_of(n, 256)), dim3(256), 0, s, mask, label, X, Y, Z,
                       dirtype, cnt, row_ptr, col);
  return hipGetLastError();
}

hipError_t launch_face_count(const int32_t* faces, int64_t F, int32_t nv, int32_t* cnt, int32_t* bad, hipStream_t s) {
  if (F > 0) hipLaunchKernelGGL(face_count_kernel, dim3(blocks_of(F, 256)), dim3(256), 0, s, faces, F, nv, cnt, bad);
  return hipGetLastError();
}

hipError_t launch_face_scatter(const int32_t* faces, int64_t F, int32_t nv, const int32_t* off, int32_t* cur,
                               int32_t* tmp, hipStream_t s) {
  if (F > 0)
    hipLaunchKernelGGL(face_scatter_kernel, dim3(blocks_of(F, 256)), dim3(256), 0, s, faces, F, nv, off, cur, tmp);
  return hipGetLastError();
}

hipError_t launch_vertex_unique(const int32_t* off, int32_t nv, int32_t* tmp, int32_t* ulen, hipStream_t s) {
  if (nv > 0) hipLaunchKernelGGL(vertex_unique_kernel, dim3(blocks_of(nv, 256)), dim3(256), 0, s, off, nv, tmp, ulen);
  return hipGetLastError();
}

hipError_t launch_vertex_copy(const int32_t* off, const int32_t* tmp, const int32_t* row_ptr, int32_t nv, int32_t* col,
                              hipStream_t s) {
  if (nv > 0)
    hipLaunchKernelGGL(vertex_copy_kernel, dim3(blocks_of(nv, 256)), dim3(256), 0, s, off, tmp, row_ptr, nv, col);
  return hipGetLastError();
}

}  // namespace tfcehip
