// C-ABI host layer of the GPU adjacency builders.
#include <hip/hip_runtime.h>

#include <climits>
#include <string>

#include "adj_internal.h"
#include "capi_common.h"
#include "tfcehip.h"

using namespace tfcehip;

namespace {

// device buffers of one builder call, freed on every return path once the
// stream has drained (an error return may leave launches in flight)
struct Buffers {
  explicit Buffers(hipStream_t s) : stream(s) {}
  hipStream_t stream;
  void* p[8] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  int n = 0;
  ~Buffers() {
    if (n > 0) (void)hipStreamSynchronize(stream);
    for (int i = 0; i < n; ++i)
      if (p[i]) (void)hipFree(p[i]);
  }
  template <typename T>
  int get(T** out, int64_t count, const char* what) {
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, (size_t)(count > 0 ? count : 1) * sizeof(T));
    if (e != hipSuccess) return set_error(-(int)e, std::string("hipMalloc(") + what + "): " + hipGetErrorString(e));
    p[n++] = q;
    *out = reinterpret_cast<T*>(q);
    return 0;
  }
};

int check(hipError_t e, const char* what) {
  return e == hipSuccess ? 0 : set_error(-(int)e, std::string(what) + ": " + hipGetErrorString(e));
}

// exclusive scan of in[0, n) into out[0, n], total returned (blocking)
int scan(const int32_t* in, int64_t n, int32_t* out, int64_t* sums, int64_t* total, hipStream_t s) {
  int rc = check(launch_exclusive_scan(in, n, out, sums, s), "scan launch");
  if (rc) return rc;
  const int64_t nb = scan_scratch_words(n) - 1;
  rc = check(hipMemcpyAsync(total, sums + nb, sizeof(int64_t), hipMemcpyDeviceToHost, s), "scan total");
  if (!rc) rc = check(hipStreamSynchronize(s), "scan sync");
  return rc;
}

}  // namespace

extern "C" {

int tfce_adj_voxel_csr(const uint8_t* d_mask, int32_t X, int32_t Y, int32_t Z, int32_t dirtype, int32_t* d_row_ptr,
                       int32_t* d_col, int32_t* V, int64_t* nnz, hipStream_t stream) {
  if (!d_mask || !V || !nnz) return set_error(TFCE_EINVAL, "NULL argument");
  if (X < 1 || Y < 1 || Z < 1) return set_error(TFCE_EINVAL, "volume dimensions must be >= 1");
  if (dirtype != 6 && dirtype != 26) return set_error(TFCE_EINVAL, "dirtype must be 6 or 26");
  if ((d_row_ptr == nullptr) != (d_col == nullptr))
    return set_error(TFCE_EINVAL, "give both d_row_ptr and d_col (fill) or neither (sizes only)");
  const int64_t n = (int64_t)X * Y * Z;
  if (n >= INT32_MAX) return set_error(TFCE_EINVAL, "the volume holds 2^31 voxels or more");
  Buffers b(stream);
  int32_t *flags = nullptr, *label = nullptr, *rp = nullptr;
  int64_t* sums = nullptr;
  int rc = 0;
  if ((rc = b.get(&flags, n, "flags")) || (rc = b.get(&label, n + 1, "labels")) ||
      (rc = b.get(&sums, scan_scratch_words(n), "scan sums")))
    return rc;
  if ((rc = check(launch_mask_flags(d_mask, n, flags, stream), "mask flags"))) return rc;
  int64_t nv = 0;
  if ((rc = scan(flags, n, label, sums, &nv, stream))) return rc;
  // labels: the in-mask voxels numbered in C order (np.where order)
  int32_t* cnt = flags;  // the flags are no longer needed; V <= n
  if ((rc = check(launch_voxel_rows(d_mask, label, X, Y, Z, dirtype, cnt, nullptr, nullptr, stream), "row counts")))
    return rc;
  rp = d_row_ptr;
  if (!rp && (rc = b.get(&rp, nv + 1, "row_ptr"))) return rc;
  int64_t total = 0;
  if ((rc = scan(cnt, nv, rp, sums, &total, stream))) return rc;
  if (total >= INT32_MAX) return set_error(TFCE_EINVAL, "the adjacency holds 2^31 entries or more");
  if (d_col) {
    if ((rc = check(launch_voxel_rows(d_mask, label, X, Y, Z, dirtype, cnt, rp, d_col, stream), "row fill"))) return rc;
    if ((rc = check(hipStreamSynchronize(stream), "row fill sync"))) return rc;
  }
  *V = (int32_t)nv;
  *nnz = total;
  return 0;
}

int tfce_adj_mesh_csr(const int32_t* d_faces, int64_t F, int32_t n_vertices, int32_t* d_row_ptr, int32_t* d_col,
                      int64_t* nnz, hipStream_t stream) {
  if (!nnz || (F > 0 && !d_faces)) return set_error(TFCE_EINVAL, "NULL argument");
  if (F < 0 || n_vertices < 0) return set_error(TFCE_EINVAL, "negative size");
  if ((d_row_ptr == nullptr) != (d_col == nullptr))
    return set_error(TFCE_EINVAL, "give both d_row_ptr and d_col (fill) or neither (sizes only)");
  if (6 * F >= INT32_MAX) return set_error(TFCE_EINVAL, "too many faces");
  const int32_t nv = n_vertices;
  Buffers b(stream);
  int32_t *cnt = nullptr, *off = nullptr, *tmp = nullptr, *bad = nullptr, *rp = nullptr;
  int64_t* sums = nullptr;
  const int64_t sw = scan_scratch_words(nv);
  int rc = 0;
  if ((rc = b.get(&cnt, nv, "counts")) || (rc = b.get(&off, (int64_t)nv + 1, "offsets")) ||
      (rc = b.get(&tmp, 6 * F, "entries")) || (rc = b.get(&bad, 1, "flag")) || (rc = b.get(&sums, sw, "scan sums")))
    return rc;
  if ((rc = check(hipMemsetAsync(cnt, 0, (size_t)(nv > 0 ? nv : 1) * sizeof(int32_t), stream), "clear counts")) ||
      (rc = check(hipMemsetAsync(bad, 0, sizeof(int32_t), stream), "clear flag")))
    return rc;
  if ((rc = check(launch_face_count(d_faces, F, nv, cnt, bad, stream), "face counts"))) return rc;
  int32_t hbad = 0;
  if ((rc = check(hipMemcpyAsync(&hbad, bad, sizeof(int32_t), hipMemcpyDeviceToHost, stream), "flag")) ||
      (rc = check(hipStreamSynchronize(stream), "face count sync")))
    return rc;
  if (hbad) return set_error(TFCE_EADJ, "a face names a vertex outside [0, n_vertices)");
  int64_t total = 0;
  if ((rc = scan(cnt, nv, off, sums, &total, stream))) return rc;
  if ((rc = check(hipMemsetAsync(cnt, 0, (size_t)(nv > 0 ? nv : 1) * sizeof(int32_t), stream), "clear cursors")))
    return rc;
  if ((rc = check(launch_face_scatter(d_faces, F, nv, off, cnt, tmp, stream), "face scatter"))) return rc;
  // cnt now holds each vertex's entry count; reuse it for the unique lengths
  if ((rc = check(launch_vertex_unique(off, nv, tmp, cnt, stream), "vertex unique"))) return rc;
  rp = d_row_ptr;
  if (!rp && (rc = b.get(&rp, (int64_t)nv + 1, "row_ptr"))) return rc;
  if ((rc = scan(cnt, nv, rp, sums, &total, stream))) return rc;
  if (d_col) {
    if ((rc = check(launch_vertex_copy(off, tmp, rp, nv, d_col, stream), "vertex copy")) ||
        (rc = check(hipStreamSynchronize(stream), "vertex copy sync")))
      return rc;
  }
  *nnz = total;
  return 0;
}

}  // extern "C"
